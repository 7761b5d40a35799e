"""Description of one string of images for the gauged string method.

Layout conventions (the reference's device layout, io.cu:23-30, 91-125):
per-site planes are (N, N) arrays indexed [j, i] with i along x, and fields
are (F, N, N, 2) arrays indexed [n, j, i, component] (component 0 = real part
or x, 1 = imaginary part or y).  The on-disk .npy files of the reference hold
the transpose of each (N, N) plane ([i][j]); mepgl_amd.npyio converts.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Union

import numpy as np

Plane = Union[float, np.ndarray]

PARAM_NAMES = ("a_1", "b_1", "m_xx_1", "m_yy_1", "a_2", "b_2", "m_xx_2", "m_yy_2", "h")
COUPLING_NAMES = ("q_1", "q_2", "eta", "gamma", "delta")


@dataclass
class GLString:
    """One string: grid, Ginzburg-Landau parameters and the F images.

    Parameters a_1 ... h are either floats (uniform over the grid) or (N, N)
    planes; q_1 ... delta are the scalar couplings of q.npy (io.cu:60-67).
    """

    N: int
    dx: float
    F: int
    multicomponent: bool
    comp_domain: np.ndarray
    sc_domain: np.ndarray
    a: np.ndarray
    psi_1: np.ndarray
    psi_2: Optional[np.ndarray] = None
    no_self_field: bool = False
    a_1: Plane = -1.0
    b_1: Plane = 1.0
    m_xx_1: Plane = 1.0
    m_yy_1: Plane = 1.0
    a_2: Plane = -1.0
    b_2: Plane = 1.0
    m_xx_2: Plane = 1.0
    m_yy_2: Plane = 1.0
    h: Plane = 0.0
    q_1: float = 1.0
    q_2: float = 1.0
    eta: float = 0.0
    gamma: float = 0.0
    delta: float = 0.0
    name: str = ""
    meta: dict = field(default_factory=dict)

    def validate(self) -> "GLString":
        N, F = int(self.N), int(self.F)
        if N < 2 or F < 1:
            raise ValueError(f"need N >= 2 and F >= 1 (got N={N}, F={F})")
        for nm in ("comp_domain", "sc_domain"):
            d = np.asarray(getattr(self, nm))
            if d.shape != (N, N):
                raise ValueError(f"{nm} must be ({N}, {N}), got {d.shape}")
        for nm in ("a", "psi_1") + (("psi_2",) if self.multicomponent else ()):
            f = getattr(self, nm)
            if f is None or np.asarray(f).shape != (F, N, N, 2):
                raise ValueError(f"{nm} must be ({F}, {N}, {N}, 2)")
        for nm in PARAM_NAMES:
            v = getattr(self, nm)
            if not np.isscalar(v) and np.asarray(v).shape != (N, N):
                raise ValueError(f"parameter {nm} must be a scalar or ({N}, {N})")
        return self

    @property
    def ncomp(self) -> int:
        return 2 if self.multicomponent else 1

    @property
    def image_cells(self) -> int:
        return int(self.F) * int(self.N) ** 2

    def copy(self) -> "GLString":
        import copy

        return copy.deepcopy(self)
