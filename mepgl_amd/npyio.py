"""The on-disk layout of a simulation, as the mepgl program reads and writes it.

``simulations/<name>/input_data/`` (io.cu:10-126) and ``.../output_data/``
(io.cu:128-291) hold

* ``x.npy``, ``y.npy`` — the mesh, (N, N) ``[i][j]`` with i along x;
* ``comp_domain.npy``, ``sc_domain.npy`` — int32 (N, N) masks;
* ``a_1 b_1 m_xx_1 m_yy_1 a_2 b_2 m_xx_2 m_yy_2 h .npy`` — (N, N) planes;
* ``q.npy`` — ``[q_1, q_2, eta, gamma, delta]``;
* ``<n>/ax ay u_1 v_1 [u_2 v_2].npy`` per image, and on output also
  ``<n>/jx_1 jy_1 b [jx_2 jy_2].npy`` and ``fenergy s s_jb .npy``;
* ``dx.npy`` — the lattice spacing.  The original keeps dx as a compile-time
  constant; here it travels with the data (the program falls back to
  ``x[1][0] - x[0][0]`` when the file is absent).

Every (N, N) file array is the transpose of the in-memory plane of
:class:`~mepgl_amd.problem.GLString` (indexed ``[j, i]``).
"""
from __future__ import annotations

import os
import shutil
from typing import Optional

import numpy as np

from .problem import COUPLING_NAMES, PARAM_NAMES, GLString


def sim_dir(simname: str, root: str = ".") -> str:
    return os.path.join(root, "simulations", simname)


def _plane(v, N):
    return np.broadcast_to(np.asarray(v, dtype=np.float64), (N, N))


def write_input(s: GLString, simname: str, root: str = ".", double_precision: bool = True,
                x: Optional[np.ndarray] = None, y: Optional[np.ndarray] = None) -> str:
    """Write ``s`` as ``simulations/<simname>/input_data`` (launcher_utils.py:194-291).

    Floats are stored as float64 when ``double_precision`` else float32, as
    the original launcher does; the masks as int32.
    """
    s.validate()
    N, F = int(s.N), int(s.F)
    dt = np.float64 if double_precision else np.float32
    d = os.path.join(sim_dir(simname, root), "input_data")
    if os.path.exists(d):
        shutil.rmtree(d)
    os.makedirs(d)
    if x is None or y is None:
        ax = (np.arange(N) - (N - 1) / 2.0) * s.dx
        x, y = np.meshgrid(ax, ax, indexing="ij")
    np.save(os.path.join(d, "x.npy"), np.asarray(x).astype(dt))
    np.save(os.path.join(d, "y.npy"), np.asarray(y).astype(dt))
    np.save(os.path.join(d, "dx.npy"), np.array([s.dx], dtype=np.float64))
    np.save(os.path.join(d, "comp_domain.npy"), np.asarray(s.comp_domain).T.astype(np.int32))
    np.save(os.path.join(d, "sc_domain.npy"), np.asarray(s.sc_domain).T.astype(np.int32))
    for nm in PARAM_NAMES:
        np.save(os.path.join(d, f"{nm}.npy"), _plane(getattr(s, nm), N).T.astype(dt))
    np.save(os.path.join(d, "q.npy"), np.array([getattr(s, nm) for nm in COUPLING_NAMES]).astype(dt))
    for n in range(F):
        fd = os.path.join(d, str(n))
        os.makedirs(fd)
        np.save(os.path.join(fd, "ax.npy"), s.a[n, :, :, 0].T.astype(dt))
        np.save(os.path.join(fd, "ay.npy"), s.a[n, :, :, 1].T.astype(dt))
        np.save(os.path.join(fd, "u_1.npy"), s.psi_1[n, :, :, 0].T.astype(dt))
        np.save(os.path.join(fd, "v_1.npy"), s.psi_1[n, :, :, 1].T.astype(dt))
        if s.multicomponent:
            np.save(os.path.join(fd, "u_2.npy"), s.psi_2[n, :, :, 0].T.astype(dt))
            np.save(os.path.join(fd, "v_2.npy"), s.psi_2[n, :, :, 1].T.astype(dt))
    return d


def _load(path):
    return np.load(path, allow_pickle=False)


def read_dir(path: str, F: Optional[int] = None, no_self_field: bool = False,
             dx: Optional[float] = None) -> tuple[GLString, dict]:
    """Read an input_data or output_data directory.

    Returns the string (planes transposed back to ``[j, i]``) and a dict of
    whatever output-only arrays exist: ``fenergy``, ``s``, ``s_jb`` (F,),
    ``b`` (F, N, N), ``j_1``, ``j_2`` (F, N, N, 2).
    """
    x = _load(os.path.join(path, "x.npy"))
    N = x.shape[0]
    if F is None:
        F = 0
        while os.path.isdir(os.path.join(path, str(F))):
            F += 1
    if dx is None:
        p = os.path.join(path, "dx.npy")
        dx = float(_load(p)[0]) if os.path.exists(p) else float(x[1, 0]) - float(x[0, 0])
    mc = os.path.exists(os.path.join(path, "0", "u_2.npy"))
    a = np.zeros((F, N, N, 2))
    p1 = np.zeros((F, N, N, 2))
    p2 = np.zeros((F, N, N, 2)) if mc else None
    extra: dict = {}
    has_j = os.path.exists(os.path.join(path, "0", "b.npy"))
    if has_j:
        extra["b"] = np.zeros((F, N, N))
        extra["j_1"] = np.zeros((F, N, N, 2))
        if mc:
            extra["j_2"] = np.zeros((F, N, N, 2))

    def rd(n, nm):
        return _load(os.path.join(path, str(n), nm + ".npy")).astype(np.float64).T

    for n in range(F):
        a[n, :, :, 0], a[n, :, :, 1] = rd(n, "ax"), rd(n, "ay")
        p1[n, :, :, 0], p1[n, :, :, 1] = rd(n, "u_1"), rd(n, "v_1")
        if mc:
            p2[n, :, :, 0], p2[n, :, :, 1] = rd(n, "u_2"), rd(n, "v_2")
        if has_j:
            extra["b"][n] = rd(n, "b")
            extra["j_1"][n, :, :, 0], extra["j_1"][n, :, :, 1] = rd(n, "jx_1"), rd(n, "jy_1")
            if mc:
                extra["j_2"][n, :, :, 0], extra["j_2"][n, :, :, 1] = rd(n, "jx_2"), rd(n, "jy_2")
    for nm in ("fenergy", "s", "s_jb"):
        p = os.path.join(path, nm + ".npy")
        if os.path.exists(p):
            extra[nm] = _load(p).astype(np.float64)
    q = _load(os.path.join(path, "q.npy")).astype(np.float64)
    kw = {nm: _load(os.path.join(path, f"{nm}.npy")).astype(np.float64).T for nm in PARAM_NAMES}
    kw.update({nm: float(q[k]) for k, nm in enumerate(COUPLING_NAMES)})
    s = GLString(N=N, dx=dx, F=F, multicomponent=mc,
                 comp_domain=_load(os.path.join(path, "comp_domain.npy")).astype(np.int32).T.copy(),
                 sc_domain=_load(os.path.join(path, "sc_domain.npy")).astype(np.int32).T.copy(),
                 a=a, psi_1=p1, psi_2=p2, no_self_field=no_self_field, **kw)
    return s, extra


def reload(simname: str, root: str = ".") -> None:
    """Make a run's output the next run's input (reload.py:1-15)."""
    d = sim_dir(simname, root)
    shutil.rmtree(os.path.join(d, "input_data"))
    shutil.copytree(os.path.join(d, "output_data"), os.path.join(d, "input_data"))
