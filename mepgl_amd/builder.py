"""Initial guesses of a Ginzburg-Landau string with the semantics of the
reference's StringBuilder (mepgl/mepgl_lib/builder.py:4-224), so that a
user config.py written for the reference builds the same arrays here.

Conventions kept from the reference:

* the grid is ``linspace`` over the closed intervals x_lim and y_lim with
  ``meshgrid(..., indexing="ij")``: every (Nx, Ny) array is indexed [i][j]
  with i along x, the layout of the on-disk .npy files (builder.py:32-34);
* domains are given by shape functions f(x, y) -> bool, evaluated on the
  grid; the default is the whole box (builder.py:37-45);
* a vortex of winding w at (xv, yv) multiplies the amplitude by
  tanh(-r^2) and adds w atan2(y - yv, x - xv) to the phase, per image, when
  w != 0 (builder.py:138-152; the negative amplitude of tanh(-r^2) is the
  reference's);
* component 2 starts from the phase difference phase_fun(x, y)
  (builder.py:160-176);
* noise is drawn from numpy's global random state with np.random.randn, one
  (Nx, Ny) draw per field in the order u_1, v_1, u_2, v_2 for each image
  (builder.py:154-184), so a seeded run reproduces the reference's arrays;
* matter fields are zero outside the superconducting domain, and the vector
  potential outside the computational domain (builder.py:154-187, 221-222);
* gauges "Landau_x" (a_x = -B y), "Landau_y" (a_y = B x) and "symmetric"
  (a = (-B y / 2, B x / 2)) (builder.py:207-219).
"""
from __future__ import annotations

from typing import Callable, List, Sequence

import numpy as np


def _mask(fun: Callable, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Boolean grid mask of a shape function, vectorising scalar ones."""
    try:
        m = np.asarray(fun(x, y))
    except (TypeError, ValueError):
        m = None
    if m is None or m.shape != x.shape:
        m = np.asarray(np.vectorize(fun)(x, y))
    return m.astype(bool)


class StringBuilder:
    """Builder of the initial string: grid, domains, vortex paths and gauge.

    F: images; Nx, Ny: grid points; x_lim, y_lim: closed coordinate
    intervals; multicomponent: two order-parameter components; B_z: the
    applied field used by generate_vector_potential.
    """

    def __init__(self, F, Nx, Ny, x_lim, y_lim, multicomponent, B_z=0):
        self.F = int(F)
        self.Nx = int(Nx)
        self.Ny = int(Ny)
        self.x_lim = x_lim
        self.y_lim = y_lim
        self.multicomponent = bool(multicomponent)
        self.B_z = B_z
        self.x_ax = np.linspace(x_lim[0], x_lim[1], self.Nx)
        self.y_ax = np.linspace(y_lim[0], y_lim[1], self.Ny)
        self.x, self.y = np.meshgrid(self.x_ax, self.y_ax, indexing="ij")
        box = self._in_box
        self.comp_domain_shape_fun = box
        self.sc_domain_shape_fun = box
        self.phase_fun = lambda x, y: 0
        self.vortices_1: List[list] = []
        self.vortices_2: List[list] = []

    def _in_box(self, x, y):
        return ((x >= self.x_lim[0]) & (x <= self.x_lim[1]) &
                (y >= self.y_lim[0]) & (y <= self.y_lim[1]))

    # -- definitions ----------------------------------------------------------
    def set_comp_domain(self, shape_fun: Callable):
        """Computational domain: shape_fun(x, y) true inside."""
        self.comp_domain_shape_fun = shape_fun

    def set_sc_domain(self, shape_fun: Callable):
        """Superconducting domain: shape_fun(x, y) true inside."""
        self.sc_domain_shape_fun = shape_fun

    def add_phase_diff(self, phase_fun: Callable):
        """Phase of component 2 relative to component 1, phase_fun(x, y)."""
        self.phase_fun = phase_fun

    def add_vortex(self, xv: Sequence, yv: Sequence, Nv: Sequence, component: int):
        """A vortex path: centre (xv[n], yv[n]) and winding Nv[n] in image n."""
        if component == 1:
            self.vortices_1.append([xv, yv, Nv])
        if component == 2:
            self.vortices_2.append([xv, yv, Nv])

    # -- arrays ----------------------------------------------------------------
    def _domain(self, fun) -> np.ndarray:
        d = np.zeros((self.Nx, self.Ny), dtype=np.intc)
        d[_mask(fun, self.x, self.y)] = 1
        return d

    def get_domain_matrices(self):
        """(comp_domain, sc_domain): int arrays (Nx, Ny), 1 inside."""
        return self._domain(self.comp_domain_shape_fun), self._domain(self.sc_domain_shape_fun)

    def _component(self, n, psi_abs, phase0, vortices):
        amp = psi_abs + 0 * self.x
        phase = phase0 + 0 * self.x
        for xv, yv, wv in vortices:
            if wv[n] != 0:
                dx, dy = self.x - xv[n], self.y - yv[n]
                r = np.sqrt(dx ** 2 + dy ** 2)
                amp = amp * np.tanh(-(r ** 2))
                phase = phase + wv[n] * np.arctan2(dy, dx)
        return amp, phase

    def generate_matter_fields(self, psi_abs_1=1.0, psi_abs_2=1.0, noise=0.0):
        """(u_1, v_1, u_2, v_2), each (F, Nx, Ny); u_2 and v_2 are ones for
        one component, as in the reference."""
        sc = self._domain(self.sc_domain_shape_fun)
        shape = (self.F, self.Nx, self.Ny)
        u_1, v_1 = np.ones(shape), np.ones(shape)
        u_2, v_2 = np.ones(shape), np.ones(shape)
        grid = (self.Nx, self.Ny)
        for n in range(self.F):
            amp, ph = self._component(n, psi_abs_1, 0.0, self.vortices_1)
            u_1[n] = (amp * np.cos(ph) + noise * np.random.randn(*grid)) * sc
            v_1[n] = (amp * np.sin(ph) + noise * np.random.randn(*grid)) * sc
            if self.multicomponent:
                amp, ph = self._component(n, psi_abs_2, self.phase_fun(self.x, self.y),
                                          self.vortices_2)
                u_2[n] = (amp * np.cos(ph) + noise * np.random.randn(*grid)) * sc
                v_2[n] = (amp * np.sin(ph) + noise * np.random.randn(*grid)) * sc
        return u_1, v_1, u_2, v_2

    def generate_vector_potential(self, gauge="symmetric"):
        """(a_x, a_y), each (F, Nx, Ny), for the field B_z in the given gauge,
        zero outside the computational domain."""
        comp = self._domain(self.comp_domain_shape_fun)
        B = self.B_z
        if gauge == "Landau_x":
            ax, ay = -B * self.y, np.zeros_like(self.x)
        elif gauge == "Landau_y":
            ax, ay = np.zeros_like(self.x), B * self.x
        elif gauge == "symmetric":
            ax, ay = -1 / 2 * B * self.y, 1 / 2 * B * self.x
        else:
            raise ValueError(f"gauge {gauge!r} is not one of Landau_x, Landau_y, symmetric")
        a_x = np.repeat((ax * comp)[None], self.F, axis=0)
        a_y = np.repeat((ay * comp)[None], self.F, axis=0)
        return a_x, a_y
