"""Initial guesses for the benchmark workloads (vortex-entry strings).

This is the project's own generator.  It follows the semantics of the
reference's StringBuilder (mepgl/mepgl_lib/builder.py:32-34 grid with
inclusive end points and ij indexing; builder.py:214-216 uniform-field gauge
A = (-H y / 2, H x / 2)), with the order-parameter profile
psi = sqrt(-a/b) tanh(r) exp(i w theta) around a vortex whose position moves
linearly from outside the sample to its centre across the images.

Workloads (seeded, deterministic):
  vortex_entry_small   64x64,   dx=0.25,  F=15,  one component,  H=0.9
  vortex_entry_256     256x256, dx=0.125, F=64,  one component,  H=1.0
  two_band_512         512x512, dx=0.125, F=101, two anisotropic components
  sweep_48             48 strings of uneven N, F, H and component count
  thin_film_disk_1024  1024x1024 thin-film disk with a notch, 4 replicas
"""
from __future__ import annotations

import numpy as np

from .problem import GLString


def grid(N: int, dx: float):
    """Coordinates x[j, i], y[j, i] of the N x N grid, centred on the origin
    with inclusive end points (builder.py:32-34)."""
    half = (N - 1) * dx / 2.0
    ax = np.linspace(-half, half, N)
    y, x = np.meshgrid(ax, ax, indexing="ij")
    return x, y


def vortex_images(N, dx, F, amp, x_start, winding=1, y_v=0.0, x_end=0.0):
    """psi for F images with a vortex core moving from x_start to x_end."""
    x, y = grid(N, dx)
    xv = np.linspace(x_start, x_end, F)
    psi = np.empty((F, N, N), dtype=np.complex128)
    for n in range(F):
        r = np.hypot(x - xv[n], y - y_v)
        th = np.arctan2(y - y_v, x - xv[n])
        psi[n] = amp * np.tanh(r) * np.exp(1j * winding * th)
    return psi


def uniform_gauge(N, dx, F, H):
    x, y = grid(N, dx)
    a = np.empty((F, N, N, 2))
    a[..., 0] = -H * y / 2.0
    a[..., 1] = H * x / 2.0
    return a


def _pack(psi):
    out = np.empty(psi.shape + (2,))
    out[..., 0] = psi.real
    out[..., 1] = psi.imag
    return out


def vortex_entry(N, dx, F, H, seed, noise=1e-3, name="vortex_entry"):
    """Single-component type-II string (a=-1, b=1, m=1, q=1, kappa=2) for the
    entry of one vortex through the vacuum boundary of a full square box."""
    L = N * dx
    amp = 1.0
    psi = vortex_images(N, dx, F, amp, x_start=-(L / 2 + 2.0))
    rng = np.random.default_rng(seed)
    p = _pack(psi)
    p[..., 0] += rng.normal(0.0, noise, (F, N, N))
    p[..., 1] += rng.normal(0.0, noise, (F, N, N))
    box = np.ones((N, N), dtype=np.int32)
    return GLString(N=N, dx=dx, F=F, multicomponent=False, comp_domain=box, sc_domain=box.copy(),
                    a=uniform_gauge(N, dx, F, H), psi_1=p, psi_2=None,
                    a_1=-1.0, b_1=1.0, m_xx_1=1.0, m_yy_1=1.0, h=float(H), q_1=1.0,
                    name=name, meta={"seed": seed, "H": H}).validate()


TWO_BAND = dict(a_1=-1.0, a_2=-0.6, b_1=1.0, b_2=0.8, m_xx_1=1.0, m_yy_1=1.0, m_xx_2=0.5,
                m_yy_2=2.0, q_1=1.0, q_2=0.7, eta=-0.2, gamma=0.3, delta=0.05)


def two_band_entry(N, dx, F, H, seed, noise=1e-3, params=None, domain=None, theta12=0.0,
                   no_self_field=False, x_start=None, name="two_band"):
    """Two-component string: a component-1 vortex enters while component 2
    stays uniform with relative phase theta12 (fractional-vortex entry)."""
    p = dict(TWO_BAND if params is None else params)
    L = N * dx
    amp1 = np.sqrt(-p["a_1"] / p["b_1"])
    amp2 = np.sqrt(-p["a_2"] / p["b_2"])
    xs = -(L / 2 + 2.0) if x_start is None else x_start
    psi1 = vortex_images(N, dx, F, amp1, x_start=xs)
    psi2 = np.full((F, N, N), amp2 * np.exp(1j * theta12), dtype=np.complex128)
    rng = np.random.default_rng(seed)
    p1, p2 = _pack(psi1), _pack(psi2)
    for arr in (p1, p2):
        arr[..., 0] += rng.normal(0.0, noise, (F, N, N))
        arr[..., 1] += rng.normal(0.0, noise, (F, N, N))
    dom = np.ones((N, N), dtype=np.int32) if domain is None else np.asarray(domain, np.int32)
    if domain is not None:
        p1[:, dom == 0] = 0.0
        p2[:, dom == 0] = 0.0
    return GLString(N=N, dx=dx, F=F, multicomponent=True, comp_domain=dom, sc_domain=dom.copy(),
                    a=uniform_gauge(N, dx, F, H), psi_1=p1, psi_2=p2,
                    no_self_field=no_self_field, h=float(H), name=name,
                    meta={"seed": seed, "H": H}, **p).validate()


def vortex_entry_small(seed=1001):
    return vortex_entry(64, 0.25, 15, 0.9, seed, name="vortex_entry_small")


def vortex_entry_256(seed=1002):
    return vortex_entry(256, 0.125, 64, 1.0, seed, name="vortex_entry_256")


def two_band_512(seed=1003):
    return two_band_entry(512, 0.125, 101, 0.8, seed, name="two_band_512")


def disk_with_notch(N, dx, radius=120.0, notch_w=10.0, notch_h=20.0):
    """Boolean cell mask: a disk with a rectangular notch cut in at +x."""
    x, y = grid(N, dx)
    disk = np.hypot(x, y) <= radius
    notch = (x >= radius - notch_h) & (np.abs(y) <= notch_w / 2.0)
    return (disk & ~notch).astype(np.int32)


THIN_FILM = dict(a_1=-1.0, a_2=-1.0, b_1=1.0, b_2=1.0, m_xx_1=1.0, m_yy_1=1.0, m_xx_2=1.0,
                 m_yy_2=1.0, q_1=1.0, q_2=1.0, eta=0.0, gamma=0.5, delta=0.1)


def thin_film_disk(replica, N=1024, dx=0.25, F=33, H=0.7):
    mask = disk_with_notch(N, dx)
    return two_band_entry(N, dx, F, H, seed=1005 * 1000 + replica, params=THIN_FILM, domain=mask,
                          theta12=np.pi / 2, no_self_field=True, x_start=-(120.0 + 2.0),
                          name=f"thin_film_disk_{replica}")


def sweep_48(seed=1004, count=48):
    """Uneven batch: N from {64,96,128,192,256,384}, F in 15..65, H in
    [0.6, 1.2], one or two components, dx = 16/N."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        N = int(rng.choice([64, 96, 128, 192, 256, 384]))
        F = int(rng.integers(15, 66))
        H = float(rng.uniform(0.6, 1.2))
        nc = int(rng.integers(1, 3))
        dx = 16.0 / N
        s = seed * 1000 + i
        if nc == 1:
            out.append(vortex_entry(N, dx, F, H, s, name=f"sweep_{i}"))
        else:
            out.append(two_band_entry(N, dx, F, H, s, name=f"sweep_{i}"))
    return out


WORKLOADS = {
    "vortex_entry_small": lambda: [vortex_entry_small()],
    "vortex_entry_256": lambda: [vortex_entry_256()],
    "two_band_512": lambda: [two_band_512()],
    "sweep_48": sweep_48,
    "thin_film_disk_1024": lambda: [thin_film_disk(r) for r in range(4)],
}
