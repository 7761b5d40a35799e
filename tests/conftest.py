import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libmepgl.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


def small_string(N=24, F=5, multicomponent=False, masked=False, no_self_field=False, seed=0,
                 H=0.7, dx=0.3, noise=0.05, planes=False):
    """A small string with a vortex path, a non-uniform gauge and noise, so
    that no term of the energy or gradient vanishes by symmetry."""
    from mepgl_amd import configs

    if multicomponent:
        dom = None
        if masked:
            x, y = configs.grid(N, dx)
            dom = (np.hypot(x, y) < 0.45 * N * dx).astype(np.int32)
            dom[N // 2 - 1:N // 2 + 1, -N // 4:] = 0
        s = configs.two_band_entry(N, dx, F, H, seed, noise=noise, domain=dom,
                                   no_self_field=no_self_field, x_start=-N * dx / 3)
    else:
        s = configs.vortex_entry(N, dx, F, H, seed, noise=noise)
        if masked:
            x, y = configs.grid(N, dx)
            dom = (np.hypot(x, y) < 0.45 * N * dx).astype(np.int32)
            s.comp_domain = dom
            s.sc_domain = dom.copy()
        s.no_self_field = no_self_field
    rng = np.random.default_rng(seed + 100)
    s.a = s.a + 0.05 * rng.standard_normal(s.a.shape)
    if planes:
        s.a_1 = -1.0 + 0.1 * rng.standard_normal((N, N))
        s.m_xx_1 = 1.0 + 0.2 * rng.random((N, N))
        s.h = H + 0.05 * rng.standard_normal((N, N))
    return s.validate()


def bondless_string(no_self_field=False):
    """N = 130 with a centred disk domain: the 2-column third strip and the
    top and bottom segments of 7-row tiles have no bond at all."""
    from mepgl_amd import configs

    N = 130
    s = small_string(N=N, F=4, seed=50, multicomponent=True, no_self_field=no_self_field)
    x, y = configs.grid(N, s.dx)
    dom = (np.hypot(x, y) < 0.3 * N * s.dx).astype(np.int32)
    s.comp_domain = dom
    s.sc_domain = dom.copy()
    return s.validate()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def normwise(a, b, axis_image=0):
    """Per-image norm-wise relative error max_n |a_n - b_n| / |b_n|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    F = a.shape[axis_image]
    out = 0.0
    for n in range(F):
        num = np.linalg.norm((a[n] - b[n]).ravel())
        den = np.linalg.norm(b[n].ravel())
        out = max(out, num / den if den > 0 else num)
    return out
