"""mepgl_free_energy_noalloc (computeFreeEnergy on caller-owned device
arrays, partials in the library's static scratch array) against the
context's mepgl_free_energy on the same string at the same segment height:
equal bit for bit, on two calls in a row, the second with a smaller input.
The test allocates the device inputs and outputs itself (the entry point
does not), through the HIP runtime the library is linked against."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import small_string

pytestmark = pytest.mark.gpu

ROWS = 16
H2D, D2H = 1, 2


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def hip_runtime():
    """The HIP runtime libmepgl.so is linked against (the one mapped into
    this process by loading it), so that test allocations and the library's
    launches share one runtime."""
    from mepgl_amd import _lib

    _lib.load()
    path = "/opt/rocm/lib/libamdhip64.so"
    with open("/proc/self/maps") as f:
        for line in f:
            p = line.split()[-1]
            if "libamdhip64" in p and "torch" not in p:
                path = p
                break
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class Device:
    """Device buffers of one call, freed on exit."""

    def __init__(self):
        self.h = hip_runtime()
        self.ptrs = []

    def put(self, arr, dtype=np.float64):
        a = np.ascontiguousarray(arr, dtype=dtype)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.ptrs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p.value

    def get(self, ptr, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.h.hipDeviceSynchronize()
        for p in self.ptrs:
            self.h.hipFree(p)


def noalloc_energy(s, rows=ROWS):
    from mepgl_amd import _lib

    L = _lib.load()
    fn = L.mepgl_free_energy_noalloc
    fn.argtypes = [C.POINTER(_lib.StringDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_void_p]
    with Device() as dev:
        d = _lib.StringDesc()
        d.N, d.F, d.ncomp = s.N, s.F, s.ncomp
        d.no_self_field = int(bool(s.no_self_field))
        d.dx = float(s.dx)
        for name in ("a_1", "b_1", "m_xx_1", "m_yy_1", "a_2", "b_2", "m_xx_2", "m_yy_2", "h"):
            v = getattr(s, name)
            p = getattr(d, name)
            if np.isscalar(v):
                p.value, p.plane = float(v), None
            else:
                p.value = 0.0
                p.plane = C.cast(C.c_void_p(dev.put(v)), C.POINTER(C.c_double))
        for name in ("q_1", "q_2", "eta", "gamma", "delta"):
            setattr(d, name, float(getattr(s, name)))
        d.comp_domain = C.cast(C.c_void_p(dev.put(s.comp_domain, np.int32)),
                               C.POINTER(C.c_int32))
        d.sc_domain = C.cast(C.c_void_p(dev.put(s.sc_domain, np.int32)), C.POINTER(C.c_int32))
        d.segment_rows = rows
        a = dev.put(s.a)
        p1 = dev.put(s.psi_1)
        p2 = dev.put(s.psi_2) if s.multicomponent else None
        out = dev.put(np.full(s.F, np.nan))
        rc = fn(C.byref(d), a, p1, p2, out, None)
        return rc, dev.get(out, (s.F,))


def context_energy(s, rows=ROWS):
    from mepgl_amd.solver import Solver

    with Solver([s], precision="fp64", segment_rows=rows) as sv:
        return sv.computeFreeEnergy()[0]


def test_noalloc_energy_equals_context_energy():
    big = small_string(N=70, F=6, seed=91, multicomponent=True, masked=True)
    rng = np.random.default_rng(5)
    big.a_1 = -1.0 + 0.1 * rng.standard_normal((70, 70))
    big.h = 0.7 + 0.05 * rng.standard_normal((70, 70))
    small = small_string(N=20, F=3, seed=92)
    for k, s in enumerate((big, small)):
        say(f"[gpu] no-alloc energy call {k}: N={s.N} F={s.F} ncomp={s.ncomp}")
        rc, e = noalloc_energy(s)
        assert rc == 0
        np.testing.assert_array_equal(e, context_energy(s))


def test_noalloc_rejects_sizes_beyond_the_scratch():
    """F x tiles above the scratch capacity (65536 work items) is refused
    before any launch: N = 1024 at 16-row segments has 16 x 64 tiles per
    image, so F = 65 needs 66560 partials."""
    from mepgl_amd import _lib

    L = _lib.load()
    with Device() as dev:
        dom = dev.put(np.ones(4, np.int32), np.int32)  # never read
        d = _lib.StringDesc()
        d.N, d.F, d.ncomp, d.dx = 1024, 65, 1, 0.1
        d.b_1.value = d.m_xx_1.value = d.m_yy_1.value = 1.0
        d.a_1.value = -1.0
        d.comp_domain = C.cast(C.c_void_p(dom), C.POINTER(C.c_int32))
        d.sc_domain = d.comp_domain
        d.segment_rows = 16
        L.mepgl_free_energy_noalloc.argtypes = [C.POINTER(_lib.StringDesc)] + [C.c_void_p] * 5
        rc = L.mepgl_free_energy_noalloc(C.byref(d), dom, dom, None, dom, None)
    assert rc == _lib.ERR_ARG
