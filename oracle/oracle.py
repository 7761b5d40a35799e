"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker and as the timed CPU
baseline; the product package never imports it.

Arrays use the reference's device layout: fields are float64 arrays of shape
(F, N, N, 2) indexed [n, j(y), i(x), component], per-site planes are (N, N)
indexed [j, i].  A "problem" is any object with the attributes of
mepgl_amd.problem.GLString (N, dx, F, multicomponent, no_self_field,
comp_domain, sc_domain, a_1 ... h, q_1 ... delta, a, psi_1, psi_2).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

TERM_NL, TERM_KIN, TERM_EM, TERM_ALL = 1, 2, 4, 7

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _System(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("dx", C.c_double), ("multicomponent", C.c_int),
        ("no_self_field", C.c_int), ("terms", C.c_int), ("poly_quirks", C.c_int),
        ("a_1", _dp), ("b_1", _dp), ("m_xx_1", _dp), ("m_yy_1", _dp),
        ("a_2", _dp), ("b_2", _dp), ("m_xx_2", _dp), ("m_yy_2", _dp),
        ("h", _dp),
        ("q_1", C.c_double), ("q_2", C.c_double), ("eta", C.c_double),
        ("gamma", C.c_double), ("delta", C.c_double),
        ("comp_nn", _ip), ("sc_nn", _ip),
    ]


class _State(C.Structure):
    _fields_ = [
        ("F", C.c_int),
        ("a", _dp), ("psi_1", _dp), ("psi_2", _dp),
        ("g_a", _dp * 2), ("g_p1", _dp * 2), ("g_p2", _dp * 2),
        ("d_a", _dp), ("d_p1", _dp), ("d_p2", _dp),
        ("g_norm2", _dp * 2),
        ("g_delta_g", _dp), ("alpha", _dp), ("fenergy", _dp), ("poly", _dp),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_line_search.restype = C.c_double
        L.or_line_search.argtypes = [C.c_double] * 5
        L.or_redistribute_cubic.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleSystem:
    """Per-string constants (parameters, neighbour maps) in oracle form."""

    def __init__(self, prob, terms: int = TERM_ALL, poly_quirks: bool = True):
        N = int(prob.N)
        self.N = N
        self.multicomponent = bool(prob.multicomponent)
        self._keep = {}
        planes = {}
        for name in ("a_1", "b_1", "m_xx_1", "m_yy_1", "a_2", "b_2", "m_xx_2", "m_yy_2", "h"):
            planes[name] = _f64(np.broadcast_to(np.asarray(getattr(prob, name), dtype=np.float64),
                                                (N, N)))
        self.comp_nn = build_nn_map(prob.comp_domain)
        self.sc_nn = build_nn_map(prob.sc_domain)
        self._keep.update(planes)
        self.s = _System(
            N=N, dx=float(prob.dx), multicomponent=int(self.multicomponent),
            no_self_field=int(bool(prob.no_self_field)), terms=int(terms),
            poly_quirks=int(bool(poly_quirks)),
            q_1=float(prob.q_1), q_2=float(prob.q_2), eta=float(prob.eta),
            gamma=float(prob.gamma), delta=float(prob.delta),
            comp_nn=_p(self.comp_nn), sc_nn=_p(self.sc_nn),
            **{k: _p(v) for k, v in planes.items()},
        )

    @property
    def ref(self):
        return C.byref(self.s)


def build_nn_map(domain) -> np.ndarray:
    d = np.ascontiguousarray(domain, dtype=np.int32)
    N = d.shape[0]
    out = np.zeros((N, N), dtype=np.int32)
    lib().or_build_nn_map(C.c_int(N), _p(d), _p(out))
    return out


def _fields(prob, F=None):
    F = prob.F if F is None else F
    N = prob.N
    a = _f64(prob.a).reshape(F, N, N, 2)
    p1 = _f64(prob.psi_1).reshape(F, N, N, 2)
    p2 = prob.psi_2
    p2 = _f64(p2).reshape(F, N, N, 2) if p2 is not None else np.zeros((F, N, N, 2))
    return a, p1, p2


def free_energy(prob, terms=TERM_ALL, sysobj=None):
    so = sysobj or OracleSystem(prob, terms)
    a, p1, p2 = _fields(prob)
    out = np.zeros(prob.F)
    lib().or_free_energy(so.ref, C.c_int(prob.F), _p(a), _p(p1), _p(p2), _p(out))
    return out


def free_energy_density(prob, terms=TERM_ALL):
    so = OracleSystem(prob, terms)
    a, p1, p2 = _fields(prob)
    out = np.zeros((prob.F, prob.N, prob.N))
    lib().or_free_energy_density(so.ref, C.c_int(prob.F), _p(a), _p(p1), _p(p2), _p(out))
    return out


def gradient(prob, terms=TERM_ALL, maxwell_solver=False, sysobj=None):
    """Return (a_g, psi_1_g, psi_2_g), each (F, N, N, 2), starting from zero
    buffers (the loop's cudaMemset initialisation, mepgl.cu:124-133)."""
    so = sysobj or OracleSystem(prob, terms)
    a, p1, p2 = _fields(prob)
    shape = a.shape
    ga, g1, g2 = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    lib().or_gradient(so.ref, C.c_int(prob.F), _p(a), _p(p1), _p(p2), _p(ga), _p(g1), _p(g2),
                      C.c_int(int(maxwell_solver)))
    return ga, g1, g2


def compute_jb(prob):
    so = OracleSystem(prob)
    a, p1, p2 = _fields(prob)
    F, N = prob.F, prob.N
    b = np.zeros((F, N, N))
    j1 = np.zeros((F, N, N, 2))
    j2 = np.zeros((F, N, N, 2))
    lib().or_compute_jb(so.ref, C.c_int(F), _p(a), _p(p1), _p(p2), _p(b), _p(j1), _p(j2))
    return b, j1, j2


def poly_expansion(prob, d_a, d_p1, d_p2, poly_quirks=True, terms=TERM_ALL):
    """Per-image sums (A, B, C, D, E) of the line-search quartic."""
    so = OracleSystem(prob, terms, poly_quirks)
    a, p1, p2 = _fields(prob)
    F, N = prob.F, prob.N
    da, d1 = _f64(d_a).reshape(F, N, N, 2), _f64(d_p1).reshape(F, N, N, 2)
    d2 = _f64(d_p2).reshape(F, N, N, 2) if d_p2 is not None else np.zeros((F, N, N, 2))
    dens = np.zeros((F, N, N))
    poly = np.zeros((F, N, N, 4))
    lib().or_poly_expansion(so.ref, C.c_int(F), _p(a), _p(p1), _p(p2), _p(da), _p(d1), _p(d2),
                            _p(dens), _p(poly))
    out = np.zeros((F, 5))
    out[:, 0] = dens.reshape(F, -1).sum(axis=1)
    out[:, 1:] = poly.reshape(F, -1, 4).sum(axis=1)
    return out


def line_search(A, B, C_, D, E) -> float:
    return lib().or_line_search(float(A), float(B), float(C_), float(D), float(E))


def arc_length(N, F, multicomponent, a, p1, p2):
    s = np.zeros(F)
    lib().or_arc_length(C.c_int(N), C.c_int(F), C.c_int(int(multicomponent)), _p(_f64(a)),
                        _p(_f64(p1)), _p(_f64(p2)), _p(s))
    return s


def arc_length_jb(N, F, multicomponent, b, j1, j2):
    s = np.zeros(F)
    lib().or_arc_length_jb(C.c_int(N), C.c_int(F), C.c_int(int(multicomponent)), _p(_f64(b)),
                           _p(_f64(j1)), _p(_f64(j2)), _p(s))
    return s


def redistribute_linear(N, F, s, f, F_new):
    out = np.zeros((F_new, N, N, 2))
    lib().or_redistribute_linear(C.c_int(N), C.c_int(F), _p(_f64(s)), _p(_f64(f)),
                                 C.c_int(F_new), _p(out))
    return out


def redistribute_cubic(N, F, s, f, F_new, textbook=False):
    out = np.zeros((F_new, N, N, 2))
    rc = lib().or_redistribute_cubic(C.c_int(N), C.c_int(F), _p(_f64(s)), _p(_f64(f)),
                                     C.c_int(F_new), _p(out), C.c_int(int(textbook)))
    if rc != 0:
        raise ValueError("cubic redistribution needs at least 4 images")
    return out


def gauge_align(N, F, multicomponent, p1, p2):
    """Constant-phase alignment; returns (psi_1, psi_2, theta)."""
    p1 = _f64(p1).copy()
    p2 = _f64(p2).copy() if p2 is not None else np.zeros_like(p1)
    th = np.zeros(F)
    lib().or_gauge_align(C.c_int(N), C.c_int(F), C.c_int(int(multicomponent)), _p(p1), _p(p2),
                         _p(th))
    return p1, p2, th


def gauge_align_local(N, F, multicomponent, dx, q1, q2, a, p1, p2, rot_1=(), rot_2=()):
    """Local gauge alignment (or_gauge_align_local) on copies; returns
    (a, psi_1, psi_2, chi, rotated state arrays of component 1, of
    component 2).  chi is (F, N, N)."""
    a = _f64(a).copy()
    p1 = _f64(p1).copy()
    p2 = _f64(p2).copy() if multicomponent else None
    r1 = [_f64(r).copy() for r in rot_1]
    r2 = [_f64(r).copy() for r in rot_2] if multicomponent else [None] * len(r1)
    chi = np.zeros((F, N, N))
    n = len(r1)
    arr1 = (_dp * max(n, 1))(*[_p(r) for r in r1])
    arr2 = (_dp * max(n, 1))(*[(_p(r) if r is not None else None) for r in r2])
    lib().or_gauge_align_local(C.c_int(N), C.c_int(F), C.c_int(int(multicomponent)),
                               C.c_double(dx), C.c_double(q1), C.c_double(q2), _p(a), _p(p1),
                               _p(p2) if p2 is not None else None, arr1, arr2, C.c_int(n),
                               _p(chi))
    return a, p1, p2, chi, r1, r2


def thomas(dl, d, du, rhs):
    x = _f64(rhs).copy()
    lib().or_thomas(C.c_int(len(x)), _p(_f64(dl)), _p(_f64(d)), _p(_f64(du)), _p(x))
    return x


class OracleString:
    """The per-string solver state of the reference loop (mepgl.cu:38-140):
    fields, two gradient buffers, direction and per-image scalars, all
    starting as the reference initialises them (zero buffers, infinite
    gradient norms)."""

    def __init__(self, prob, poly_quirks=True, cubic_textbook=False):
        self.prob = prob
        self.sys = OracleSystem(prob, TERM_ALL, poly_quirks)
        self.cubic_textbook = cubic_textbook
        self.N, self.F = prob.N, prob.F
        self.multicomponent = bool(prob.multicomponent)
        a, p1, p2 = _fields(prob)
        self.a, self.psi_1, self.psi_2 = a.copy(), p1.copy(), p2.copy()
        self._alloc()

    def _alloc(self):
        F, N = self.F, self.N
        shape = (F, N, N, 2)
        self.g_a = [np.zeros(shape), np.zeros(shape)]
        self.g_p1 = [np.zeros(shape), np.zeros(shape)]
        self.g_p2 = [np.zeros(shape), np.zeros(shape)]
        self.d_a, self.d_p1, self.d_p2 = np.zeros(shape), np.zeros(shape), np.zeros(shape)
        self.g_norm2 = [np.full(F, np.inf), np.full(F, np.inf)]
        self.g_delta_g = np.zeros(F)
        self.alpha = np.zeros(F)
        self.fenergy = np.zeros(F)
        self.poly = np.zeros((F, 4))
        self.s = np.zeros(F)
        self.s_jb = np.zeros(F)

    def _state(self):
        st = _State(F=self.F, a=_p(self.a), psi_1=_p(self.psi_1), psi_2=_p(self.psi_2),
                    d_a=_p(self.d_a), d_p1=_p(self.d_p1), d_p2=_p(self.d_p2),
                    g_delta_g=_p(self.g_delta_g), alpha=_p(self.alpha),
                    fenergy=_p(self.fenergy), poly=_p(self.poly))
        for k in range(2):
            st.g_a[k] = _p(self.g_a[k])
            st.g_p1[k] = _p(self.g_p1[k])
            st.g_p2[k] = _p(self.g_p2[k])
            st.g_norm2[k] = _p(self.g_norm2[k])
        return st

    def nlcg_steps(self, iter_n, nlcg=True, maxwell_solver=False):
        """nlcgSteps, nlcg.cu:1462-1586."""
        st = self._state()
        lib().or_nlcg_steps(self.sys.ref, C.byref(st), C.c_int(iter_n), C.c_int(int(nlcg)),
                            C.c_int(int(maxwell_solver)))

    def free_energy(self):
        out = np.zeros(self.F)
        lib().or_free_energy(self.sys.ref, C.c_int(self.F), _p(self.a), _p(self.psi_1),
                             _p(self.psi_2), _p(out))
        return out

    def compute_jb(self):
        F, N = self.F, self.N
        b = np.zeros((F, N, N))
        j1, j2 = np.zeros((F, N, N, 2)), np.zeros((F, N, N, 2))
        lib().or_compute_jb(self.sys.ref, C.c_int(F), _p(self.a), _p(self.psi_1), _p(self.psi_2),
                            _p(b), _p(j1), _p(j2))
        return b, j1, j2

    def arc_lengths(self):
        """arcLength + computeJB + arcLengthJB, mepgl.cu:193-201."""
        self.s = arc_length(self.N, self.F, self.multicomponent, self.a, self.psi_1, self.psi_2)
        b, j1, j2 = self.compute_jb()
        self.s_jb = arc_length_jb(self.N, self.F, self.multicomponent, b, j1, j2)
        return self.s, self.s_jb

    def redistribute(self, cubic=False, F_new=None):
        """redistributeLinear / redistributeCubic on s_jb (mepgl.cu:211, 264)."""
        F_new = self.F if F_new is None else F_new
        fields = [self.a, self.psi_1] + ([self.psi_2] if self.multicomponent else [])
        new = []
        for f in fields:
            if cubic:
                new.append(redistribute_cubic(self.N, self.F, self.s_jb, f, F_new,
                                              self.cubic_textbook))
            else:
                new.append(redistribute_linear(self.N, self.F, self.s_jb, f, F_new))
        if F_new != self.F:
            self.F = F_new
            self._alloc()
            if not self.multicomponent:
                new.append(np.zeros_like(new[0]))
        elif not self.multicomponent:
            new.append(self.psi_2)
        self.a, self.psi_1, self.psi_2 = new

    def gauge_align(self):
        """Constant-phase alignment of neighbouring images, applied to the
        fields, the direction and both gradient buffers."""
        p1, p2, th = gauge_align(self.N, self.F, self.multicomponent, self.psi_1,
                                 self.psi_2 if self.multicomponent else None)
        self.psi_1 = p1
        if self.multicomponent:
            self.psi_2 = p2
        rot = np.exp(1j * th)[:, None, None]

        def rotate(arr):
            z = (arr[..., 0] + 1j * arr[..., 1]) * rot
            arr[..., 0], arr[..., 1] = z.real, z.imag

        for arr in [self.d_p1, self.g_p1[0], self.g_p1[1]] + \
                ([self.d_p2, self.g_p2[0], self.g_p2[1]] if self.multicomponent else []):
            rotate(arr)
        return th

    def gauge_align_local(self):
        """Local gauge alignment of neighbouring images (or_gauge_align_local):
        fields, and the psi parts of the direction and both gradient
        buffers.  Strings with a fixed A (no self field) are left as they
        are.  Returns chi (F, N, N)."""
        if self.prob.no_self_field:
            return np.zeros((self.F, self.N, self.N))
        mc = self.multicomponent
        r1 = [self.d_p1, self.g_p1[0], self.g_p1[1]]
        r2 = [self.d_p2, self.g_p2[0], self.g_p2[1]] if mc else []
        a, p1, p2, chi, r1, r2 = gauge_align_local(
            self.N, self.F, mc, self.prob.dx, self.prob.q_1, self.prob.q_2, self.a, self.psi_1,
            self.psi_2 if mc else None, r1, r2)
        self.a, self.psi_1 = a, p1
        self.d_p1, self.g_p1[0], self.g_p1[1] = r1
        if mc:
            self.psi_2 = p2
            self.d_p2, self.g_p2[0], self.g_p2[1] = r2
        return chi

    def reset_interior_beta(self):
        """mepgl.cu:231-234."""
        self.g_norm2[0][1:-1] = np.inf
        self.g_norm2[1][1:-1] = np.inf

    def loop(self, iterations, relaxation_steps, mode="L", nlcg=True, align=False,
             align_local=False):
        """The outer loop of mepgl.cu:148-360 (without printing); align adds
        the optional constant-phase alignment before each arclength, and
        align_local the local gauge alignment before that."""
        for _ in range(iterations):
            if mode == "S":
                self.nlcg_steps(relaxation_steps, nlcg, False)
            elif mode == "M":
                self.nlcg_steps(relaxation_steps, nlcg, True)
            elif mode in ("L", "C"):
                if align_local:
                    self.gauge_align_local()
                if align:
                    self.gauge_align()
                self.arc_lengths()
                self.redistribute(cubic=(mode == "C"))
                self.reset_interior_beta()
                self.nlcg_steps(relaxation_steps, nlcg, False)
            else:
                raise ValueError(mode)
