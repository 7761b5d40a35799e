"""Pinning the oracle's line-search quartic, line search, NLCG step pairing,
arclength and reparametrisation (oracle/mepgl_oracle.c) against exact
expansions, independent scans, closed forms and banded solves."""
import math

import numpy as np
import pytest
from scipy.linalg import solve_banded

from conftest import small_string

M = float(np.float32(1.025))


def random_direction(s, seed, with_a=True):
    rng = np.random.default_rng(seed)
    da = 0.05 * rng.standard_normal(s.a.shape) if with_a else np.zeros(s.a.shape)
    d1 = 0.05 * rng.standard_normal(s.psi_1.shape)
    d2 = 0.05 * rng.standard_normal(s.psi_1.shape) if s.multicomponent else None
    return da, d1, d2


def energy_along(oracle, s, d, alpha):
    t = s.copy()
    t.a = s.a + alpha * d[0]
    t.psi_1 = s.psi_1 + alpha * d[1]
    if s.multicomponent:
        t.psi_2 = s.psi_2 + alpha * d[2]
    return oracle.free_energy(t)


def vandermonde_fit(oracle, s, d):
    al = np.array([-0.2, -0.1, 0.0, 0.1, 0.2])
    E = np.stack([energy_along(oracle, s, d, a) for a in al])  # (5, F)
    V = np.vander(al, 5, increasing=True)
    return np.linalg.solve(V, E).T  # (F, 5): A..E


@pytest.mark.parametrize("mc", [False, True])
def test_exact_quartic_is_the_energy_polynomial(oracle, mc):
    s = small_string(N=12, F=2, multicomponent=mc, seed=31)
    d = random_direction(s, 1)
    got = oracle.poly_expansion(s, *d, poly_quirks=False)
    fit = vandermonde_fit(oracle, s, d)
    scale = np.max(np.abs(fit), axis=1, keepdims=True)
    assert np.max(np.abs(got - fit) / scale) < 1e-9


def test_reference_quartic_equals_exact_when_a_is_frozen(oracle):
    s = small_string(N=12, F=2, multicomponent=True, seed=32)
    d = random_direction(s, 2, with_a=False)
    q = oracle.poly_expansion(s, *d, poly_quirks=True)
    e = oracle.poly_expansion(s, *d, poly_quirks=False)
    np.testing.assert_allclose(q, e, rtol=1e-13, atol=1e-15)
    fit = vandermonde_fit(oracle, s, d)
    assert np.max(np.abs(q - fit)) / np.max(np.abs(fit)) < 1e-9


def test_reference_quartic_differs_when_a_moves(oracle):
    """The reference's x-bond first-order term (nlcg.cu:1162-1165) and the
    y-bond cubic factor (nlcg.cu:1251) change B and D once d_A != 0; A, C
    (apart from the x-bond cross term) and E keep their exact values."""
    s = small_string(N=12, F=2, seed=33)
    d = random_direction(s, 3)
    q = oracle.poly_expansion(s, *d, poly_quirks=True)
    e = oracle.poly_expansion(s, *d, poly_quirks=False)
    np.testing.assert_allclose(q[:, 0], e[:, 0], rtol=1e-14)
    np.testing.assert_allclose(q[:, 4], e[:, 4], rtol=1e-14)
    assert np.all(np.abs(q[:, 1] - e[:, 1]) > 1e-12 * np.abs(e[:, 1]))
    assert np.all(np.abs(q[:, 3] - e[:, 3]) > 1e-12 * np.abs(e[:, 3]))


def with_charges(s, q):
    t = s.copy()
    t.q_1 = q
    if t.multicomponent:
        t.q_2 = q
    return t.validate()


@pytest.mark.parametrize("mc", [False, True])
def test_reference_x_bond_term_at_charge_minus_one(oracle, mc):
    """Q1 (nlcg.cu:1162-1165, 1269-1272): the reference's x-bond first-order
    term is +a_d,x psi where the exact expansion has -q a_d,x psi.  At
    q = -1 the two coincide, so with the y-bond cubic factor out of play
    (a_d,y = 0) the reference quartic is the energy polynomial.  The term
    with the charge kept, or with the exact sign, would differ here."""
    s = with_charges(small_string(N=12, F=2, multicomponent=mc, seed=35), -1.0)
    da, d1, d2 = random_direction(s, 5)
    da[..., 1] = 0.0
    d = (da, d1, d2)
    q = oracle.poly_expansion(s, *d, poly_quirks=True)
    e = oracle.poly_expansion(s, *d, poly_quirks=False)
    np.testing.assert_allclose(q, e, rtol=1e-12, atol=1e-14 * np.max(np.abs(e)))
    fit = vandermonde_fit(oracle, s, d)
    scale = np.max(np.abs(fit), axis=1, keepdims=True)
    assert np.max(np.abs(q - fit) / scale) < 1e-9
    # The same direction at the charge +1 does move B (the test has teeth).
    s1 = with_charges(s, 1.0)
    q1 = oracle.poly_expansion(s1, *d, poly_quirks=True)
    e1 = oracle.poly_expansion(s1, *d, poly_quirks=False)
    assert np.all(np.abs(q1[:, 1] - e1[:, 1]) > 1e-9 * np.abs(e1[:, 1]))


@pytest.mark.parametrize("mc", [False, True])
def test_reference_y_bond_cubic_is_halved(oracle, mc):
    """Q2 (nlcg.cu:1251, 1358): the reference's y-bond cubic term carries an
    extra factor 0.5.  With a_d,x = 0 the x-bond terms are exact.  The only
    part of the exact D that needs a_d,y is the y-bond cubic term, so that
    term is D(d) - D(d with a_d,y = 0), both from Vandermonde fits of the
    energy.  The reference's D is then D(d) minus half of it, and A, B, C
    and E are exact."""
    s = small_string(N=12, F=2, multicomponent=mc, seed=36)
    da, d1, d2 = random_direction(s, 6)
    da[..., 0] = 0.0
    da[..., 1] *= 10.0  # a y-bond cubic term well above the fits' rounding
    d = (da, d1, d2)
    d0 = (np.zeros_like(da), d1, d2)
    q = oracle.poly_expansion(s, *d, poly_quirks=True)
    e = oracle.poly_expansion(s, *d, poly_quirks=False)
    fit = vandermonde_fit(oracle, s, d)
    fit0 = vandermonde_fit(oracle, s, d0)
    ybond = fit[:, 3] - fit0[:, 3]
    scale = np.max(np.abs(fit), axis=1)
    # The term stands a thousand times above the fits' tolerance, so a wrong
    # factor on it cannot pass the check below.
    assert np.all(np.abs(ybond) > 1e-6 * scale)
    want = fit[:, 3] - 0.5 * ybond
    assert np.max(np.abs(q[:, 3] - want) / scale) < 1e-9
    keep = [0, 1, 2, 4]
    np.testing.assert_allclose(q[:, keep], e[:, keep], rtol=1e-12,
                               atol=1e-14 * np.max(np.abs(e)))


def test_quartic_constant_term_is_the_energy(oracle):
    s = small_string(N=12, F=3, multicomponent=True, masked=True, seed=34)
    d = random_direction(s, 4)
    q = oracle.poly_expansion(s, *d)
    np.testing.assert_allclose(q[:, 0], oracle.free_energy(s), rtol=1e-13)


def scan(A, B, C, D, E):
    """Independent statement of the scan: candidates x_k = 1e-10 m^k,
    first k with f(x_{k+1}) > f(x_k), alpha = x_{k+1}/m."""
    xs = [1e-10]
    for _ in range(2000):
        xs.append(xs[-1] * M)

    def f(x):
        return (((E * math.pow(x, 4) + D * x * x * x) + C * x * x) + B * x) + A

    for k in range(2000):
        if f(xs[k + 1]) > f(xs[k]):
            return xs[k + 1] / M
    return xs[2000] / M


def test_line_search_against_independent_scan(oracle):
    rng = np.random.default_rng(7)
    for _ in range(300):
        A = rng.normal() * 10
        B = -abs(rng.normal()) * 10 ** rng.uniform(-3, 2)
        C = abs(rng.normal()) * 10 ** rng.uniform(-2, 2)
        D, E = rng.normal(size=2) * 10 ** rng.uniform(-3, 1, size=2)
        assert oracle.line_search(A, B, C, D, E) == scan(A, B, C, D, E)


def test_line_search_special_cases(oracle):
    # Increasing from the start: the first candidate pair already rises.
    assert oracle.line_search(1.0, 1.0, 1.0, 0.0, 0.0) == (1e-10 * M) / M
    # Never rising (linear decrease): the scan runs to the end.
    x = 1e-10
    for _ in range(2000):
        x *= M
    assert oracle.line_search(0.0, -1.0, 0.0, 0.0, 0.0) == x / M
    assert 2.5e11 < x < 3.1e11
    # The candidate table reaches 1 near k = 933.
    k = math.ceil(math.log(1e10) / math.log(M))
    assert 925 < k < 940
    # Pure quadratic: the result brackets the minimiser -B/(2C).
    for B, C in ((-2.0, 3.0), (-1e-3, 5.0), (-40.0, 0.5)):
        a = oracle.line_search(7.0, B, C, 0.0, 0.0)
        f = lambda t: 7.0 + B * t + C * t * t  # noqa: E731
        assert f(a) <= f(a / M) and f(a * M) > f(a)
        astar = -B / (2 * C)
        assert astar / M ** 2 < a < astar * M ** 2


def test_first_step_is_steepest_descent_and_lowers_energy(oracle):
    s = small_string(N=14, F=3, seed=35)
    st = oracle.OracleString(s)
    g = oracle.gradient(s)
    e0 = oracle.free_energy(s)
    st.nlcg_steps(1)
    np.testing.assert_array_equal(st.d_p1, -g[1])
    np.testing.assert_array_equal(st.d_a, -g[0])
    assert np.all(st.free_energy() < e0)
    # x' = x + alpha d with one rounding (fused multiply-add), checked with
    # exact rational arithmetic.
    from fractions import Fraction

    def fma_exact(a, b, c):
        return float(Fraction(a) * Fraction(b) + Fraction(c))

    for n in range(s.F):
        np.testing.assert_array_equal(st.psi_1[n], np.vectorize(fma_exact)(
            st.alpha[n], -g[1][n], s.psi_1[n]))


def test_polak_ribiere_beta_closed_form(oracle):
    """Second step of one call: beta = max(g1.(g1 - g0) / |g0|^2, 0) from the
    gradients at x0 and x1, and d1 = -g1 + beta d0 with d0 = -g0."""
    s = small_string(N=14, F=2, seed=36)
    one = oracle.OracleString(s)
    one.nlcg_steps(1)
    t = s.copy()
    t.a, t.psi_1 = one.a.copy(), one.psi_1.copy()
    g0 = oracle.gradient(s)
    g1 = oracle.gradient(t)
    two = oracle.OracleString(s)
    two.nlcg_steps(2)
    for n in range(s.F):
        num = sum(np.sum(g1[k][n] * (g1[k][n] - g0[k][n])) for k in range(2))
        den = sum(np.sum(g0[k][n] ** 2) for k in range(2))
        beta = max(num / den, 0.0)
        assert beta > 0
        for k, dk in ((0, two.d_a), (1, two.d_p1)):
            np.testing.assert_allclose(dk[n], -g1[k][n] - beta * g0[k][n], rtol=1e-12,
                                       atol=1e-15 * np.max(np.abs(g1[k][n])))


def test_call_boundary_pairing(oracle):
    """With an even step count a second call continues exactly where one
    long call would; with an odd count the next call's first step pairs with
    the gradient from two steps back (the pointer swap acts on local copies)."""
    s = small_string(N=12, F=2, seed=37)
    one, two = oracle.OracleString(s), oracle.OracleString(s)
    one.nlcg_steps(8)
    two.nlcg_steps(4)
    two.nlcg_steps(4)
    np.testing.assert_array_equal(one.psi_1, two.psi_1)
    one, two = oracle.OracleString(s), oracle.OracleString(s)
    one.nlcg_steps(6)
    two.nlcg_steps(3)
    two.nlcg_steps(3)
    assert not np.array_equal(one.psi_1, two.psi_1)


def test_arclength_of_affine_string(oracle):
    """Images affine in n are equally spaced in both metrics: the plain
    metric directly, and the (b, j) metric for psi = 0 with A affine in n."""
    N, F = 10, 7
    rng = np.random.default_rng(11)
    base = rng.standard_normal((N, N, 2))
    step = rng.standard_normal((N, N, 2))
    a = np.stack([base + n * step for n in range(F)])
    p1 = np.stack([0.3 * base + 0.5 * n * step for n in range(F)])
    s = oracle.arc_length(N, F, False, a, p1, np.zeros_like(a))
    np.testing.assert_allclose(s, np.arange(F) / (F - 1), rtol=1e-12, atol=1e-15)
    assert s[-1] == 1.0
    b = np.stack([base[..., 0] + n * step[..., 0] for n in range(F)])
    sj = oracle.arc_length_jb(N, F, False, b, np.zeros_like(a), np.zeros_like(a))
    np.testing.assert_allclose(sj, np.arange(F) / (F - 1), rtol=1e-12, atol=1e-15)


def test_linear_redistribution_is_exact_on_linear_data(oracle):
    N, F = 6, 8
    rng = np.random.default_rng(12)
    sn = np.sort(rng.uniform(0, 1, F))
    sn[0], sn[-1] = 0.0, 1.0
    c0, c1 = rng.standard_normal((N, N, 2)), rng.standard_normal((N, N, 2))
    f = np.stack([c0 + c1 * x for x in sn])
    for F_new in (F, 11, 4):
        out = oracle.redistribute_linear(N, F, sn, f, F_new)
        for n in range(F_new):
            np.testing.assert_allclose(out[n], c0 + c1 * (n / (F_new - 1)), rtol=1e-12,
                                       atol=1e-13)
    uniform = np.arange(F) / (F - 1)
    np.testing.assert_array_equal(oracle.redistribute_linear(N, F, uniform, f, F), f)


def test_thomas_against_banded_solve(oracle):
    rng = np.random.default_rng(13)
    for n in (4, 7, 30):
        dl, du = rng.uniform(0.1, 1, n), rng.uniform(0.1, 1, n)
        d = dl + du + rng.uniform(0.5, 1.5, n)
        dl[0] = du[-1] = 0.0
        rhs = rng.standard_normal(n)
        ab = np.zeros((3, n))
        ab[0, 1:] = du[:-1]
        ab[1] = d
        ab[2, :-1] = dl[1:]
        np.testing.assert_allclose(oracle.thomas(dl, d, du, rhs), solve_banded((1, 1), ab, rhs),
                                   rtol=1e-12)


def test_spline_system_matches_banded_solve(oracle):
    """The first and last rows of the spline system are not diagonally
    dominant; the textbook solve still matches a pivoting banded solver."""
    rng = np.random.default_rng(14)
    F = 9
    h = np.concatenate([[0.0], rng.uniform(0.05, 0.3, F - 1)])
    dl, d, du = np.zeros(F), np.zeros(F), np.zeros(F)
    d[0], du[0] = -h[1], h[1]
    dl[-1], d[-1] = h[-1], -h[-1]
    for n in range(1, F - 1):
        dl[n], d[n], du[n] = h[n], 2 * (h[n] + h[n + 1]), h[n + 1]
    rhs = rng.standard_normal(F)
    ab = np.zeros((3, F))
    ab[0, 1:] = du[:-1]
    ab[1] = d
    ab[2, :-1] = dl[1:]
    np.testing.assert_allclose(oracle.thomas(dl, d, du, rhs), solve_banded((1, 1), ab, rhs),
                               rtol=1e-10)


@pytest.mark.parametrize("textbook", [False, True])
def test_cubic_redistribution_interpolates_nodes(oracle, textbook):
    N, F = 4, 9
    rng = np.random.default_rng(15)
    sn = np.arange(F) / (F - 1)
    f = rng.standard_normal((F, N, N, 2))
    out = oracle.redistribute_cubic(N, F, sn, f, F, textbook=textbook)
    np.testing.assert_allclose(out, f, rtol=1e-12, atol=1e-12)


def test_reference_cubic_sigma_sign_on_quadratics(oracle):
    """Q7 (interpolation.cu:394-395, 399-400): the reference weighs the
    sigma_{k-1} term by (w-1)^3 - (w-1) where the Forsythe-Malcolm-Moler
    spline has (1-w)^3 - (1-w).  On data c0 + c2 s^2 the spline's second
    derivative is 2 c2 at every node (the end conditions reproduce cubics),
    so the flipped term moves the value at w by exactly
    -2 (h^2 / 6) (2 c2) ((1-w)^3 - (1-w)), with h the segment length and w
    the fraction of the segment from node k-1."""
    N, F, F_new = 3, 9, 14
    rng = np.random.default_rng(17)
    sn = np.sort(rng.uniform(0, 1, F))
    sn[0], sn[-1] = 0.0, 1.0
    c0 = rng.standard_normal((N, N, 2))
    c2 = rng.standard_normal((N, N, 2))
    f = np.stack([c0 + c2 * x ** 2 for x in sn])
    out = oracle.redistribute_cubic(N, F, sn, f, F_new, textbook=False)
    for n in range(F_new):
        x = n / (F_new - 1)
        k = min(max(int(np.searchsorted(sn, x, side="right")), 1), F - 1)
        h = sn[k] - sn[k - 1]
        w = (x - sn[k - 1]) / h
        want = c0 + c2 * x ** 2 - 2.0 * (h * h / 6.0) * (2.0 * c2) * ((1 - w) ** 3 - (1 - w))
        np.testing.assert_allclose(out[n], want, rtol=1e-10, atol=1e-10)


def test_textbook_cubic_reproduces_cubic_polynomials(oracle):
    """The Forsythe-Malcolm-Moler end conditions make the spline exact on
    cubics; the reference's sign of the sigma_{k-1} weight does not."""
    N, F = 3, 10
    rng = np.random.default_rng(16)
    sn = np.sort(rng.uniform(0, 1, F))
    sn[0], sn[-1] = 0.0, 1.0
    co = rng.standard_normal((4, N, N, 2))
    p = lambda x: co[0] + co[1] * x + co[2] * x ** 2 + co[3] * x ** 3  # noqa: E731
    f = np.stack([p(x) for x in sn])
    out = oracle.redistribute_cubic(N, F, sn, f, 13, textbook=True)
    for n in range(13):
        np.testing.assert_allclose(out[n], p(n / 12), rtol=1e-10, atol=1e-10)
    ref = oracle.redistribute_cubic(N, F, sn, f, 13, textbook=False)
    assert np.max(np.abs(ref - out)) > 1e-6
