"""Pinning the constant-phase alignment of neighbouring images and the
barrier observable."""
import numpy as np

from conftest import small_string


def test_alignment_removes_known_phases(oracle):
    s = small_string(N=12, F=5, seed=41, multicomponent=True)
    base1 = s.psi_1[0].copy()
    base2 = s.psi_2[0].copy()
    phases = np.array([0.0, 0.7, -1.1, 2.0, 0.4])
    for n, ph in enumerate(phases):
        for arr, base in ((s.psi_1, base1), (s.psi_2, base2)):
            z = (base[..., 0] + 1j * base[..., 1]) * np.exp(1j * ph)
            arr[n, ..., 0], arr[n, ..., 1] = z.real, z.imag
    p1, p2, th = oracle.gauge_align(s.N, s.F, True, s.psi_1, s.psi_2)
    np.testing.assert_allclose(th, -phases, atol=1e-12)
    for n in range(s.F):
        np.testing.assert_allclose(p1[n], base1, atol=1e-12)
        np.testing.assert_allclose(p2[n], base2, atol=1e-12)


def test_alignment_keeps_energy_and_gauge_invariant_metric(oracle):
    s = small_string(N=14, F=6, seed=42, multicomponent=True, masked=True)
    ref = oracle.OracleString(s)
    e0 = ref.free_energy()
    _, sjb0 = ref.arc_lengths()
    d_before = [np.sum((np.concatenate([ref.psi_1[n], ref.psi_2[n]]) -
                        np.concatenate([ref.psi_1[n - 1], ref.psi_2[n - 1]])) ** 2)
                for n in range(1, s.F)]
    ref.gauge_align()
    np.testing.assert_allclose(ref.free_energy(), e0, rtol=1e-13)
    _, sjb1 = ref.arc_lengths()
    np.testing.assert_allclose(sjb1, sjb0, rtol=1e-11, atol=1e-14)
    d_after = [np.sum((np.concatenate([ref.psi_1[n], ref.psi_2[n]]) -
                       np.concatenate([ref.psi_1[n - 1], ref.psi_2[n - 1]])) ** 2)
               for n in range(1, s.F)]
    assert all(a <= b * (1 + 1e-12) for a, b in zip(d_after, d_before))


def test_alignment_commutes_with_relaxation(oracle):
    """The gradient is equivariant under the rotation, so aligning fields,
    directions and gradients before a step equals aligning after it."""
    s = small_string(N=12, F=4, seed=43)
    a, b = oracle.OracleString(s), oracle.OracleString(s)
    a.nlcg_steps(2)
    b.nlcg_steps(2)
    th = a.gauge_align()
    a.nlcg_steps(2)
    b.nlcg_steps(2)
    z = (b.psi_1[..., 0] + 1j * b.psi_1[..., 1]) * np.exp(1j * th)[:, None, None]
    np.testing.assert_allclose(a.psi_1[..., 0], z.real, atol=1e-10)
    np.testing.assert_allclose(a.psi_1[..., 1], z.imag, atol=1e-10)


def test_barriers_closed_form():
    from mepgl_amd.observables import barriers

    e = np.array([3.0, 4.5, 6.0, 2.0, 1.0])
    assert barriers(e) == (3.0, 5.0)
    assert barriers([1.0, 0.5, 0.25]) == (0.0, 0.75)


# ---------------------------------------------------------------------------
# Local gauge alignment (or_gauge_align_local).

def _fwd(chi, dx):
    """Forward differences of (F, N, N) arrays as (F, N, N, 2), zero in the
    last column (x) and the last row (y)."""
    g = np.zeros(chi.shape + (2,))
    g[:, :, :-1, 0] = (chi[:, :, 1:] - chi[:, :, :-1]) / dx
    g[:, :-1, :, 1] = (chi[:, 1:, :] - chi[:, :-1, :]) / dx
    return g


def _fwd_T(v, dx):
    """The adjoint of _fwd applied to (F, N, N, 2)."""
    out = np.zeros(v.shape[:-1])
    out[:, :, :-1] -= v[:, :, :-1, 0] / dx
    out[:, :, 1:] += v[:, :, :-1, 0] / dx
    out[:, :-1, :] -= v[:, :-1, :, 1] / dx
    out[:, 1:, :] += v[:, :-1, :, 1] / dx
    return out


def _plaquette_curl(a, dx):
    """Forward curl of every plaquette (i, j)-(i+1, j+1): the quadrant-0 B of
    the discrete energy (nlcg.cu:191-203)."""
    ax, ay = a[..., 0], a[..., 1]
    return ((ay[:, :-1, 1:] - ay[:, :-1, :-1]) - (ax[:, 1:, :-1] - ax[:, :-1, :-1])) / dx


def _local(oracle, s):
    mc = s.multicomponent
    return oracle.gauge_align_local(s.N, s.F, mc, s.dx, s.q_1, s.q_2, s.a, s.psi_1,
                                    s.psi_2 if mc else None)


def test_local_alignment_solves_its_least_squares_problem(oracle):
    """xi_n = chi_n - chi_{n-1} minimises |A_n + D xi - A_{n-1}|^2: the normal
    equations D^T (A'_n - A'_{n-1}) = 0 hold for every pair, and xi equals a
    sparse direct solve of the same problem (mean fixed to zero)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    s = small_string(N=14, F=5, seed=44, multicomponent=True)
    a, p1, p2, chi, _, _ = _local(oracle, s)
    res = _fwd_T(a[1:] - a[:-1], s.dx)
    scale = np.max(np.abs(_fwd_T(s.a[1:] - s.a[:-1], s.dx)))
    assert np.max(np.abs(res)) < 1e-10 * scale
    # Independent solve: the least-squares problem for the pair (0, 1).
    N = s.N
    rows, cols, vals = [], [], []
    eq = 0
    rhs = []
    R = (s.a[0] - s.a[1]) * s.dx
    for j in range(N):
        for i in range(N):
            k = j * N + i
            if i < N - 1:
                rows += [eq, eq]; cols += [k + 1, k]; vals += [1.0, -1.0]
                rhs.append(R[j, i, 0]); eq += 1
            if j < N - 1:
                rows += [eq, eq]; cols += [k + N, k]; vals += [1.0, -1.0]
                rhs.append(R[j, i, 1]); eq += 1
    rows += [eq] * (N * N); cols += list(range(N * N)); vals += [1.0] * (N * N)
    rhs.append(0.0)  # zero mean
    Dm = sp.csr_matrix((vals, (rows, cols)), shape=(eq + 1, N * N))
    xi = spl.lsqr(Dm, np.array(rhs), atol=1e-15, btol=1e-15, iter_lim=20000)[0]
    np.testing.assert_allclose(chi[1].ravel(), xi, atol=1e-9 * np.max(np.abs(xi)))


def test_local_alignment_removes_a_pure_gauge_difference(oracle):
    """Images that differ only by gauge transformations (A_n = A_0 + D eta_n,
    psi_n = psi_0 exp(i q eta_n)) become equal to image 0, up to the
    constant phase the constant-phase alignment then removes."""
    s = small_string(N=16, F=4, seed=45, multicomponent=True)
    s.q_2 = 0.7
    rng = np.random.default_rng(46)
    x = np.arange(s.N) * s.dx
    X, Y = np.meshgrid(x, x)
    etas = [np.zeros((s.N, s.N))] + [0.8 * np.sin(0.3 * k * X + 0.2 * Y) + 0.3 * k * X * Y / 8
                                     + 0.05 * rng.standard_normal((s.N, s.N))
                                     for k in range(1, s.F)]
    for n in range(1, s.F):
        s.a[n] = s.a[0] + _fwd(etas[n][None], s.dx)[0]
        for arr, q in ((s.psi_1, s.q_1), (s.psi_2, s.q_2)):
            z = (arr[0, ..., 0] + 1j * arr[0, ..., 1]) * np.exp(1j * q * etas[n])
            arr[n, ..., 0], arr[n, ..., 1] = z.real, z.imag
    a, p1, p2, chi, _, _ = _local(oracle, s)
    for n in range(1, s.F):
        np.testing.assert_allclose(a[n], a[0], atol=1e-10)
        # chi_n = -eta_n up to a constant
        d = chi[n] + etas[n]
        np.testing.assert_allclose(d, np.mean(d), atol=1e-10)
    # Equal charges would leave one common constant phase; with q_2 != q_1 the
    # constant c_n rotates the components by q_1 c_n and q_2 c_n.
    for n in range(1, s.F):
        c = np.mean(chi[n] + etas[n])
        for arr, base, q in ((p1, s.psi_1[0], s.q_1), (p2, s.psi_2[0], s.q_2)):
            z = (base[..., 0] + 1j * base[..., 1]) * np.exp(1j * q * c)
            np.testing.assert_allclose(arr[n, ..., 0], z.real, atol=1e-10)
            np.testing.assert_allclose(arr[n, ..., 1], z.imag, atol=1e-10)


def test_local_alignment_invariants(oracle):
    """Plaquette curls and |psi| of every image unchanged; no neighbour A
    distance grows; the direction and gradients rotate with psi."""
    s = small_string(N=18, F=6, seed=47, multicomponent=True, masked=True)
    ref = oracle.OracleString(s)
    ref.nlcg_steps(3)
    a0, p10, p20 = ref.a.copy(), ref.psi_1.copy(), ref.psi_2.copy()
    d0 = ref.d_p1.copy()
    chi = ref.gauge_align_local()
    np.testing.assert_allclose(_plaquette_curl(ref.a, s.dx), _plaquette_curl(a0, s.dx),
                               atol=1e-11)
    for new, old in ((ref.psi_1, p10), (ref.psi_2, p20)):
        np.testing.assert_allclose(np.hypot(new[..., 0], new[..., 1]),
                                   np.hypot(old[..., 0], old[..., 1]), rtol=1e-13, atol=1e-15)
    before = np.sum((a0[1:] - a0[:-1]) ** 2, axis=(1, 2, 3))
    after = np.sum((ref.a[1:] - ref.a[:-1]) ** 2, axis=(1, 2, 3))
    assert np.all(after <= before * (1 + 1e-12))
    z = (d0[..., 0] + 1j * d0[..., 1]) * np.exp(1j * s.q_1 * chi)
    np.testing.assert_allclose(ref.d_p1[..., 0], z.real, atol=1e-12)
    np.testing.assert_allclose(ref.d_p1[..., 1], z.imag, atol=1e-12)


def test_local_gauge_transformation_changes_energy_at_first_order(oracle):
    """The discrete energy is gauge covariant only up to terms of first order
    in dx: each bond pairs A at one end with psi at both ends (nlcg.cu:148-185),
    so exp(i q chi) factors out of a transformed bond term only up to
    i q chi' (psi(i+1) - psi(i)) and chi'' dx psi, and the forward
    differences of chi stop at the last column and row.  A smooth gauge
    transformation of a smooth state (same physical box) changes the energy
    by an amount that halves when dx halves (measured ratios 1.89 and 1.95
    over dx = 0.4, 0.2, 0.1)."""
    from mepgl_amd import configs

    rel = []
    for N, dx in ((24, 0.4), (48, 0.2)):
        s = configs.vortex_entry(N, dx, 2, 0.6, seed=48, noise=0.0)
        x, y = configs.grid(N, dx)
        eta = 0.5 * np.sin(0.4 * x) * np.cos(0.3 * y) + 0.1 * x
        s.a[1] = s.a[0] + _fwd(eta[None], dx)[0]
        z = (s.psi_1[0, ..., 0] + 1j * s.psi_1[0, ..., 1]) * np.exp(1j * s.q_1 * eta)
        s.psi_1[1, ..., 0], s.psi_1[1, ..., 1] = z.real, z.imag
        e = oracle.free_energy(s)
        rel.append(abs(e[1] - e[0]) / abs(e[0]))
    ratio = rel[0] / rel[1]
    assert 1.6 < ratio < 2.6, (rel, ratio)
