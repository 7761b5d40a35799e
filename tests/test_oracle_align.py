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
