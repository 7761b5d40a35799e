"""The post-processing summary (post.py:125-152): boundary winding, flux,
barriers and the .npz written next to the simulation."""
import os

import numpy as np
import pytest

from conftest import small_string


def field(N, kind):
    """psi on an N x N grid in the file layout [i][j] (i along x)."""
    c = (np.arange(N) - (N - 1) / 2.0)
    x, y = np.meshgrid(c, c, indexing="ij")
    z = x + 1j * y
    if kind == "vortex":
        p = z / np.abs(z)
    elif kind == "antivortex":
        p = np.conj(z) / np.abs(z)
    elif kind == "double":
        p = (z / np.abs(z)) ** 2
    elif kind == "uniform":
        p = np.full((N, N), 0.7 + 0.2j)
    elif kind == "offcentre":
        w = z - (N / 4.0 + 0.3j)
        p = w / np.abs(w)
    return p.real[None], p.imag[None]


@pytest.mark.parametrize("kind,turns", [("vortex", 1), ("antivortex", -1), ("double", 2),
                                        ("uniform", 0), ("offcentre", 1)])
def test_winding_closed_forms(kind, turns):
    from mepgl_amd import observables

    u, v = field(64, kind)
    w = observables.winding(u, v)
    assert w.shape == (1,)
    # sum of sin(dtheta) over the boundary: 2 pi n up to O(dtheta^2)
    assert abs(w[0] - 2 * np.pi * turns) < 0.02 * max(1, abs(turns))


def test_winding_matches_edge_by_edge_walk():
    """Against an explicit walk of the four sides (brute force, 7x7)."""
    from mepgl_amd import observables

    rng = np.random.default_rng(3)
    u = rng.normal(size=(3, 7, 7)) + 2.0
    v = rng.normal(size=(3, 7, 7))
    N = 7
    path = [(i, 0) for i in range(N - 1)] + [(N - 1, j) for j in range(N - 1)] + \
           [(i, N - 1) for i in range(N - 1, 0, -1)] + [(0, j) for j in range(N - 1, 0, -1)]
    want = np.zeros(3)
    for n in range(3):
        for k, (i, j) in enumerate(path):
            i2, j2 = path[(k + 1) % len(path)]
            a, b = u[n, i, j], v[n, i, j]
            want[n] += (a * (v[n, i2, j2] - b) - b * (u[n, i2, j2] - a)) / (a * a + b * b)
    np.testing.assert_allclose(observables.winding(u, v), want, rtol=1e-13)


def test_winding_is_gauge_phase_invariant():
    from mepgl_amd import observables

    u, v = field(32, "vortex")
    p = (u + 1j * v) * np.exp(0.8j)
    np.testing.assert_allclose(observables.winding(p.real, p.imag),
                               observables.winding(u, v), rtol=1e-12)


def test_flux_closed_form():
    from mepgl_amd import observables

    b = np.full((2, 10, 10), 0.7)
    b[1] = 0.0
    np.testing.assert_allclose(observables.flux(b, 0.25), [0.7 * 100 * 0.0625, 0.0])


def test_barriers():
    from mepgl_amd import observables

    assert observables.barriers([1.0, 3.0, 2.0, 0.5]) == (2.0, 2.5)
    assert observables.barriers([3.0, 2.0, 1.0]) == (0.0, 2.0)


def test_precision_error():
    """Energies off by known amounts: per-image errors against the FP64
    values, and barriers whose errors follow from the closed forms
    (entry 2.0 vs 2.5, exit 2.5 vs 2.5)."""
    from mepgl_amd import observables

    e64 = np.array([1.0, 3.5, 2.0, 1.0])
    em = np.array([1.0, 3.5, 2.0, 0.5])
    r = observables.precision_error(em, e64)
    np.testing.assert_allclose(r["fenergy_rel_err"], [0.0, 0.0, 0.0, 0.5])
    np.testing.assert_allclose(r["barriers_mixed"], [2.5, 3.0])
    np.testing.assert_allclose(r["barriers_fp64"], [2.5, 2.5])
    np.testing.assert_allclose(r["barrier_rel_err"], [0.0, 0.2])
    # A zero FP64 value gives the absolute difference.
    z = observables.precision_error([0.25, 0.0], [0.0, 0.0])
    np.testing.assert_allclose(z["fenergy_rel_err"], [0.25, 0.0])
    np.testing.assert_allclose(z["barrier_rel_err"], [0.0, 0.25])


@pytest.mark.parametrize("mc", [False, True])
def test_summary_file(tmp_path, mc):
    """output_data fabricated with the layout writer, then summarised."""
    from mepgl_amd import npyio, observables

    s = small_string(N=12, F=3, multicomponent=mc, seed=8)
    d = npyio.write_input(s, "t", str(tmp_path))
    out = os.path.join(os.path.dirname(d), "output_data")
    os.rename(d, out)
    e = np.array([10.0, 12.5, 9.0])
    np.save(os.path.join(out, "fenergy.npy"), e)
    np.save(os.path.join(out, "s.npy"), np.array([0.0, 0.4, 1.0]))
    np.save(os.path.join(out, "s_jb.npy"), np.array([0.0, 0.5, 1.0]))
    for n in range(3):
        np.save(os.path.join(out, str(n), "b.npy"), np.full((12, 12), 0.1 * (n + 1)))
        for nm in ("jx_1", "jy_1") + (("jx_2", "jy_2") if mc else ()):
            np.save(os.path.join(out, str(n), nm + ".npy"), np.zeros((12, 12)))
    path = observables.write_summary("t", str(tmp_path))
    assert path.endswith(os.path.join("t", "t.npz"))
    z = np.load(path, allow_pickle=True)  # written by this package; w_2 may be None
    np.testing.assert_array_equal(z["fenergy"], e)
    np.testing.assert_array_equal(z["s_ph"], [0.0, 0.5, 1.0])
    np.testing.assert_allclose(z["phi"], [0.1 * (n + 1) * 144 * s.dx ** 2 for n in range(3)])
    u1 = np.swapaxes(s.psi_1[..., 0], 1, 2)
    v1 = np.swapaxes(s.psi_1[..., 1], 1, 2)
    np.testing.assert_allclose(z["w_1"], observables.winding(u1, v1))
    if mc:
        assert z["w_2"].shape == (3,)
    else:
        assert z["w_2"].item() is None
    assert "fenergy_rel_err" not in z
    # A mixed-precision run adds its energies against the FP64 path's.
    e64 = e * (1 + np.array([1e-7, -2e-7, 0.0]))
    np.save(os.path.join(out, "fp64_check.npy"), np.stack([e, e64], axis=1))
    z = np.load(observables.write_summary("t", str(tmp_path)), allow_pickle=True)
    np.testing.assert_allclose(z["fenergy_rel_err"], np.abs(e - e64) / np.abs(e64), rtol=1e-12)
    np.testing.assert_array_equal(z["barriers_fp64"], observables.barriers(e64))
