"""mepgl_amd.builder against vectors of the reference's StringBuilder
(tests/golden/builder_vectors.npz, written by
tests/golden/make_builder_vectors.py from mepgl/mepgl_lib/builder.py), plus
closed forms that hold whatever the noise."""
import json
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)


def _vectors():
    z = np.load(os.path.join(GOLDEN, "builder_vectors.npz"), allow_pickle=False)
    return json.loads(str(z["cases"])), z


@pytest.mark.parametrize("k", [0, 1])
def test_builder_matches_reference_vectors(k):
    from builder_cases import build

    from mepgl_amd.builder import StringBuilder

    cases, z = _vectors()
    got = build(StringBuilder, cases[k])
    names = [n[len(f"c{k}_"):] for n in z.files if n.startswith(f"c{k}_")]
    assert sorted(names) == sorted(got)
    for n in names:
        ref = z[f"c{k}_{n}"]
        assert got[n].shape == ref.shape, n
        if ref.dtype.kind in "iu":
            np.testing.assert_array_equal(got[n], ref, err_msg=n)
        else:
            np.testing.assert_allclose(got[n], ref, rtol=1e-14, atol=1e-15, err_msg=n)


def test_builder_closed_forms():
    """Without noise: |psi| = psi_abs prod |tanh(r_k^2)|, the phase winds
    by w around each core, fields vanish outside the sc domain, and the
    symmetric gauge has curl A = B_z."""
    from mepgl_amd.builder import StringBuilder

    b = StringBuilder(2, Nx=21, Ny=21, x_lim=[-2, 2], y_lim=[-2, 2], multicomponent=True,
                      B_z=0.9)
    b.set_sc_domain(lambda x, y: np.hypot(x, y) < 1.8)
    b.add_vortex([0.4, 0.4], [0.0, 0.0], [1, 2], 1)
    u1, v1, u2, v2 = b.generate_matter_fields(psi_abs_1=0.7, psi_abs_2=0.5)
    comp, sc = b.get_domain_matrices()
    r2 = (b.x - 0.4) ** 2 + b.y ** 2
    np.testing.assert_allclose(np.hypot(u1[0], v1[0]), 0.7 * np.abs(np.tanh(r2)) * sc,
                               rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(np.hypot(u2, v2), np.broadcast_to(0.5 * sc, u2.shape), rtol=1e-12)
    assert np.all(u1[:, sc == 0] == 0) and np.all(v2[:, sc == 0] == 0)
    # Winding of image 1 (w = 2) around a small loop about the core.
    # The core sits on grid point (12, 10); walk the square of half-width 2
    # around it, one grid step at a time.
    loop = ([(i, 8) for i in range(10, 14)] + [(14, j) for j in range(8, 12)] +
            [(i, 12) for i in range(14, 10, -1)] + [(10, j) for j in range(12, 7, -1)])
    ii, jj = np.array(loop).T
    th = np.unwrap(np.arctan2(v1[1, ii, jj], u1[1, ii, jj]))
    assert abs((th[-1] - th[0]) - 2 * 2 * np.pi) < 1e-9
    ax, ay = b.generate_vector_potential("symmetric")
    dx = b.x_ax[1] - b.x_ax[0]
    curl = np.gradient(ay[0], dx, axis=0) - np.gradient(ax[0], dx, axis=1)
    np.testing.assert_allclose(curl, 0.9, rtol=1e-12)
    with pytest.raises(Exception):
        b.generate_vector_potential("coulomb")
