"""The MIXED precision policy against the FP64 oracle.

Fields, directions and gradients are stored and updated in float; every
reduction, NLCG scalar, line search, arclength and reparametrisation is in
double.  Tolerances (relative to the FP64 oracle on the same inputs):
  energy, single evaluation                1e-5
  gradient, per-image norm-wise            1e-4
  energies after a short relaxation loop   1e-5
  energy barrier max(F - F_0)              1e-3 of the barrier
"""
import sys

import numpy as np
import pytest

from conftest import bondless_string, normwise, rel_err, small_string

pytestmark = pytest.mark.gpu

CASES = {
    "one_comp": dict(multicomponent=False),
    "one_comp_planes": dict(multicomponent=False, masked=True, planes=True),
    "two_comp_masked": dict(multicomponent=True, masked=True),
    "two_comp_thin_film": dict(multicomponent=True, masked=True, no_self_field=True),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rounded(s):
    """The string with its fields as the float store holds them."""
    t = s.copy()
    t.a = t.a.astype(np.float32).astype(np.float64)
    t.psi_1 = t.psi_1.astype(np.float32).astype(np.float64)
    if t.psi_2 is not None:
        t.psi_2 = t.psi_2.astype(np.float32).astype(np.float64)
    return t


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_energy_and_gradient(case, oracle):
    from mepgl_amd.solver import Solver

    s = small_string(N=24, F=4, seed=31, **CASES[case])
    say(f"[gpu] mixed energy/gradient case={case}")
    with Solver([s], precision="mixed") as sv:
        e = sv.computeFreeEnergy()[0]
        ga, g1, g2 = sv.gradient(0)
    t = rounded(s)
    assert rel_err(e, oracle.free_energy(t)) < 1e-5
    ra, r1, r2 = oracle.gradient(t)
    assert normwise(g1, r1) < 1e-4
    if not s.no_self_field:
        assert normwise(ga, ra) < 1e-4
    if s.multicomponent:
        assert normwise(g2, r2) < 1e-4


@pytest.mark.parametrize("case", ["one_comp", "two_comp_masked"])
def test_mixed_loop_energies_and_barrier(case, oracle):
    from mepgl_amd import configs, observables
    from mepgl_amd.solver import Solver

    if case == "one_comp":
        s = configs.vortex_entry(32, 0.25, 8, 0.9, seed=2)
    else:
        s = small_string(N=28, F=8, seed=32, **CASES[case])
    say(f"[gpu] mixed loop case={case}")
    with Solver([s], precision="mixed") as sv:
        sv.loop(4, 5, "L")
        e = sv.computeFreeEnergy()[0]
    ref = oracle.OracleString(s)
    ref.loop(4, 5, "L")
    e_ref = ref.free_energy()
    assert rel_err(e, e_ref) < 1e-5
    b, _ = observables.barriers(e)
    b_ref, _ = observables.barriers(e_ref)
    assert abs(b - b_ref) <= 1e-3 * max(abs(b_ref), 1e-12) + 1e-5 * np.max(np.abs(e_ref))


def test_mixed_tiles_without_bonds(oracle):
    """The mixed path with tiles whose stencils are skipped (see
    test_gpu_parity.test_tiles_without_bonds) against the FP64 oracle."""
    from mepgl_amd.solver import Solver

    s = bondless_string()
    say("[gpu] mixed tiles without bonds")
    with Solver([s], precision="mixed") as sv:
        sv.set_segment_rows(0, 7)
        sv.loop(2, 3, "L")
        e = sv.computeFreeEnergy()[0]
    ref = oracle.OracleString(s)
    ref.loop(2, 3, "L")
    assert rel_err(e, ref.free_energy()) < 1e-5


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_arclength_applies_pending_update(case):
    """In mixed precision the arclength pass applies a pending update itself
    and writes the new fields to the other buffer.  The fields, s and s_jb
    are bit-identical to those of the update applied first (by the download,
    which flushes it) and the arclength taken after.  N = 70 gives two
    strips and, at 10 rows, segments of both march directions."""
    from mepgl_amd.solver import Solver

    s = small_string(N=70, F=5, seed=37, **CASES[case])
    say(f"[gpu] mixed arclength with a pending update case={case}")
    with Solver([s], precision="mixed", segment_rows=10) as fused, \
            Solver([s], precision="mixed", segment_rows=10) as first:
        fused.nlcgSteps(3)
        first.nlcgSteps(3)
        want_fields = first.download(0)
        (s_f, sjb_f), = fused.arcLength()
        (s_w, sjb_w), = first.arcLength()
        got_fields = fused.download(0)
        # The next steps continue from the same state.
        fused.nlcgSteps(2)
        first.nlcgSteps(2)
        got_after = fused.download(0)
        want_after = first.download(0)
    nf = 3 if s.multicomponent else 2  # (a, psi_1[, psi_2])
    for k in range(nf):
        np.testing.assert_array_equal(got_fields[k], want_fields[k])
        np.testing.assert_array_equal(got_after[k], want_after[k])
    np.testing.assert_array_equal(s_f, s_w)
    np.testing.assert_array_equal(sjb_f, sjb_w)
    assert np.all(np.diff(s_f) > 0)


@pytest.mark.parametrize("mode", ["L", "C"])
def test_mixed_loop_graph_and_eager(mode, oracle):
    """Mixed-precision outer iterations (the arclength pass applies the
    pending update) replayed from a captured graph equal eager launches bit
    for bit, and agree with the FP64 oracle within the mixed tolerance."""
    from mepgl_amd.solver import Solver

    s = small_string(N=20, F=6, seed=11, multicomponent=False)
    say(f"[gpu] mixed loop mode={mode} graph and eager")
    out, fen = [], []
    for use_graph in (True, False):
        with Solver([s], precision="mixed", use_graph=use_graph) as sv:
            sv.loop(3, 4, mode)
            out.append(sv.download(0))
            fen.append(sv.scalars(0, "fenergy"))
    for k in range(2):
        np.testing.assert_array_equal(out[0][k], out[1][k])
    np.testing.assert_array_equal(fen[0], fen[1])
    ref = oracle.OracleString(s)
    ref.loop(3, 4, mode)
    assert rel_err(fen[0], ref.fenergy) < 1e-5
    assert normwise(out[0][1], ref.psi_1) < 1e-4


def test_solver_fp64_check(oracle):
    """Solver.fp64_check relaxes the mixed state further in FP64: its FP64
    energies are the oracle's after the same iterations from the downloaded
    state (1e-11), its mixed energies are the solver's own, and its FP64
    barriers are those of the oracle's energies."""
    from mepgl_amd import observables
    from mepgl_amd.solver import Solver

    s = small_string(N=20, F=5, seed=26, multicomponent=True, masked=True)
    say("[gpu] Solver.fp64_check")
    with Solver([s], precision="mixed") as sv:
        sv.loop(2, 3, "L")
        r = sv.fp64_check(iterations=2, relaxation_steps=3, mode="L")[0]
        a, p1, p2 = sv.download(0)
        e_mixed = sv.computeFreeEnergy()[0]
    np.testing.assert_array_equal(r["fenergy_mixed"], e_mixed)
    t = s.copy()
    t.a, t.psi_1, t.psi_2 = a, p1, p2
    ref = oracle.OracleString(t)
    ref.loop(2, 3, "L")
    e64 = ref.free_energy()
    assert rel_err(r["fenergy_fp64"], e64) < 1e-11
    np.testing.assert_allclose(r["barriers_fp64"], observables.barriers(e64), rtol=1e-9,
                               atol=1e-11 * np.max(np.abs(e64)))
    # This string is still relaxing: the two further FP64 iterations lower
    # every image's energy, which the statement counts as an upper bound.
    assert r["fenergy_rel_err"].shape == (5,) and np.all(np.isfinite(r["fenergy_rel_err"]))
    assert np.all(r["fenergy_fp64"] < r["fenergy_mixed"])


@pytest.mark.parametrize("N,multicomponent,masked", [(256, False, True), (250, True, False),
                                                     (200, False, False), (320, True, True)])
def test_mixed_edge_exchange_against_oracle(N, multicomponent, masked, oracle):
    """Mixed-precision G and P trade the edge columns of the four strips of a
    workgroup through LDS when the four are one segment of one image.  Rows
    of 4 strips (N = 256, 250: the last strip partial, its lane 63 taking its
    own clamped column), 4 strips of which the last holds 8 columns (N = 200)
    and 5 strips (N = 320: groups that straddle two segments load their
    halos), with and without a domain mask: after two outer iterations of
    mode L the fields and energies agree with the FP64 oracle within the mixed
    tolerance (energies 1e-5, fields norm-wise 1e-4).  A wrong halo value
    would put O(1) errors on the strip edges."""
    from mepgl_amd.solver import Solver

    s = small_string(N=N, F=4, seed=61 + N, multicomponent=multicomponent, masked=masked)
    say(f"[gpu] mixed edge exchange N={N} mc={multicomponent} masked={masked}")
    with Solver([s], precision="mixed", segment_rows=16) as sv:
        sv.loop(2, 3, "L")
        e = sv.computeFreeEnergy()[0]
        a, p1, p2 = sv.download(0)
    ref = oracle.OracleString(s)
    ref.loop(2, 3, "L")
    assert rel_err(e, ref.free_energy()) < 1e-5
    assert normwise(p1, ref.psi_1) < 1e-4
    assert normwise(a, ref.a) < 1e-4
    if multicomponent:
        assert normwise(p2, ref.psi_2) < 1e-4


@pytest.mark.parametrize("N,multicomponent,masked", [(256, False, True), (250, True, False),
                                                     (320, True, True)])
def test_mixed_edge_exchange_equals_halo_loads(N, multicomponent, masked):
    """The edge exchange (MEPGL_CONTROL_EDGE_EXCHANGE = 1, the default) and
    halo loads from memory (0) give the same fields, direction and per-image
    scalars bit for bit through graph replays of modes L and C: the traded
    values are the neighbours' own, formed by the same operations from the
    same inputs."""
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    s = small_string(N=N, F=5, seed=71 + N, multicomponent=multicomponent, masked=masked)
    say(f"[gpu] edge exchange vs halo loads N={N} mc={multicomponent} masked={masked}")
    names = ("fenergy", "alpha", "ls_index", "g_norm2", "beta", "poly")
    out = []
    for edges in (1, 0):
        with Solver([s], precision="mixed", segment_rows=16) as sv:
            sv.set_control(_lib.CONTROL_EDGE_EXCHANGE, edges)
            sv.loop(2, 4, "L")
            sv.loop(1, 3, "C")
            out.append((sv.download(0), sv.download_state(0, "d"),
                        [sv.scalars(0, k) for k in names]))
    (f1, d1, s1), (f0, d0, s0) = out
    for x, y in zip(f1 + d1, f0 + d0):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(s1, s0):
        np.testing.assert_array_equal(x, y)
