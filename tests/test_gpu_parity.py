"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (FP64 path, identical inputs):
  per-image energy, A..E, s, s_jb          relative 1e-12
  gradient, JB, direction                  per-image norm-wise relative 1e-12
  line-search step alpha                   bit-identical when the GPU's
                                           coefficients are fed to the oracle's
                                           scan, and equal to the oracle's
  fields after steps / reparametrisation   per-image norm-wise relative 1e-11
Mixed path: energies within 1e-5 relative of the FP64 oracle on the same
inputs (float storage of the fields), see test_gpu_mixed.py.
"""
import sys

import numpy as np
import pytest

from conftest import bondless_string, normwise, rel_err, small_string

pytestmark = pytest.mark.gpu

CASES = {
    "one_comp": dict(multicomponent=False),
    "one_comp_masked_planes": dict(multicomponent=False, masked=True, planes=True),
    "two_comp": dict(multicomponent=True),
    "two_comp_masked": dict(multicomponent=True, masked=True),
    "two_comp_thin_film": dict(multicomponent=True, masked=True, no_self_field=True),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def solver_for(strings, **kw):
    from mepgl_amd.solver import Solver

    return Solver(strings, **kw)


@pytest.mark.parametrize("case", list(CASES))
def test_energy_gradient_jb(case, oracle):
    s = small_string(N=24, F=4, seed=3, **CASES[case])
    say(f"[gpu] energy/gradient/jb case={case} N={s.N} F={s.F}")
    with solver_for([s]) as sv:
        e_gpu = sv.computeFreeEnergy()[0]
        ga, g1, g2 = sv.gradient(0)
        b, j1, j2 = sv.computeJB(0)
    e_ref = oracle.free_energy(s)
    assert rel_err(e_gpu, e_ref) < 1e-12, (e_gpu, e_ref)
    ra, r1, r2 = oracle.gradient(s)
    if not s.no_self_field:
        assert normwise(ga, ra) < 1e-12
    assert normwise(g1, r1) < 1e-12
    if s.multicomponent:
        assert normwise(g2, r2) < 1e-12
    rb, rj1, rj2 = oracle.compute_jb(s)
    assert normwise(b, rb) < 1e-12
    assert normwise(j1, rj1) < 1e-12
    if s.multicomponent:
        assert normwise(j2, rj2) < 1e-12


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("exact", [False, True])
def test_one_nlcg_step(case, exact, oracle):
    s = small_string(N=20, F=4, seed=5, **CASES[case])
    say(f"[gpu] one NLCG step case={case} exact={exact}")
    ref = oracle.OracleString(s, poly_quirks=not exact)
    ref.nlcg_steps(1)
    with solver_for([s], poly_exact=exact) as sv:
        sv.nlcgSteps(1)
        fen = sv.scalars(0, "fenergy")
        poly = sv.scalars(0, "poly")
        alpha = sv.scalars(0, "alpha")
        gn = sv.scalars(0, "g_norm2_old")  # step 0 writes the caller's g_norm2_old
        a, p1, p2 = sv.download(0)
        da, d1, d2 = sv.download_state(0, "d")
    assert rel_err(fen, ref.fenergy) < 1e-12
    for q in range(4):
        assert rel_err(poly[:, q], ref.poly[:, q]) < 1e-11, q
    assert rel_err(gn, ref.g_norm2[1]) < 1e-12
    # The scan on the GPU's own coefficients is the oracle's scan bit for bit.
    for n in range(s.F):
        assert alpha[n] == oracle.line_search(fen[n], *poly[n]), n
    np.testing.assert_array_equal(alpha, ref.alpha)
    assert normwise(d1, ref.d_p1) < 1e-12
    if not s.no_self_field:
        assert normwise(da, ref.d_a) < 1e-12
    assert normwise(p1, ref.psi_1) < 1e-12
    assert normwise(a, ref.a) < 1e-12
    if s.multicomponent:
        assert normwise(p2, ref.psi_2) < 1e-12


@pytest.mark.parametrize("case", ["one_comp", "two_comp_masked"])
@pytest.mark.parametrize("steps", [6, 5])
def test_nlcg_steps_call_boundaries(case, steps, oracle):
    """Two calls in a row, with even and odd step counts (the reference's
    pointer-swap pairing at call boundaries)."""
    s = small_string(N=20, F=4, seed=7, **CASES[case])
    say(f"[gpu] nlcg steps case={case} steps={steps}x2")
    ref = oracle.OracleString(s)
    ref.nlcg_steps(steps)
    ref.nlcg_steps(steps)
    with solver_for([s]) as sv:
        sv.nlcgSteps(steps)
        sv.nlcgSteps(steps)
        a, p1, p2 = sv.download(0)
        gn0 = sv.scalars(0, "g_norm2")
        gn1 = sv.scalars(0, "g_norm2_old")
        gdg = sv.scalars(0, "g_delta_g")
        beta = sv.scalars(0, "beta")
        g = sv.download_state(0, "g")
    assert normwise(p1, ref.psi_1) < 1e-11
    assert normwise(a, ref.a) < 1e-11
    assert rel_err(gn0, ref.g_norm2[0]) < 1e-10
    assert rel_err(gn1, ref.g_norm2[1]) < 1e-10
    assert normwise(g[1], ref.g_p1[0]) < 1e-10
    # The last step's g.(g - g_old) and Polak-Ribiere beta: step k of a call
    # stores |g|^2 in slot 1 when k is even, and beta divides by the other.
    assert rel_err(gdg, ref.g_delta_g) < 1e-9
    cur = 1 if (steps - 1) % 2 == 0 else 0
    want = np.maximum(ref.g_delta_g / ref.g_norm2[1 - cur], 0.0)
    assert np.any(want > 0)
    assert rel_err(beta, want) < 1e-9


def test_steepest_descent_steps(oracle):
    """nlcg = 0 (the stdin `C` toggle): beta is 0 on every step, so each
    direction is -g."""
    s = small_string(N=20, F=4, seed=9, multicomponent=False)
    say("[gpu] steepest descent steps")
    ref = oracle.OracleString(s)
    ref.nlcg_steps(3, nlcg=False)
    with solver_for([s]) as sv:
        sv.nlcgSteps(3, nlcg=False)
        a, p1, _ = sv.download(0)
        beta = sv.scalars(0, "beta")
        gdg = sv.scalars(0, "g_delta_g")
    np.testing.assert_array_equal(beta, np.zeros(s.F))
    assert rel_err(gdg, ref.g_delta_g) < 1e-9
    assert normwise(p1, ref.psi_1) < 1e-11
    assert normwise(a, ref.a) < 1e-11


def test_maxwell_mode(oracle):
    s = small_string(N=20, F=4, seed=8, multicomponent=False)
    say("[gpu] maxwell mode")
    ref = oracle.OracleString(s)
    ref.nlcg_steps(4)
    ref.nlcg_steps(4, maxwell_solver=True)
    with solver_for([s]) as sv:
        sv.nlcgSteps(4)
        sv.nlcgSteps(4, maxwell_solver=True)
        a, p1, _ = sv.download(0)
    assert normwise(a, ref.a) < 1e-11
    assert normwise(p1, ref.psi_1) < 1e-11


@pytest.mark.parametrize("case", ["one_comp", "one_comp_masked_planes", "two_comp",
                                  "two_comp_masked", "two_comp_thin_film"])
def test_arclength_and_linear_redistribution(case, oracle):
    s = small_string(N=20, F=6, seed=9, **CASES[case])
    say(f"[gpu] arclength + linear case={case}")
    ref = oracle.OracleString(s)
    s_ref, sjb_ref = ref.arc_lengths()
    ref.redistribute(cubic=False)
    with solver_for([s]) as sv:
        (s_gpu, sjb_gpu), = sv.arcLength()
        sv.redistributeLinear()
        a, p1, p2 = sv.download(0)
    assert rel_err(s_gpu, s_ref) < 1e-12
    assert rel_err(sjb_gpu, sjb_ref) < 1e-12
    assert normwise(a, ref.a) < 1e-12
    assert normwise(p1, ref.psi_1) < 1e-12
    if s.multicomponent:
        assert normwise(p2, ref.psi_2) < 1e-12


@pytest.mark.parametrize("rows", [4, 5])
def test_arclength_multi_strip(rows, oracle):
    """Two strips (the second 6 columns wide) and short segments: the halo
    columns and rows of the arclength pass's row window at every seam."""
    s = small_string(N=70, F=5, seed=49, multicomponent=True, masked=True)
    say(f"[gpu] arclength N=70 rows={rows}")
    ref = oracle.OracleString(s)
    s_ref, sjb_ref = ref.arc_lengths()
    with solver_for([s]) as sv:
        sv.set_segment_rows(0, rows)
        (s_gpu, sjb_gpu), = sv.arcLength()
    assert rel_err(s_gpu, s_ref) < 1e-12
    assert rel_err(sjb_gpu, sjb_ref) < 1e-12


@pytest.mark.parametrize("nsf", [False, True])
def test_tiles_without_bonds(nsf, oracle):
    """The step kernels skip the stencils of tiles without bonds and only
    carry the fields over; the loop must still match the oracle, arclength
    over every cell and reparametrisation included."""
    s = bondless_string(no_self_field=nsf)
    N, rows = s.N, 7
    # The geometry does leave tiles (64 columns x 7 rows) whose neighbour
    # maps are empty in both domains.
    nn = (oracle.build_nn_map(s.comp_domain) != 0).T
    empty = [not nn[j0:j0 + rows, i0:i0 + 64].any()
             for j0 in range(0, N, rows) for i0 in range(0, N, 64)]
    assert sum(empty) >= 6
    say(f"[gpu] tiles without bonds N={N} rows={rows} empty={sum(empty)} nsf={nsf}")
    ref = oracle.OracleString(s)
    ref.loop(2, 3, "L")
    with solver_for([s]) as sv:
        sv.set_segment_rows(0, rows)
        sv.loop(2, 3, "L")
        a, p1, p2 = sv.download(0)
        fen = sv.scalars(0, "fenergy")
    assert normwise(a, ref.a) < 1e-10
    assert normwise(p1, ref.psi_1) < 1e-10
    assert normwise(p2, ref.psi_2) < 1e-10
    assert rel_err(fen, ref.fenergy) < 1e-12


@pytest.mark.parametrize("mode,precision", [("L", "fp64"), ("C", "fp64"), ("L", "mixed")])
def test_shared_fixed_a(mode, precision, oracle):
    """A thin-film string whose fixed A is the same in every image keeps one
    A for all images: the step kernels read image 0's A planes and the
    reparametrisation does not interpolate it, so A stays exactly the input.
    The reference interpolates the (identical) A as well, which moves it by
    at most a few ulps; fields and energies match the oracle to 1e-10 after
    three outer iterations (FP64), to 1e-5 in mixed precision."""
    s = small_string(N=24, F=6, seed=94, multicomponent=True, masked=True, no_self_field=True)
    s.a[:] = s.a[0]
    say(f"[gpu] shared fixed A, mode {mode}, {precision}")
    ref = oracle.OracleString(s)
    ref.loop(3, 4, mode)
    with solver_for([s], precision=precision) as sv:
        sv.loop(3, 4, mode)
        a, p1, p2 = sv.download(0)
        e = sv.computeFreeEnergy()[0]
    a_in = s.a if precision == "fp64" else s.a.astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(a, a_in)
    tol = 1e-10 if precision == "fp64" else 1e-5
    if precision == "fp64":
        assert normwise(a, ref.a) < 1e-14
    assert normwise(p1, ref.psi_1) < tol
    assert normwise(p2, ref.psi_2) < tol
    assert rel_err(e, ref.free_energy()) < tol


@pytest.mark.parametrize("textbook", [False, True])
def test_cubic_redistribution(textbook, oracle):
    s = small_string(N=20, F=7, seed=10, multicomponent=True)
    say(f"[gpu] cubic textbook={textbook}")
    ref = oracle.OracleString(s, cubic_textbook=textbook)
    ref.arc_lengths()
    ref.redistribute(cubic=True)
    with solver_for([s], cubic_textbook=textbook) as sv:
        sv.arcLength()
        sv.redistributeCubic()
        a, p1, p2 = sv.download(0)
    assert normwise(a, ref.a) < 1e-11
    assert normwise(p1, ref.psi_1) < 1e-11
    assert normwise(p2, ref.psi_2) < 1e-11


@pytest.mark.parametrize("mode", ["L", "C", "S"])
def test_loop_graph_and_eager(mode, oracle):
    s = small_string(N=20, F=6, seed=11, multicomponent=False)
    say(f"[gpu] loop mode={mode} graph and eager")
    ref = oracle.OracleString(s)
    ref.loop(3, 4, mode)
    out = []
    for use_graph in (True, False):
        with solver_for([s], use_graph=use_graph) as sv:
            sv.loop(3, 4, mode)
            out.append(sv.download(0))
            fen = sv.scalars(0, "fenergy")
    for k in range(2):
        np.testing.assert_array_equal(out[0][k], out[1][k])
    assert normwise(out[0][1], ref.psi_1) < 1e-10
    assert normwise(out[0][0], ref.a) < 1e-10
    assert rel_err(fen, ref.fenergy) < 1e-12


def test_batched_strings_match_single(oracle):
    strings = [small_string(N=20, F=4, seed=12, multicomponent=False),
               small_string(N=28, F=6, seed=13, multicomponent=True, masked=True),
               small_string(N=70, F=3, seed=14, multicomponent=False)]
    say("[gpu] batched 3 strings")
    with solver_for(strings) as sv:
        sv.loop(2, 4, "L")
        got = [sv.download(i) for i in range(3)]
        en = sv.computeFreeEnergy()
    for i, s in enumerate(strings):
        ref = oracle.OracleString(s)
        ref.loop(2, 4, "L")
        assert normwise(got[i][1], ref.psi_1) < 1e-10, i
        e_ref = ref.free_energy()
        assert rel_err(en[i], e_ref) < 1e-11, i


def test_resample(oracle):
    s = small_string(N=20, F=5, seed=15, multicomponent=True)
    say("[gpu] resample to 8 images")
    ref = oracle.OracleString(s)
    ref.arc_lengths()
    ref.redistribute(F_new=8)
    with solver_for([s]) as sv:
        sv.resample(0, 8)
        a, p1, p2 = sv.download(0)
    assert a.shape[0] == 8
    assert normwise(p1, ref.psi_1) < 1e-12
    assert normwise(p2, ref.psi_2) < 1e-12
    assert normwise(a, ref.a) < 1e-12


@pytest.mark.parametrize("mc", [False, True])
def test_gauge_alignment(mc, oracle):
    s = small_string(N=20, F=6, seed=44, multicomponent=mc)
    say(f"[gpu] gauge alignment mc={mc}")
    ref = oracle.OracleString(s)
    ref.nlcg_steps(2)
    th_ref = ref.gauge_align()
    ref.nlcg_steps(2)
    with solver_for([s]) as sv:
        sv.nlcgSteps(2)
        th, = sv.gauge_align()
        sv.nlcgSteps(2)
        a, p1, p2 = sv.download(0)
    np.testing.assert_allclose(th, th_ref, rtol=1e-11, atol=1e-13)
    assert normwise(p1, ref.psi_1) < 1e-10
    assert normwise(a, ref.a) < 1e-10
    if mc:
        assert normwise(p2, ref.psi_2) < 1e-10


def test_loop_with_alignment(oracle):
    s = small_string(N=20, F=6, seed=45, multicomponent=True)
    say("[gpu] loop with alignment")
    ref = oracle.OracleString(s)
    ref.loop(2, 4, "L", align=True)
    with solver_for([s], gauge_align=True) as sv:
        sv.loop(2, 4, "L")
        a, p1, p2 = sv.download(0)
    assert normwise(p1, ref.psi_1) < 1e-10
    assert normwise(p2, ref.psi_2) < 1e-10


@pytest.mark.parametrize("case", ["one_comp", "two_comp_masked", "two_strings"])
def test_local_gauge_alignment(case, oracle):
    """mepgl_gauge_align_local (rhs, four DCT products, prefix, apply) against
    the oracle's restatement, after two NLCG steps so that the direction and
    gradients carry psi parts that must rotate; FP64, 1e-11.  Two strings of
    different sizes share the context in the last case."""
    if case == "two_strings":
        strings = [small_string(N=20, F=5, seed=95, multicomponent=True),
                   small_string(N=70, F=4, seed=96)]
    else:
        strings = [small_string(N=24, F=6, seed=97, **CASES[case])]
    for st in strings:
        st.q_2 = 0.6
    say(f"[gpu] local gauge alignment case={case}")
    refs = [oracle.OracleString(st) for st in strings]
    for r in refs:
        r.nlcg_steps(2)
        r.gauge_align_local()
        r.nlcg_steps(2)
    with solver_for(strings) as sv:
        sv.nlcgSteps(2)
        sv.gauge_align_local()
        sv.nlcgSteps(2)
        got = [sv.download(i) for i in range(len(strings))]
    for (a, p1, p2), r, st in zip(got, refs, strings):
        assert normwise(a, r.a) < 1e-11
        assert normwise(p1, r.psi_1) < 1e-11
        if st.multicomponent:
            assert normwise(p2, r.psi_2) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_loop_with_local_alignment(precision, oracle):
    """Mode L with the local and the constant-phase alignment before every
    arclength, graph replay included, against the oracle's loop."""
    s = small_string(N=20, F=6, seed=98, multicomponent=True)
    say(f"[gpu] loop with local alignment, {precision}")
    ref = oracle.OracleString(s)
    ref.loop(3, 4, "L", align=True, align_local=True)
    with solver_for([s], precision=precision, gauge_align=True,
                    gauge_align_local=True) as sv:
        sv.loop(3, 4, "L")
        a, p1, p2 = sv.download(0)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert normwise(a, ref.a) < tol
    assert normwise(p1, ref.psi_1) < tol
    assert normwise(p2, ref.psi_2) < tol


# N = 40: the last segment has 1 (13 rows), 2 (19), 3 (37), 4 (6), 5 (7)
# or a full height; heights of 3k+1, 3k+2 and 3k+3 exercise every padding
# of the kernels' last rotation.
@pytest.mark.parametrize("rows", [4, 5, 6, 7, 8, 13, 19, 37, 64])
def test_segment_rows_do_not_change_results(rows, oracle):
    s = small_string(N=40, F=4, seed=46, multicomponent=True, masked=True)
    say(f"[gpu] segment rows {rows}")
    ref = oracle.OracleString(s)
    ref.loop(2, 3, "L")
    with solver_for([s]) as sv:
        sv.set_segment_rows(0, rows)
        assert sv.info(0)["tile_rows"] == min(rows, s.N)
        sv.loop(2, 3, "L")
        a, p1, p2 = sv.download(0)
        fen = sv.scalars(0, "fenergy")
    assert normwise(p1, ref.psi_1) < 1e-10
    assert normwise(p2, ref.psi_2) < 1e-10
    assert rel_err(fen, ref.fenergy) < 1e-12


def test_autotune_keeps_the_reparametrisation_plan():
    """The autotuner times outer iterations, which recompute the arclength
    plan; it restores the caller's plan, so a redistribution requested after
    tuning equals one without tuning (redistribution is per element)."""
    s = small_string(N=20, F=5, seed=49, multicomponent=False)
    say("[gpu] autotune keeps the reparametrisation plan")
    with solver_for([s]) as tuned, solver_for([s]) as plain:
        for sv in (tuned, plain):
            sv.nlcgSteps(3)
            sv.arcLength()
        tuned.autotune(steps=2)
        for sv in (tuned, plain):
            sv.redistributeLinear()
        a, b = tuned.download(0), plain.download(0)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_autotune_leaves_state_unchanged(oracle):
    strings = [small_string(N=20, F=4, seed=47, multicomponent=False),
               small_string(N=36, F=5, seed=48, multicomponent=True)]
    say("[gpu] autotune on 2 strings")
    with solver_for(strings) as sv:
        sv.nlcgSteps(3)
        before = [sv.download(i) for i in range(2)] + \
                 [sv.download_state(i, "d") for i in range(2)] + \
                 [sv.scalars(i, "g_norm2") for i in range(2)]
        rows = sv.autotune(steps=2)
        after = [sv.download(i) for i in range(2)] + \
                [sv.download_state(i, "d") for i in range(2)] + \
                [sv.scalars(i, "g_norm2") for i in range(2)]
        for x, y in zip(before, after):
            if isinstance(x, tuple):
                for u, v in zip(x, y):
                    if u is not None:
                        np.testing.assert_array_equal(u, v)
            else:
                np.testing.assert_array_equal(x, y)
        assert [sv.info(i)["tile_rows"] for i in range(2)] == rows
        sv.nlcgSteps(3)
        got = [sv.download(i)[1] for i in range(2)]
    for i, s in enumerate(strings):
        ref = oracle.OracleString(s)
        ref.nlcg_steps(3)
        ref.nlcg_steps(3)
        assert normwise(got[i], ref.psi_1) < 1e-10, i


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_checkpoint_resume_is_exact(tmp_path, precision):
    """Relax, checkpoint, relax more; a fresh context restored from the
    checkpoint relaxes to the same fields bit for bit."""
    strings = [small_string(N=24, F=5, seed=50, multicomponent=True),
               small_string(N=20, F=4, seed=51, multicomponent=False)]
    say(f"[gpu] checkpoint resume {precision}")
    path = str(tmp_path / "ckpt.npz")
    with solver_for(strings, precision=precision) as sv:
        sv.loop(2, 3, "L")
        sv.nlcgSteps(3)
        sv.save_checkpoint(path)
        sv.loop(2, 3, "L")
        sv.nlcgSteps(2)
        want = [sv.download(i) for i in range(2)] + [sv.scalars(i, "g_norm2") for i in range(2)]
    with solver_for(strings, precision=precision) as sv:
        sv.load_checkpoint(path)
        sv.loop(2, 3, "L")
        sv.nlcgSteps(2)
        got = [sv.download(i) for i in range(2)] + [sv.scalars(i, "g_norm2") for i in range(2)]
    for x, y in zip(want, got):
        if isinstance(x, tuple):
            for u, v in zip(x, y):
                np.testing.assert_array_equal(u, v)
        else:
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("N", [65, 129])
def test_one_column_last_strip(N, oracle):
    """N = 64k + 1: the last 64-column strip holds a single column, whose
    x-neighbours come from clamped lanes."""
    s = small_string(N=N, F=3, seed=52, multicomponent=True, masked=True)
    say(f"[gpu] N={N} one-column strip")
    ref = oracle.OracleString(s)
    ref.loop(1, 3, "L")
    with solver_for([s]) as sv:
        e = sv.computeFreeEnergy()[0]
        ga, g1, g2 = sv.gradient(0)
        sv.loop(1, 3, "L")
        a, p1, p2 = sv.download(0)
    assert rel_err(e, oracle.free_energy(s)) < 1e-12
    ra, r1, r2 = oracle.gradient(s)
    assert normwise(g1, r1) < 1e-12 and normwise(g2, r2) < 1e-12 and normwise(ga, ra) < 1e-12
    assert normwise(p1, ref.psi_1) < 1e-10
    assert normwise(p2, ref.psi_2) < 1e-10
    assert normwise(a, ref.a) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_winding_kernel_matches_summary(precision):
    """mepgl_winding against observables.winding (pinned by closed forms)
    on fields with a vortex, an antivortex and no vortex."""
    from mepgl_amd import observables

    s = small_string(N=33, F=3, seed=53, multicomponent=True)
    c = (np.arange(33) - 16.0)
    y, x = np.meshgrid(c, c, indexing="ij")
    z = x + 1j * y + 0.3 + 0.2j
    for n, p in enumerate([z / np.abs(z), np.conj(z) / np.abs(z), np.full((33, 33), 0.8 + 0.1j)]):
        s.psi_1[n, ..., 0], s.psi_1[n, ..., 1] = p.real, p.imag
        q = (z / np.abs(z)) ** 2
        s.psi_2[n, ..., 0], s.psi_2[n, ..., 1] = q.real, q.imag
    say(f"[gpu] winding kernel {precision}")
    with solver_for([s], precision=precision) as sv:
        (w1, w2), = sv.winding()
        _, p1, p2 = sv.download(0)
    # The reference's layout [n][i][j] is the transpose of [n][j][i].
    r1 = observables.winding(np.swapaxes(p1[..., 0], 1, 2), np.swapaxes(p1[..., 1], 1, 2))
    r2 = observables.winding(np.swapaxes(p2[..., 0], 1, 2), np.swapaxes(p2[..., 1], 1, 2))
    np.testing.assert_allclose(w1, r1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(w2, r2, rtol=0, atol=1e-12)
    np.testing.assert_allclose(w1, [2 * np.pi, -2 * np.pi, 0.0], atol=0.05)
    np.testing.assert_allclose(w2, [4 * np.pi] * 3, atol=0.1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("N,rows", [(70, 16), (130, 7)])
def test_folded_line_search_equals_finishing_launch(precision, N, rows):
    """P's line search done by the next G (the default, with at most 64
    tiles per image) and by a finishing launch after every P
    (MEPGL_CONTROL_FOLD_LINE_SEARCH = 0) give the same fields, direction and
    per-image scalars bit for bit, through graph replays of mode L and an
    nlcgSteps call.  130 columns at 7-row segments: 57 tiles per image, the
    last strip two columns wide.  The folded run makes one finishing launch
    per call, the other one per step."""
    from mepgl_amd import _lib

    s = small_string(N=N, F=5, seed=41, multicomponent=True, masked=True)
    say(f"[gpu] folded line search vs finishing launch: {precision} N={N} rows={rows}")
    out = []
    for fold in (1, 0):
        with solver_for([s], precision=precision, segment_rows=rows) as sv:
            sv.set_control(_lib.CONTROL_FOLD_LINE_SEARCH, fold)
            sv.loop(2, 6, "L")
            sv.profile(True)
            sv.nlcgSteps(5)
            finishes = sv.profile_read(4)[1]
            sv.profile(False)
            assert finishes == (1 if fold else 5), (fold, finishes)
            names = ("fenergy", "alpha", "ls_index", "g_norm2", "beta", "poly")
            sc = [sv.scalars(0, k) for k in names]
            out.append((sv.download(0), sv.download_state(0, "d"), sc))
    (f1, d1, s1), (f0, d0, s0) = out
    for x, y in zip(f1 + d1, f0 + d0):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(s1, s0):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reset_makes_the_run_independent_of_history(precision):
    """mepgl_reset_solver makes the current field and direction buffers the
    context's buffer 0 (pointer swaps in the descriptors).  A context that
    ran outer iterations and NLCG steps of both parities before an upload
    and a reset then continues exactly as a fresh context from the same
    upload: fields, direction and scalars bit for bit, through graph
    replays."""
    s = small_string(N=40, F=5, seed=43, multicomponent=True, masked=True)
    say(f"[gpu] reset after history vs fresh context: {precision}")
    out = []
    for history in (False, True):
        with solver_for([s], precision=precision) as sv:
            if history:
                sv.loop(3, 5, "L")
                sv.nlcgSteps(3)
                sv.upload(0, s)
                sv.reset_solver()
            sv.loop(3, 5, "L")
            out.append((sv.download(0), sv.download_state(0, "d"),
                        [sv.scalars(0, k) for k in ("fenergy", "alpha", "s_jb")]))
    (f1, d1, s1), (f0, d0, s0) = out
    for x, y in zip(f1 + d1, f0 + d0):
        if x is not None:
            np.testing.assert_array_equal(x, y)
    for x, y in zip(s1, s0):
        np.testing.assert_array_equal(x, y)


def test_sites_without_bonds_across_height_changes(oracle):
    """G with the pending update skips tiles without bonds, relying on both
    field buffers holding the same values at every site without bonds (G
    without the update copies them after a reparametrisation).  Changing the
    segment height between calls moves sites between tiles with and without
    bonds; fields everywhere, outside the domain included (where the
    redistribution moves the non-zero initial values), must still follow the
    oracle through linear and cubic reparametrisations and plain steps."""
    s = bondless_string()
    say("[gpu] sites without bonds across segment-height changes (fp64)")
    ref = oracle.OracleString(s)
    ref.loop(1, 3, "L")
    ref.nlcg_steps(2)
    ref.loop(1, 3, "C")
    ref.nlcg_steps(3)
    with solver_for([s]) as sv:
        sv.set_segment_rows(0, 7)
        sv.loop(1, 3, "L")
        sv.set_segment_rows(0, 16)
        sv.nlcgSteps(2)
        sv.set_segment_rows(0, 4)
        sv.loop(1, 3, "C")
        sv.set_segment_rows(0, 7)
        sv.nlcgSteps(3)
        a, p1, p2 = sv.download(0)
    assert normwise(a, ref.a) < 1e-10
    assert normwise(p1, ref.psi_1) < 1e-10
    assert normwise(p2, ref.psi_2) < 1e-10
    # Sites without bonds (both domains are the same disk), [i][j] as the
    # fields' file layout [n][i][j].
    bondless = oracle.build_nn_map(s.comp_domain) == 0
    assert bondless.any()
    for got, want in ((p1, ref.psi_1), (p2, ref.psi_2)):
        g, w = got[:, bondless], want[:, bondless]
        assert np.max(np.abs(g - w)) <= 1e-10 * np.max(np.abs(w))


@pytest.mark.parametrize("fold", [1, 0])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("parts", [2, 4])
def test_split_images_equals_one_sequence(precision, use_graph, fold, parts):
    """The NLCG steps of an outer iteration launched as two or four parts of
    each class's images on as many streams (MEPGL_CONTROL_SPLIT_IMAGES = 2 or
    4) and as one launch sequence (0) give the same fields, direction
    and per-image scalars bit for bit: modes L, C and S, graph replay and
    eager launches, three strings in two classes (one class of 5 + 4 images,
    one of 5, both split unevenly), with the line search in G or in a
    finishing launch per part after every P."""
    from mepgl_amd import _lib

    strings = [small_string(N=40, F=5, seed=51),
               small_string(N=40, F=4, seed=52),
               small_string(N=24, F=5, seed=53, multicomponent=True, masked=True)]
    say(f"[gpu] split images vs one sequence: {precision} graph={use_graph} fold={fold} "
        f"parts={parts}")
    names = ("fenergy", "alpha", "ls_index", "g_norm2", "beta", "poly")
    out = []
    for split in (parts, 0):
        with solver_for(strings, precision=precision, use_graph=use_graph) as sv:
            sv.set_control(_lib.CONTROL_SPLIT_IMAGES, split)
            sv.set_control(_lib.CONTROL_FOLD_LINE_SEARCH, fold)
            sv.loop(2, 6, "L")
            sv.loop(1, 5, "C")
            sv.loop(1, 4, "S")
            out.append([(sv.download(i), sv.download_state(i, "d"),
                         [sv.scalars(i, k) for k in names]) for i in range(3)])
    for (f1, d1, s1), (f0, d0, s0) in zip(*out):
        for x, y in zip(f1 + d1, f0 + d0):
            np.testing.assert_array_equal(x, y)
        for x, y in zip(s1, s0):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_split_images_with_empty_parts(precision):
    """With four parts, classes of 2 and 3 images leave some parts empty (no
    step launches, no finishing launch).  Split and one launch sequence
    still give the same fields, direction and per-image scalars bit for bit,
    in modes L and S, by graph replay."""
    from mepgl_amd import _lib

    strings = [small_string(N=40, F=2, seed=61),
               small_string(N=24, F=3, seed=62, multicomponent=True)]
    say(f"[gpu] split images with empty parts: {precision}")
    names = ("fenergy", "alpha", "ls_index", "g_norm2", "beta", "poly")
    out = []
    for split in (4, 0):
        with solver_for(strings, precision=precision, use_graph=True) as sv:
            sv.set_control(_lib.CONTROL_SPLIT_IMAGES, split)
            sv.loop(2, 5, "L")
            sv.loop(1, 4, "S")
            out.append([(sv.download(i), sv.download_state(i, "d"),
                         [sv.scalars(i, k) for k in names]) for i in range(2)])
    for (f1, d1, s1), (f0, d0, s0) in zip(*out):
        for x, y in zip(f1 + d1, f0 + d0):
            np.testing.assert_array_equal(x, y)
        for x, y in zip(s1, s0):
            np.testing.assert_array_equal(x, y)
