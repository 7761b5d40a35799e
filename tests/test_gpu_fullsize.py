"""The benchmark workloads at their full sizes, through properties that do
not depend on size:

* vortex_entry_small (64x64, 15 images): the whole 200-step run with a
  reparametrisation every 10 steps against the oracle.
* vortex_entry_256, two_band_512, thin_film_disk_1024: one FP64 NLCG step
  against the oracle on the full grid (two_band_512 and the thin film on
  the images of one string); graph replay equal to eager launches bit for
  bit; mixed energies within 1e-5 of FP64.
* sweep_48: the batched context equals one context per string bit for bit
  (same schedule), and a tuned schedule stays within 1e-12 of the default.

Tolerances: one step 1e-12 (energies, per-image norm-wise fields); 200
steps 1e-9; mixed vs FP64 1e-5 relative per image.
"""
import sys

import numpy as np
import pytest

from conftest import normwise, rel_err

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def solver(strings, **kw):
    from mepgl_amd.solver import Solver

    return Solver(strings, **kw)


_SWEEP = []


def sweep():
    """The 48-string sweep, generated once per session (about 20 s)."""
    from mepgl_amd import configs

    if not _SWEEP:
        _SWEEP.append(configs.sweep_48())
    return _SWEEP[0]


def test_vortex_entry_small_full_run(oracle):
    from mepgl_amd import configs, observables

    s = configs.vortex_entry_small()
    say("[gpu] vortex_entry_small: 20 x 10 steps, mode L, vs oracle")
    ref = oracle.OracleString(s)
    ref.loop(20, 10, "L")
    e_ref = ref.free_energy()
    with solver([s]) as sv:
        sv.loop(20, 10, "L")
        a, p1, _ = sv.download(0)
        e = sv.computeFreeEnergy()[0]
    assert rel_err(e, e_ref) < 1e-9
    assert normwise(p1, ref.psi_1) < 1e-9
    assert normwise(a, ref.a) < 1e-9
    b, b_ref = observables.barriers(e), observables.barriers(e_ref)
    scale = np.max(np.abs(e_ref))
    assert abs(b[0] - b_ref[0]) <= 1e-9 * scale and abs(b[1] - b_ref[1]) <= 1e-9 * scale


@pytest.mark.parametrize("name", ["vortex_entry_256", "two_band_512", "thin_film_disk_1024"])
def test_one_step_full_grid(name, oracle):
    from mepgl_amd import configs

    s = configs.thin_film_disk(0) if name == "thin_film_disk_1024" else configs.WORKLOADS[name]()[0]
    say(f"[gpu] {name}: one FP64 NLCG step, N={s.N} F={s.F}")
    ref = oracle.OracleString(s)
    ref.nlcg_steps(1)
    with solver([s]) as sv:
        sv.nlcgSteps(1)
        fen = sv.scalars(0, "fenergy")
        alpha = sv.scalars(0, "alpha")
        a, p1, p2 = sv.download(0)
    assert rel_err(fen, ref.fenergy) < 1e-12
    np.testing.assert_array_equal(alpha, ref.alpha)
    assert normwise(p1, ref.psi_1) < 1e-12
    if s.multicomponent:
        assert normwise(p2, ref.psi_2) < 1e-12
    if not s.no_self_field:
        assert normwise(a, ref.a) < 1e-12


@pytest.mark.parametrize("name", ["vortex_entry_256", "two_band_512"])
def test_graph_equals_eager_and_mixed_close(name):
    from mepgl_amd import configs

    s = configs.WORKLOADS[name]()[0]
    say(f"[gpu] {name}: graph vs eager, mixed vs fp64")
    out = []
    for use_graph in (True, False):
        with solver([s], use_graph=use_graph, segment_rows=8) as sv:
            sv.loop(2, 10, "L")
            out.append(sv.download(0))
    for k in range(2):
        np.testing.assert_array_equal(out[0][k], out[1][k])
    with solver([s], precision="fp64") as sv:
        sv.loop(3, 10, "L")
        e64 = sv.computeFreeEnergy()[0]
    with solver([s], precision="mixed") as sv:
        sv.loop(3, 10, "L")
        e32 = sv.computeFreeEnergy()[0]
    assert rel_err(e32, e64) < 1e-5


def test_sweep_batched_equals_separate():
    strings = sweep()
    say(f"[gpu] sweep_48: batched vs one context per string ({len(strings)} strings)")
    rows = 16
    with solver(strings, segment_rows=rows) as sv:
        sv.loop(2, 5, "L")
        batched = [sv.download(i) for i in range(len(strings))]
    for i in (0, 7, 23, 47):
        with solver([strings[i]], segment_rows=rows) as sv:
            sv.loop(2, 5, "L")
            alone = sv.download(0)
        for k in range(3):
            if alone[k] is not None:
                np.testing.assert_array_equal(batched[i][k], alone[k])


def test_sweep_tuned_schedule_close_to_default():
    strings = sweep()[:12]
    say("[gpu] sweep_48 (12 strings): autotuned schedule vs default")
    with solver(strings) as sv:
        sv.loop(1, 5, "L")
        ref = [sv.download(i)[1] for i in range(len(strings))]
    with solver(strings) as sv:
        rows = sv.autotune(steps=2)
        assert all(r >= 1 for r in rows)
        sv.loop(1, 5, "L")
        got = [sv.download(i)[1] for i in range(len(strings))]
    for i in range(len(strings)):
        assert normwise(got[i], ref[i]) < 1e-12, i


def test_mixed_barrier_at_convergence():
    """Mixed vs FP64 barriers at convergence (the outer iteration changes
    the energies by less than 1e-10 relative): within 1e-6 relative."""
    from mepgl_amd import configs, observables

    s = configs.vortex_entry_small()
    say("[gpu] vortex_entry_small: mixed vs fp64 barriers at convergence")
    bars = {}
    for prec in ("fp64", "mixed"):
        with solver([s], precision=prec) as sv:
            e_old = None
            for it in range(400):
                sv.loop(1, 10, "L")
                e = sv.computeFreeEnergy()[0]
                if e_old is not None and np.max(np.abs(e - e_old) / np.abs(e)) < 1e-10:
                    break
                e_old = e
            bars[prec] = observables.barriers(e)
    for k in range(2):
        ref = bars["fp64"][k]
        if ref:
            assert abs(bars["mixed"][k] - ref) <= 1e-6 * abs(ref), (k, bars)
