"""Host logic of the measurement tools, on synthetic rocprofv3 outputs (CPU):
mepgl_amd.profiling (kernel classes, counter calibration, algorithmic
bytes, traffic ratios, the per-height traffic table) and bench.py's
roofline arithmetic and traffic lookup."""
import argparse
import csv
import json
import os

import numpy as np
import pytest

from mepgl_amd import configs, profiling


def _write_csv(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


G_UPD = "void mepgl::k_grad<float, 1, true, false, false, false>(mepgl::StepArgs)"
G_FIRST = "void mepgl::k_grad<float, 1, false, false, false, false>(mepgl::StepArgs)"
POLY = "void mepgl::k_poly<float, 1, false, false>(mepgl::StepArgs)"
COPY = "void copy_kernel<16>(double2 const*, double2*, long long)"


def test_kernel_classes_from_names():
    assert profiling._classify(G_UPD) == ("grad_update", ["float", "1", "true", "false", "false",
                                                          "false"])
    assert profiling._classify(G_FIRST)[0] == "grad_first"
    assert profiling._classify(POLY)[0] == "poly"
    assert profiling._classify("void mepgl::k_arclen<float, 1, false, true>(mepgl::AuxArgs)") is None
    # Trace and counter outputs print the name with or without the return
    # type and parameter list: both map to one key.
    assert profiling._key(G_UPD) == profiling._key("mepgl::k_grad<float, 1, true, false, false, "
                                                   "false>")


def test_summary_and_traffic_table(tmp_path, monkeypatch):
    """A synthetic profile of vortex_entry_small (15 images x 64^2, one
    component, mixed, 16-row segments): the counters report 1.05x the
    algorithmic reads of G and 1.0x its writes; the copy kernel's FETCH_SIZE
    reads half the bytes it moved (factor 2)."""
    out = tmp_path / "prof"
    s = configs.vortex_entry_small()
    cells = s.F * s.N * s.N
    cf, sb = 4, 4
    g_read = cells * 3 * cf * sb
    g_write = cells * 2 * cf * sb
    p_read = cells * 3 * cf * sb
    p_write = cells * cf * sb
    _write_csv(str(out / "trace" / "trace_kernel_stats.csv"), ["Name", "Calls", "AverageNs"],
               [[G_UPD, 9, 50000.0], [POLY, 10, 40000.0]])
    kib = 1024.0
    _write_csv(str(out / "fetch" / "fetch_counter_collection.csv"),
               ["Kernel_Name", "Counter_Name", "Counter_Value"],
               [[G_UPD, "FETCH_SIZE", 1.05 * g_read / kib / 2]] * 3 +
               [[POLY, "FETCH_SIZE", p_read / kib / 2]] * 3)
    _write_csv(str(out / "write" / "write_counter_collection.csv"),
               ["Kernel_Name", "Counter_Name", "Counter_Value"],
               [[G_UPD, "WRITE_SIZE", g_write / kib]] * 3 + [[POLY, "WRITE_SIZE", p_write / kib]] * 3)
    copy_kib = profiling.CALIB_MIB * 1024.0
    _write_csv(str(out / "cfetch" / "cfetch_counter_collection.csv"),
               ["Kernel_Name", "Counter_Name", "Counter_Value"], [[COPY, "FETCH_SIZE", copy_kib / 2]])
    _write_csv(str(out / "cwrite" / "cwrite_counter_collection.csv"),
               ["Kernel_Name", "Counter_Name", "Counter_Value"], [[COPY, "WRITE_SIZE", copy_kib]])
    a = argparse.Namespace(workload="vortex_entry_small", precision="mixed", rows=16, steps=20,
                           warmup=10, timeout=60, passes="", out=str(out))
    profiling.summarize(a)
    sm = json.load(open(out / "summary.json"))
    assert sm["calibration"]["fetch_factor"] == pytest.approx(2.0)
    assert sm["calibration"]["write_factor"] == pytest.approx(1.0)
    g = sm["kernels"]["grad_update"][0]
    assert g["algorithmic_bytes"] == cells * 5 * cf * sb
    assert g["achieved_GBs"] == pytest.approx(g["algorithmic_bytes"] / 50000.0)
    assert g["read_over_algorithmic"] == pytest.approx(1.05)
    assert g["traffic_over_algorithmic"] == pytest.approx((1.05 * g_read + g_write) /
                                                          (g_read + g_write))
    p = sm["kernels"]["poly"][0]
    assert p["read_over_algorithmic"] == pytest.approx(1.0)
    assert p["traffic_over_algorithmic"] == pytest.approx(1.0)
    # The recorded commands run committed files from the repository root.
    for cmd in sm["commands"].values():
        assert profiling.ROOT not in cmd
        assert "--legs ''" in cmd or "copy_calib" in cmd
    # The traffic table, keyed workload -> precision -> height -> kernel.
    table = tmp_path / "traffic.json"
    table.write_text(json.dumps({"vortex_entry_small": {"mixed": {"grad_update": 1}}}))
    monkeypatch.setattr(profiling, "TRAFFIC", str(table))
    profiling.merge(str(out / "summary.json"))
    t = json.load(open(table))
    ent = t["vortex_entry_small"]["mixed"]
    assert list(ent) == ["16"]  # the entry of the older flat form is dropped
    assert ent["16"]["grad_update"]["bytes"] == round(1.05 * g_read + g_write)
    assert ent["16"]["poly"]["read_over_algorithmic"] == pytest.approx(1.0)


def test_bench_roofline_and_traffic_lookup():
    import bench

    s = configs.vortex_entry_small()
    times = {"grad_update": (0.05, 90), "poly": (0.04, 100), "grad_first": (0.03, 10),
             "finish_poly": (0.005, 10)}
    r = bench.roofline([s], {"times": times}, 4)
    cells = s.F * s.N * s.N
    # G has the larger share of the step (0.05 ms x 90 > 0.04 ms x 100).
    assert r["kernel"] == "grad_update"
    assert r["algorithmic_bytes_per_step"] == cells * 5 * 4 * 4
    assert r["achieved"] == pytest.approx(cells * 80 / 0.05e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert "step_achieved" not in r
    # With the timed run's time per step: G's and P's bytes (9 C_f s per
    # image*cell) over it.
    rs = bench.roofline([s], {"times": times, "seconds": 0.02, "steps": 200}, 4)
    assert rs["step_achieved"] == pytest.approx(cells * 9 * 4 * 4 / 0.1e-3 / 1e9)
    assert rs["step_frac"] == pytest.approx(rs["step_achieved"] / 8000.0)
    # Two classes: one launch per class and step, bytes per string's C_f.
    s2 = configs.two_band_entry(64, 0.25, 5, 0.8, seed=3)
    r2 = bench.roofline([s, s2], {"times": times}, 4)
    assert r2["launches_per_step"] == 2
    assert r2["algorithmic_bytes_per_step"] == (cells * 5 * 4 + s2.F * 64 * 64 * 5 * 6) * 4
    assert r2["achieved"] == pytest.approx(r2["algorithmic_bytes_per_step"] / (2 * 0.05e-3) / 1e9)
    # The committed traffic table covers the headline workload at the heights
    # the autotuner picks, and nothing for heights never profiled.
    with open(os.path.join(bench.ROOT, "profiles", "traffic.json")) as f:
        table = json.load(f)
    for prec in ("mixed", "fp64"):
        for rows in table["vortex_entry_256"][prec]:
            b = bench.measured_traffic("vortex_entry_256", prec, "grad_update", int(rows))
            alg = table["vortex_entry_256"][prec][rows]["grad_update"]["algorithmic"]
            assert 1.0 <= b / alg < 1.2
    assert bench.measured_traffic("vortex_entry_256", "mixed", "poly", 5) is None
    assert np.isfinite(bench.barrier([3.0, 4.0, 2.5])[0])


def test_bench_json_line_contract(monkeypatch, capsys):
    """bench.main's JSON line, with the GPU runs replaced by synthetic
    results: the keys the benchmark contract names, a leg with its roofline
    and traffic fields, and a refusal for more than one GPU."""
    import bench

    def fake_run_path(strings, precision, steps, warmup, segment_rows=0, **kw):
        iters = max(steps // bench.R, 1)
        ms = {"mixed": 0.135, "fp64": 0.265}[precision]
        times = {"grad_update": (0.068, 27), "poly": (0.060, 30), "grad_first": (0.044, 3),
                 "finish_poly": (0.008, 3)}
        rows = segment_rows or 16
        return dict(seconds=iters * bench.R * ms * 1e-3, steps=iters * bench.R,
                    energies=[np.linspace(1.0, 2.0, s.F) for s in strings], times=times,
                    info={"tile_rows": rows, "rows": [rows] * len(strings), "parts": 2},
                    profile_steps=30)

    monkeypatch.setattr(bench, "run_path", fake_run_path)
    monkeypatch.setitem(configs.WORKLOADS, "two_band_small",
                        lambda: [configs.two_band_entry(64, 0.25, 5, 0.8, seed=3)])
    monkeypatch.setattr(bench, "streaming_ceilings",
                        lambda: {"copy": 6200.0, "triad": 5650.0, "triad_4B": 6150.0,
                                 "triad_8B": 5900.0})
    rc = bench.main(["--steps", "20", "--warmup", "5", "--no-cpu", "--workload",
                     "vortex_entry_small", "--legs", "vortex_entry_small,two_band_small"])
    assert rc == 0
    line = [l for l in capsys.readouterr().out.splitlines() if l.startswith("{")][-1]
    d = json.loads(line)
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "fp64", "mixed_over_fp64"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 20 and d["warmup"] == 5
    assert d["ms_per_step"] == pytest.approx(0.135)
    roof = d["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic", "step_kernels",
                "step_achieved", "step_frac"):
        assert key in roof, key
    assert d["config"]["split_images"] is True
    assert set(roof["step_kernels"]) == {"grad_update", "poly"}
    assert d["config"]["workload"] == "vortex_entry_small"
    # The headline workload is not repeated as a leg.
    assert list(d["legs"]) == ["two_band_small"]
    leg = d["legs"]["two_band_small"]
    for prec in ("mixed", "fp64"):
        assert leg[prec]["value"] > 0 and leg[prec]["roofline"]["frac"] > 0
        assert "traffic" in leg[prec]["roofline"]
        assert "traffic" in leg[prec]["roofline"]["step_kernels"]["poly"]
    assert leg["mixed_over_fp64"] == pytest.approx(0.265 / 0.135)
    assert bench.main(["--gpus", "2"]) == 2
