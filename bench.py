"""Benchmark of the relaxation loop: image*cells per second per NLCG step.

Workload (default): vortex_entry_256 - one-component string, 256x256 grid
(dx = 0.125), 64 images, H = 1.0, seeded vortex-entry initial guess; mode L
(linear reparametrisation every R = 10 NLCG steps, included in the timing).
A "step" is one NLCG iteration over all images; the timed region runs
`--steps` of them (default 1000 = 100 outer iterations) on fields already
resident in HBM.  `value` is the mixed-precision path's throughput; the FP64
path is timed on identical inputs and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
R = 10


def barrier(e):
    """post.py:188-189: max(F - F_0) and max(F - F_last)."""
    e = np.asarray(e)
    return float(np.max(e - e[0])), float(np.max(e - e[-1]))


KERNEL_NAMES = ((0, "grad_update"), (1, "poly"), (2, "grad_first"), (4, "finish_poly"))


def run_path(strings, precision, steps, warmup, profile_iters=3, segment_rows=0, tune=True):
    from mepgl_amd.solver import Solver

    sv = Solver(strings, precision=precision, segment_rows=segment_rows)
    iters = max(steps // R, 1)
    w_iters = max(warmup // R, 1)
    sv.loop(w_iters, R, "L")
    sv.synchronize()
    # The library's autotuner picks each string class's segment height on
    # the live problem (the state is restored afterwards).
    rows = sv.autotune(steps=10) if (tune and not segment_rows) else None
    # Re-start from the initial guess so that both precisions time the same
    # trajectory; warm-up only compiled graphs and touched memory.
    for i, s in enumerate(strings):
        sv.upload(i, s)
    sv.reset_solver()
    sv.synchronize()
    t0 = time.perf_counter()
    sv.loop(iters, R, "L")
    sv.synchronize()
    dt = time.perf_counter() - t0
    energies = sv.computeFreeEnergy()
    # Per-kernel timing with HIP events on the library's stream (eager
    # launches, same sequence), for the roofline of the dominant kernel.
    sv.profile(True)
    sv.loop(profile_iters, R, "L")
    sv.synchronize()
    times = {}
    for k, name in KERNEL_NAMES:
        ms, cnt = sv.profile_read(k)
        times[name] = (ms, cnt)
        print(f"# {precision} {name}: {1e3 * ms:.2f} us x {cnt}", flush=True)
    sv.profile(False)
    info = sv.info(0)
    info["rows"] = rows
    sv.close()
    return dict(seconds=dt, steps=iters * R, energies=energies, times=times, info=info)


def run_separately(strings, precision, steps, warmup):
    """The same strings, one context (and so one launch sequence) each, run
    one after another: the unbatched reference point for a sweep."""
    from mepgl_amd.solver import Solver

    iters = max(steps // R, 1)
    total = 0.0
    for s in strings:
        sv = Solver([s], precision=precision)
        sv.loop(max(warmup // R, 1), R, "L")
        sv.autotune(steps=2)
        sv.upload(0, s)
        sv.reset_solver()
        sv.synchronize()
        t0 = time.perf_counter()
        sv.loop(iters, R, "L")
        sv.synchronize()
        total += time.perf_counter() - t0
        sv.close()
    return dict(seconds=total, steps=iters * R)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(strings, seconds):
    """The FP64 CPU oracle on a bounded sample of the same workload: after
    2 warm-up steps, NLCG steps of string 0 in chunks of 5 until about
    `seconds` have passed; the rate is the median over chunks."""
    from oracle import oracle as O

    O.build()
    s = strings[0]
    ref = O.OracleString(s)
    ref.nlcg_steps(2)  # warm-up (page-in, thread pool)
    rates, done = [], 0
    t0 = time.perf_counter()
    while True:
        t1 = time.perf_counter()
        ref.nlcg_steps(5)
        rates.append(s.image_cells * 5 / (time.perf_counter() - t1))
        done += 5
        dt = time.perf_counter() - t0
        if dt >= seconds and len(rates) >= 5:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": float(np.median(rates)),
        "unit": "image*cells/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} NLCG steps ({dt:.1f} s, median of {len(rates)} chunks of 5) of "
                  f"string 0 ({s.F} images x {s.N}^2), oracle/mepgl_oracle.c FP64 -O3 "
                  f"-fopenmp, {threads} threads, {cpu_model()}",
    }


def streaming_ceilings():
    """Plain copy and 3-read/1-write kernels of 1 GiB on this GPU
    (mepgl_amd/copy_calib), for the roofline's measured ceilings."""
    import re
    import subprocess

    exe = os.path.join(ROOT, "mepgl_amd", "copy_calib")
    out = {}
    for mode in ("copy", "triad"):
        try:
            r = subprocess.run([exe, "1024", "10", mode], capture_output=True, text=True,
                               timeout=60)
            m = re.search(r"([0-9.]+) GB/s", r.stdout)
            out[mode] = float(m.group(1)) if (r.returncode == 0 and m) else None
        except (OSError, subprocess.SubprocessError):
            out[mode] = None
    return out


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` from the calibrated rocprofv3 counter
    runs recorded in profiles/traffic.json (null when not recorded)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
        return table[workload][precision][kernel]
    except (OSError, KeyError, ValueError):
        return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="vortex_entry_256")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--compare-batching", action="store_true",
                    help="multi-string workloads: batched vs one string per context")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="both", choices=("both", "fp64", "mixed"))
    ap.add_argument("--segment-rows", type=int, default=0)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: multi-GPU runs are not part of "
                                   "it (strings and image ranges would shard with no exchange "
                                   "between reparametrisations)"}))
        return 2
    from mepgl_amd import configs

    strings = configs.WORKLOADS[args.workload]()
    cells = sum(s.image_cells for s in strings)
    C_f = [4 if s.no_self_field else 2 + 2 * s.ncomp for s in strings]
    cf = C_f[0]

    if args.compare_batching:
        steps = args.steps
        bat = run_path(strings, "mixed", steps, args.warmup)
        bat64 = run_path(strings, "fp64", steps, args.warmup)
        sep = run_separately(strings, "mixed", steps, args.warmup)
        v_b = cells * bat["steps"] / bat["seconds"]
        print(json.dumps({
            "workload": args.workload, "strings": len(strings), "image_cells": cells,
            "batched_mixed": v_b,
            "batched_fp64": cells * bat64["steps"] / bat64["seconds"],
            "separate_mixed": cells * sep["steps"] / sep["seconds"],
            "batched_over_separate": v_b / (cells * sep["steps"] / sep["seconds"]),
            "mixed_over_fp64": v_b / (cells * bat64["steps"] / bat64["seconds"]),
            "rows_mixed": bat["info"]["rows"], "unit": "image*cells/s"}))
        return 0

    res = {}
    precs = ("fp64", "mixed") if args.precision == "both" else (args.precision,)
    for prec in precs:
        res[prec] = run_path(strings, prec, args.steps, args.warmup,
                             segment_rows=args.segment_rows)
        print(f"# {prec}: {res[prec]['steps']} steps in {res[prec]['seconds']:.4f} s", flush=True)
    if args.precision != "both":
        r = res[args.precision]
        print(json.dumps({"precision": args.precision, "ms_per_step": 1e3 * r["seconds"] / r["steps"],
                          "value": cells * r["steps"] / r["seconds"],
                          "segment_rows": r["info"]["tile_rows"]}))
        return 0

    mixed, fp64 = res["mixed"], res["fp64"]
    value = cells * mixed["steps"] / mixed["seconds"]
    value64 = cells * fp64["steps"] / fp64["seconds"]

    # Roofline of the dominant kernel (mixed path), algorithmic bytes:
    # G with update reads x, d, g_old and writes x, g: 5 C_f s per image*cell;
    # P reads x, g, d and writes d: 4 C_f s.
    def roof(r, sbytes):
        g_ms, g_n = r["times"]["grad_update"]
        p_ms, p_n = r["times"]["poly"]
        cand = [("grad_update", g_ms, g_n, 5 * cf * sbytes), ("poly", p_ms, p_n, 4 * cf * sbytes)]
        name, ms, n, per = max(cand, key=lambda t: t[1] * t[2])
        ach = cells * per / (ms * 1e-3) / 1e9
        return name, ms, ach, per

    kname, kms, ach, per = roof(mixed, 4)
    kname64, kms64, ach64, per64 = roof(fp64, 8)
    b_m = barrier(mixed["energies"][0])
    b_d = barrier(fp64["energies"][0])
    rel_b = [abs(m - d) / abs(d) if d else None for m, d in zip(b_m, b_d)]
    e_rel = float(np.max(np.abs(mixed["energies"][0] - fp64["energies"][0]) /
                         np.abs(fp64["energies"][0])))
    traffic = measured_traffic(args.workload, "mixed", kname)
    ceil = streaming_ceilings()
    # Cells inside the superconducting domain (the active fraction of masked
    # geometries; the kernels stream every cell).
    active = sum(s.F * float(np.count_nonzero(s.sc_domain)) for s in strings) / cells

    out = {
        "metric": "image*cells/s per NLCG step (mixed precision)",
        "value": value,
        "unit": "image*cells/s",
        "n_gpus": 1,
        "steps": mixed["steps"],
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["seconds"] / mixed["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields / f64 reductions",
        "data": "synthetic seeded vortex-entry initial guess",
        "config": {"workload": args.workload, "strings": len(strings),
                   "images": sum(s.F for s in strings), "N": strings[0].N,
                   "components": strings[0].ncomp, "mode": "L", "reparam_every": R,
                   "segment_rows": mixed["info"]["tile_rows"],
                   "segment_rows_fp64": fp64["info"]["tile_rows"]},
        "fp64": {"value": value64, "value_active_cells": value64 * active,
                 "ms_per_step": 1e3 * fp64["seconds"] / fp64["steps"],
                 "roofline": {"kernel": kname64, "achieved": ach64, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": ach64 / HBM_PEAK_GBS,
                              "avg_ms": kms64}},
        "mixed_over_fp64": value / value64,
        "energy_rel_err_mixed_vs_fp64": e_rel,
        "barrier_fp64": {"entry": b_d[0], "exit": b_d[1]},
        "barrier_mixed": {"entry": b_m[0], "exit": b_m[1]},
        "barrier_rel_err_mixed_vs_fp64": {"entry": rel_b[0], "exit": rel_b[1]},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_image_cell": per,
                     "algorithmic_bytes_per_launch": per * cells, "avg_ms": kms,
                     "measured_ceilings_GBs": ceil,
                     "frac_of_triad_ceiling": ach / ceil["triad"] if ceil.get("triad") else None},
        "active_cell_fraction": active,
        # For masked geometries the metric counts mask-active image*cells
        # only; equal to value when the domain is the full box.
        "value_active_cells": value * active,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(strings, args.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
