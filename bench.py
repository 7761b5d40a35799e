"""Benchmark of the relaxation loop: image*cells per second per NLCG step.

Workload (default): vortex_entry_256 - one-component string, 256x256 grid
(dx = 0.125), 64 images, H = 1.0, seeded vortex-entry initial guess; mode L
(linear reparametrisation every R = 10 NLCG steps, included in the timing).
A "step" is one NLCG iteration over all images; the timed region runs
`--steps` of them (default 1000 = 100 outer iterations) on fields already
resident in HBM.  Before it: `--warmup` steps, the autotuner's trial
rounds (the state is restored), a restart from the initial guess and at
least 4 untimed outer iterations and 50 ms from there, after which the
timed steps follow at once.  `value` is the mixed-precision path's throughput; the FP64
path is timed on identical inputs and reported beside it.  The energies and barriers of
the two are compared after exactly the timed number of steps from the initial guess, in
a run of their own after the timing.

Further workloads (--legs, default two_band_512, sweep_48 and
thin_film_disk_1024) run after it for min(K, 50) steps each and are
reported under "legs": both precisions, the dominant kernel's roofline
(mask-active cells only for the thin film) with its calibrated traffic at
the leg's segment height (fixed for the single-class legs), and for the
48-string sweep the batched context against one context per string in both
precisions.

The timed steps run the images in 2 or 4 parts on their own streams (the
library's default; the autotuner picks the count, reported as
`config.image_parts`); `--no-split` keeps one launch sequence, as the profiled
runs do.  `roofline` describes the dominant step kernel from a per-kernel
timing pass (one launch sequence, HIP events); `roofline.step_achieved` is G's
and P's algorithmic bytes per step over the timed time per step, where the
parts overlap.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--legs a,b,c] [--leg-steps K2] [--no-split]
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


WARM_S = 0.05  # least untimed device time right before a timed region


def warm_up(sv, iters):
    """Untimed outer iterations right before a timed region: at least
    `iters`, in pairs, and at least WARM_S seconds of them.  The device
    reaches its steady speed only after some 30 ms of load (vortex_entry_256
    mixed: 0.143-0.151 ms per step after 4 outer iterations, 0.136 after 20
    or more); the outer iterations' buffer parity repeats within two, so
    every graph the timed steps replay has been captured by the end."""
    t0, done = time.perf_counter(), 0
    while done < iters or time.perf_counter() - t0 < WARM_S:
        sv.loop(2, R, "L")
        sv.synchronize()
        done += 2


def run_path(strings, precision, steps, warmup, profile_iters=3, segment_rows=0, tune=True,
             tune_steps=10, split=True):
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    sv = Solver(strings, precision=precision, segment_rows=segment_rows)
    # Split: the steps of 2 or 4 parts of the images run on their own streams
    # (the library's default).  The per-kernel timing pass below always runs
    # one launch sequence.
    if not split:
        sv.set_control(_lib.CONTROL_SPLIT_IMAGES, 0)
    iters = max(steps // R, 1)
    w_iters = max(warmup // R, 1)
    # The warm-up steps, on a device brought to its steady speed before the
    # autotuner times its candidates.
    warm_up(sv, w_iters)
    # The library's autotuner picks each string class's segment height on
    # the live problem (the state is restored afterwards).
    rows = sv.autotune(steps=tune_steps) if (tune and not segment_rows) else None
    # Re-start from the initial guess, so that both precisions run the same
    # trajectory, and warm up from there (a new schedule also drops the
    # captured graphs); the timed steps follow at once.
    for i, s in enumerate(strings):
        sv.upload(i, s)
    sv.reset_solver()
    warm_up(sv, max(w_iters, 4))
    t0 = time.perf_counter()
    sv.loop(iters, R, "L")
    sv.synchronize()
    dt = time.perf_counter() - t0
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
    # The energies the two precisions are compared on: exactly the timed
    # number of steps from the initial guess.  The warm-up is time-based, so
    # it runs about twice as many iterations in mixed precision as in FP64,
    # and the state after the timed steps is not comparable between them.
    for i, s in enumerate(strings):
        sv.upload(i, s)
    sv.reset_solver()
    sv.loop(iters, R, "L")
    energies = sv.computeFreeEnergy()
    info = sv.info(0)
    info["rows"] = rows if rows is not None else [sv.info(i)["tile_rows"]
                                                  for i in range(len(strings))]
    # Image parts of the split step sequence (0: one sequence).
    info["parts"] = sv.get_control(_lib.CONTROL_SPLIT_IMAGES)
    sv.close()
    return dict(seconds=dt, steps=iters * R, energies=energies, times=times, info=info,
                profile_steps=profile_iters * R)


def run_separately(strings, precision, steps, warmup, tune_steps=2):
    """The same strings, one context (and so one launch sequence) each, run
    one after another: the unbatched reference point for a sweep."""
    from mepgl_amd.solver import Solver

    iters = max(steps // R, 1)
    total = 0.0
    for s in strings:
        sv = Solver([s], precision=precision)
        sv.loop(max(warmup // R, 1), R, "L")
        sv.autotune(steps=tune_steps)
        # As in run_path: restart, warm up, then time at once.
        sv.upload(0, s)
        sv.reset_solver()
        warm_up(sv, max(warmup // R, 4))
        t0 = time.perf_counter()
        sv.loop(iters, R, "L")
        sv.synchronize()
        total += time.perf_counter() - t0
        sv.close()
    return dict(seconds=total, steps=iters * R)


def c_f(s):
    """Streamed reals per image*cell: A (unless fixed) and psi."""
    return (0 if s.no_self_field else 2) + 2 * s.ncomp


def active_cells(s):
    """Image*cells inside the domains (every cell for the full box)."""
    dom = (np.asarray(s.sc_domain) != 0) | (np.asarray(s.comp_domain) != 0)
    return s.F * int(np.count_nonzero(dom))


def kernel_classes(strings):
    """The G and P kernels launch once per class of strings (components,
    fixed A, parameter planes)."""
    from mepgl_amd.problem import PARAM_NAMES

    return len({(s.ncomp, bool(s.no_self_field),
                 any(not np.isscalar(getattr(s, n)) for n in PARAM_NAMES)) for s in strings})


def roofline(strings, r, sbytes):
    """Achieved algorithmic GB/s of the dominant step kernel.

    Algorithmic bytes per mask-active image*cell: G with the pending update
    reads x, d, g_old and writes x, g (5 C_f s); P reads x, g, d and writes
    d (4 C_f s).  The kernel time per step is the average launch time times
    the number of item classes (each class is one launch per step)."""
    ncls = kernel_classes(strings)
    act = [active_cells(s) for s in strings]
    out, others = None, {}
    for name, per_cf in (("grad_update", 5), ("poly", 4)):
        ms, n = r["times"][name]
        if not n:
            continue
        step_ms = ms * ncls
        nbytes = sum(a * per_cf * c_f(s) * sbytes for a, s in zip(act, strings))
        cand = {"kernel": name, "achieved": nbytes / (step_ms * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "avg_ms": ms, "launches_per_step": ncls,
                "algorithmic_bytes_per_step": nbytes,
                "algorithmic_bytes_per_image_cell": per_cf * c_f(strings[0]) * sbytes
                if len({c_f(s) for s in strings}) == 1 else None}
        cand["frac"] = cand["achieved"] / HBM_PEAK_GBS
        # The dominant kernel: the larger share of the step's time.
        share = ms * n
        if out is None or share > out[0]:
            out = (share, cand)
        others[name] = {"achieved": cand["achieved"], "frac": cand["frac"], "avg_ms": ms}
    if out is None:
        return None
    # Both step kernels' figures, the dominant one included.
    out[1]["step_kernels"] = others
    # The step as a whole: G's and P's algorithmic bytes over the timed time
    # per step (reparametrisation and finishing launches included).  With
    # the image parts on their own streams, one part's G and another's P
    # overlap, so this is the figure that sees that overlap.
    if r.get("steps"):
        step_bytes = sum(a * 9 * c_f(s) * sbytes for a, s in zip(act, strings))
        step_s = r["seconds"] / r["steps"]
        out[1]["step_achieved"] = step_bytes / step_s / 1e9
        out[1]["step_frac"] = out[1]["step_achieved"] / HBM_PEAK_GBS
    return out[1]


def barrier_errors(e_mixed, e_fp64):
    """Mixed vs FP64 barriers and energies of every string (post.py:188-189)."""
    rel_b, e_rel = [], 0.0
    for m, d in zip(e_mixed, e_fp64):
        bm, bd = barrier(m), barrier(d)
        rel_b.append([abs(x - y) / abs(y) if y else None for x, y in zip(bm, bd)])
        e_rel = max(e_rel, float(np.max(np.abs(m - d) / np.abs(d))))
    vals = [v for b in rel_b for v in b if v is not None]
    return (max(vals) if vals else None), e_rel


def leg(name, steps, warmup, tune_steps=4, separate=False):
    """One bounded run of a further workload: mixed and FP64 paths on the
    same seeded strings (and, for a sweep, one context per string)."""
    from mepgl_amd import configs

    t0 = time.perf_counter()
    strings = configs.WORKLOADS[name]()
    gen_s = time.perf_counter() - t0
    cells = sum(s.image_cells for s in strings)
    act = sum(active_cells(s) for s in strings)
    out = {"workload": name, "strings": len(strings), "image_cells": cells,
           "active_cell_fraction": act / cells}
    res = {}
    for prec in ("mixed", "fp64"):
        r = run_path(strings, prec, steps, warmup, tune_steps=tune_steps,
                     segment_rows=LEG_ROWS.get(name, 0))
        res[prec] = r
        v = cells * r["steps"] / r["seconds"]
        roof = roofline(strings, r, 4 if prec == "mixed" else 8)
        rows = sorted(set(r["info"]["rows"]))
        if roof is not None:
            # Calibrated HBM bytes per launch at the leg's height, when a
            # single class ran at one recorded height.
            single = len(rows) == 1 and roof["launches_per_step"] == 1
            roof["traffic"] = measured_traffic(name, prec, roof["kernel"], rows[0]) if single else None
            for kname, kfig in roof["step_kernels"].items():
                kfig["traffic"] = measured_traffic(name, prec, kname, rows[0]) if single else None
        out[prec] = {"value": v, "value_active_cells": v * act / cells,
                     "ms_per_step": 1e3 * r["seconds"] / r["steps"], "steps": r["steps"],
                     "segment_rows": rows, "roofline": roof}
        print(f"# {name} {prec}: {r['steps']} steps in {r['seconds']:.4f} s", flush=True)
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    b_err, e_err = barrier_errors(res["mixed"]["energies"], res["fp64"]["energies"])
    out["after_steps"] = res["mixed"]["steps"]
    out["barrier_rel_err_mixed_vs_fp64_unconverged"] = b_err
    out["energy_rel_err_mixed_vs_fp64"] = e_err
    if separate:
        for prec in ("mixed", "fp64"):
            r = run_separately(strings, prec, steps, warmup)
            out[f"separate_{prec}"] = cells * r["steps"] / r["seconds"]
            print(f"# {name} separate {prec}: {r['steps']} steps in {r['seconds']:.4f} s",
                  flush=True)
        out["batched_over_separate_mixed"] = out["mixed"]["value"] / out["separate_mixed"]
        # The target ratio: the mixed batched solver over the plain FP64 path
        # (one string per context).
        out["batched_mixed_over_plain_fp64"] = out["mixed"]["value"] / out["separate_fp64"]
    out["wall_s"] = time.perf_counter() - t0
    out["generate_s"] = gen_s
    return out


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
                  f"string 0 ({s.F} images x {s.N}^2), oracle/mepgl_oracle.c FP64 built "
                  f"-O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (oracle/Makefile), "
                  f"{threads} threads, {cpu_model()}",
    }


def streaming_ceilings():
    """Plain copy and 3-read/1-write kernels of 1 GiB per array on this GPU
    (mepgl_amd/copy_calib), for the roofline's measured ceilings: copy and
    triad with 16-byte accesses per lane, and triad with the 4- and 8-byte
    accesses of the mixed and FP64 step kernels."""
    import re
    import subprocess

    exe = os.path.join(ROOT, "mepgl_amd", "copy_calib")
    out = {}
    for key, mode, width in (("copy", "copy", 16), ("triad", "triad", 16),
                             ("triad_4B", "triad", 4), ("triad_8B", "triad", 8)):
        try:
            r = subprocess.run([exe, "1024", "10", mode, str(width)], capture_output=True,
                               text=True, timeout=60)
            m = re.search(r"([0-9.]+) GB/s", r.stdout)
            out[key] = float(m.group(1)) if (r.returncode == 0 and m) else None
        except (OSError, subprocess.SubprocessError):
            out[key] = None
    return out


def measured_traffic(workload, precision, kernel, rows):
    """HBM bytes per launch of `kernel` at segment height `rows` from the
    calibrated rocprofv3 counter runs recorded in profiles/traffic.json
    (null when that height was not recorded)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
        return table[workload][precision][str(rows)][kernel]["bytes"]
    except (OSError, KeyError, ValueError, TypeError):
        return None


LEGS = ("two_band_512", "sweep_48", "thin_film_disk_1024")
# Segment heights of the single-class legs, fixed at heights whose HBM
# traffic profiles/traffic.json records (the autotuner's picks for these
# workloads vary from run to run within about 1 % of these heights' times);
# the sweep's classes are autotuned.
LEG_ROWS = {"two_band_512": 16, "thin_film_disk_1024": 25}


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
    ap.add_argument("--no-split", action="store_true",
                    help="one launch sequence for all images (profiling: no overlapping kernels)")
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="further workloads run after the headline one ('' for none)")
    ap.add_argument("--leg-steps", type=int, default=0,
                    help="NLCG steps of each further workload (default: min(steps, 50))")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: multi-GPU runs are not part of "
                                   "it (strings and image ranges would shard with no exchange "
                                   "between reparametrisations)"}))
        return 2
    from mepgl_amd import configs

    strings = configs.WORKLOADS[args.workload]()
    cells = sum(s.image_cells for s in strings)
    act_cells = sum(active_cells(s) for s in strings)
    active = act_cells / cells

    if args.compare_batching:
        print(json.dumps(leg(args.workload, args.steps, args.warmup, tune_steps=10,
                             separate=True)))
        return 0

    res = {}
    precs = ("fp64", "mixed") if args.precision == "both" else (args.precision,)
    for prec in precs:
        res[prec] = run_path(strings, prec, args.steps, args.warmup,
                             segment_rows=args.segment_rows, split=not args.no_split)
        print(f"# {prec}: {res[prec]['steps']} steps in {res[prec]['seconds']:.4f} s", flush=True)
    if args.precision != "both":
        r = res[args.precision]
        print(json.dumps({"precision": args.precision, "ms_per_step": 1e3 * r["seconds"] / r["steps"],
                          "value": cells * r["steps"] / r["seconds"],
                          "segment_rows": r["info"]["tile_rows"],
                          "roofline": roofline(strings, r, 4 if args.precision == "mixed" else 8)}))
        return 0

    mixed, fp64 = res["mixed"], res["fp64"]
    value = cells * mixed["steps"] / mixed["seconds"]
    value64 = cells * fp64["steps"] / fp64["seconds"]
    roof = roofline(strings, mixed, 4)
    roof64 = roofline(strings, fp64, 8)
    b_m = barrier(mixed["energies"][0])
    b_d = barrier(fp64["energies"][0])
    rel_b = [abs(m - d) / abs(d) if d else None for m, d in zip(b_m, b_d)]
    e_rel = float(np.max(np.abs(mixed["energies"][0] - fp64["energies"][0]) /
                         np.abs(fp64["energies"][0])))
    rows_m = mixed["info"]["tile_rows"]
    ceil = streaming_ceilings()
    roof.update({
        "bound": "hbm",
        "traffic": measured_traffic(args.workload, "mixed", roof["kernel"], rows_m),
        "traffic_segment_rows": rows_m,
        "algorithmic_bytes_per_launch": roof["algorithmic_bytes_per_step"],
        "measured_ceilings_GBs": ceil,
        "frac_of_triad_ceiling": roof["achieved"] / ceil["triad"] if ceil.get("triad") else None})
    roof64["traffic"] = measured_traffic(args.workload, "fp64", roof64["kernel"],
                                         fp64["info"]["tile_rows"])
    for rf, prec, rows in ((roof, "mixed", rows_m), (roof64, "fp64", fp64["info"]["tile_rows"])):
        for kname, kfig in rf["step_kernels"].items():
            kfig["traffic"] = measured_traffic(args.workload, prec, kname, rows)

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
                   "segment_rows": rows_m, "split_images": not args.no_split,
                   "image_parts": mixed["info"]["parts"],
                   "segment_rows_fp64": fp64["info"]["tile_rows"],
                   "image_parts_fp64": fp64["info"]["parts"]},
        "fp64": {"value": value64, "value_active_cells": value64 * active,
                 "ms_per_step": 1e3 * fp64["seconds"] / fp64["steps"], "roofline": roof64},
        "mixed_over_fp64": value / value64,
        "energy_rel_err_mixed_vs_fp64": e_rel,
        "barrier_fp64": {"entry": b_d[0], "exit": b_d[1]},
        "barrier_mixed": {"entry": b_m[0], "exit": b_m[1]},
        # After `steps` steps from the initial guess, not at convergence
        # (the converged comparison is a GPU test).
        "barrier_rel_err_mixed_vs_fp64": {"entry": rel_b[0], "exit": rel_b[1]},
        "roofline": roof,
        "active_cell_fraction": active,
        # For masked geometries the metric counts mask-active image*cells
        # only; equal to value when the domain is the full box.
        "value_active_cells": value * active,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(strings, args.cpu_seconds)
    legs = [w for w in args.legs.split(",") if w and w != args.workload]
    if legs:
        steps = args.leg_steps or min(args.steps, 50)
        out["legs"] = {}
        for w in legs:
            print(f"# leg {w}: {steps} steps", flush=True)
            out["legs"][w] = leg(w, steps, min(args.warmup, 10), separate=(w == "sweep_48"))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
