"""rocprofv3 measurement of the step kernels: durations, HBM traffic and
roofline fractions at a fixed segment height.

    python -m mepgl_amd.profiling run --workload vortex_entry_256 \
        --precision mixed --rows 16 --out prof/vortex_entry_256_mixed_16
    python -m mepgl_amd.profiling merge prof/vortex_entry_256_mixed_16/summary.json

`run` (on a GPU host) profiles `bench.py` at a fixed segment height (no
autotuning, so every launch has the same shape) in three rocprofv3 passes:
kernel trace with statistics, FETCH_SIZE and WRITE_SIZE, each counter in a
run of its own.  It also calibrates the two counters on the plain copy
kernel of copy_calib (256 MiB, 16 bytes per lane): the factor is the known
byte count over what the counter reports for it.  It writes summary.json and
summary.md under --out.

`merge` records a summary's per-launch HBM bytes in profiles/traffic.json
under workload -> precision -> segment height -> kernel, where bench.py
looks them up for the height its autotuner picked.

Per-launch algorithmic bytes count mask-active image*cells only:
G with the pending update 5 C_f s, G without it 3 C_f s, P 4 C_f s per
image*cell (C_f streamed reals: A unless fixed, and psi; s = 4 or 8 bytes).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import shlex
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
HBM_PEAK_GBS = 8000.0
CALIB_MIB = 256

# Per-image*cell algorithmic traffic of each step kernel, in units of C_f s.
KERNEL_UNITS = {"grad_update": 5, "grad_first": 3, "poly": 4}


def _classify(name: str):
    """(kernel id, template arguments) of a step-kernel name, or None."""
    m = re.search(r"k_(grad|poly)<([^>]*)>", name)
    if not m:
        return None
    args = [a.strip() for a in m.group(2).split(",")]
    if m.group(1) == "poly":
        return "poly", args
    return ("grad_update" if args[2] == "true" else "grad_first"), args


def _rocprof(args, out_dir, prefix, timeout):
    cmd = ["timeout", "-s", "KILL", str(timeout), "rocprofv3", *args, "-d", out_dir, "-o", prefix,
           "--output-format", "csv", "--"]
    return cmd


def _run(cmd, log):
    print("[profiling]", " ".join(cmd), flush=True)
    with open(log, "w") as f:
        r = subprocess.run(cmd, stdout=f, stderr=subprocess.STDOUT)
    return r.returncode


def _find(out_dir, pattern):
    hits = sorted(glob.glob(os.path.join(out_dir, "**", pattern), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no {pattern} under {out_dir}")
    return hits[-1]


def _key(name: str) -> str:
    """A kernel name without its return type and parameter list, so that the
    trace and counter outputs match whichever form each prints."""
    return name.split("(")[0].replace("void ", "").replace(" ", "")


def _col(row, *names):
    for n in names:
        if n in row and row[n] != "":
            return row[n]
    raise KeyError(names)


def kernel_stats(out_dir):
    """Kernel name -> (calls, average ns) from the --stats summary."""
    out = {}
    with open(_find(out_dir, "*kernel_stats.csv")) as f:
        for row in csv.DictReader(f):
            out[_key(_col(row, "Name", "KernelName", "Kernel_Name"))] = (
                int(float(_col(row, "Calls"))), float(_col(row, "AverageNs", "Average_Ns")))
    return out


def counter_per_dispatch(out_dir, counter):
    """Kernel name -> list of per-dispatch counter values (KiB)."""
    out = {}
    with open(_find(out_dir, "*counter_collection.csv")) as f:
        for row in csv.DictReader(f):
            if _col(row, "Counter_Name", "CounterName") != counter:
                continue
            out.setdefault(_key(_col(row, "Kernel_Name", "KernelName", "Name")), []).append(
                float(_col(row, "Counter_Value", "CounterValue")))
    return out


def _bench_cmd(a):
    # One launch sequence (--no-split): the kernels of the profiled run do not
    # overlap, so each launch's duration is its own.
    return [sys.executable, os.path.join(ROOT, "bench.py"), "--no-cpu", "--no-split", "--legs", "",
            "--workload", a.workload, "--precision", a.precision, "--segment-rows", str(a.rows),
            "--steps", str(a.steps), "--warmup", str(a.warmup)]


def _class_cells(strings):
    """Mask-active image*cells and C_f per kernel class (components, fixed A,
    parameter planes)."""
    import numpy as np

    from .problem import PARAM_NAMES

    out = {}
    for s in strings:
        pl = any(not np.isscalar(getattr(s, n)) for n in PARAM_NAMES)
        key = (s.ncomp, bool(s.no_self_field), pl)
        dom = (np.asarray(s.sc_domain) != 0) | (np.asarray(s.comp_domain) != 0)
        cf = (0 if s.no_self_field else 2) + 2 * s.ncomp
        cells, _ = out.get(key, (0, cf))
        out[key] = (cells + s.F * int(np.count_nonzero(dom)), cf)
    return out


def _portable(cmd):
    """A recorded command with the repository root and interpreter written
    portably (run from the repository root)."""
    return shlex.join(cmd).replace(sys.executable, "python3").replace(ROOT + os.sep, "")


def _passes(a):
    bench = _bench_cmd(a)
    calib = [os.path.join(HERE, "copy_calib"), str(CALIB_MIB), "3", "copy", "16"]
    passes = [("trace", ["--kernel-trace", "--stats"], bench),
              ("fetch", ["--pmc", "FETCH_SIZE"], bench),
              ("write", ["--pmc", "WRITE_SIZE"], bench),
              ("cfetch", ["--pmc", "FETCH_SIZE"], calib),
              ("cwrite", ["--pmc", "WRITE_SIZE"], calib)]
    if a.passes:
        keep = set(a.passes.split(","))
        passes = [p for p in passes if p[0] in keep or p[0] in ("cfetch", "cwrite")]
    return [(name, _rocprof(opts, os.path.join(a.out, name), name, a.timeout) + prog)
            for name, opts, prog in passes]


def run(a):
    os.makedirs(a.out, exist_ok=True)
    for name, cmd in _passes(a):
        rc = _run(cmd, os.path.join(a.out, name + ".log"))
        # rocprofv3 itself may end with a segmentation fault after a complete
        # run; the outputs are checked below instead of its status.
        print(f"[profiling] pass {name}: rc {rc}", flush=True)
        if rc in (124, 137):
            raise SystemExit(f"pass {name} hit its time limit")
    summarize(a)


def summarize(a):
    """summary.json and summary.md from the pass outputs under a.out."""
    from . import configs

    cmds = {name: _portable(cmd[4:]) for name, cmd in _passes(a)
            if os.path.isdir(os.path.join(a.out, name))}

    def opt(fn, *args):
        try:
            return fn(*args)
        except FileNotFoundError:
            return {}

    stats = opt(kernel_stats, os.path.join(a.out, "trace"))
    fetch = opt(counter_per_dispatch, os.path.join(a.out, "fetch"), "FETCH_SIZE")
    write = opt(counter_per_dispatch, os.path.join(a.out, "write"), "WRITE_SIZE")
    if not stats:
        # Without the trace pass, the counter runs name the kernels.
        stats = {k: (len(v), 0.0) for k, v in (fetch or write).items()}
    cf = counter_per_dispatch(os.path.join(a.out, "cfetch"), "FETCH_SIZE")
    cw = counter_per_dispatch(os.path.join(a.out, "cwrite"), "WRITE_SIZE")
    copy_kib = CALIB_MIB * 1024.0

    def calib_factor(vals):
        v = [x for k, xs in vals.items() if "copy_kernel" in k for x in xs]
        return copy_kib / (sum(v) / len(v)) if v else None

    f_fac, w_fac = calib_factor(cf), calib_factor(cw)
    strings = configs.WORKLOADS[a.workload]()
    classes = _class_cells(strings)
    sbytes = 4 if a.precision == "mixed" else 8
    kernels = {}
    for name, (calls, avg_ns) in stats.items():
        cl = _classify(name)
        if cl is None:
            continue
        kid, targs = cl
        key = (int(targs[1]), targs[3 if kid != "poly" else 2] == "true",
               targs[4 if kid != "poly" else 3] == "true")
        cells, c_f = classes.get(key, (0, 0))
        alg = cells * KERNEL_UNITS[kid] * c_f * sbytes
        rec = {"name": name, "calls": calls, "avg_us": avg_ns / 1e3,
               "algorithmic_bytes": alg,
               "achieved_GBs": alg / avg_ns if avg_ns else None}
        rec["frac"] = rec["achieved_GBs"] / HBM_PEAK_GBS if rec["achieved_GBs"] else None
        fv, wv = fetch.get(name), write.get(name)
        if fv and f_fac:
            rec["read_bytes"] = sum(fv) / len(fv) * 1024 * f_fac
            rec["dispatches"] = len(fv)
            # Reads: x, d and g_old (G with update), x and g_old (G first),
            # x, g and d (P).
            reads = {"grad_update": 3, "grad_first": 2, "poly": 3}[kid]
            rec["read_over_algorithmic"] = rec["read_bytes"] / (cells * reads * c_f * sbytes)
        if wv and w_fac:
            rec["write_bytes"] = sum(wv) / len(wv) * 1024 * w_fac
        if "read_bytes" in rec and "write_bytes" in rec:
            rd, wr = rec["read_bytes"], rec["write_bytes"]
            rec.update({"bytes": rd + wr,
                        "traffic_over_algorithmic": (rd + wr) / alg if alg else None})
        kernels.setdefault(kid, []).append(rec)
    summary = {"workload": a.workload, "precision": a.precision, "segment_rows": a.rows,
               "steps": a.steps, "warmup": a.warmup, "calibration":
               {"fetch_factor": f_fac, "write_factor": w_fac, "copy_kib": copy_kib},
               "commands": cmds, "kernels": kernels, "when": time.strftime("%Y-%m-%d %H:%M")}
    with open(os.path.join(a.out, "summary.json"), "w") as f:
        json.dump(summary, f, indent=1)
    with open(os.path.join(a.out, "summary.md"), "w") as f:
        f.write(markdown(summary))
    print(markdown(summary), flush=True)


def markdown(sm):
    lines = [f"### {sm['workload']}, {sm['precision']}, {sm['segment_rows']}-row segments",
             "", "| Kernel | Calls | Average (us) | Algorithmic MB | GB/s (algorithmic) | "
             "of 8 TB/s | HBM read MB | read / algorithmic read | HBM write MB | "
             "traffic / algorithmic |",
             "|---|---|---|---|---|---|---|---|---|---|"]

    def mb(r, k):
        return f"{r[k] / 1e6:.1f}" if r.get(k) is not None else "-"

    def num(r, k, fmt):
        return format(r[k], fmt) if r.get(k) is not None else "-"

    for recs in sm["kernels"].values():
        for r in recs:
            timed = bool(r["avg_us"])
            lines.append(
                f"| `{r['name']}` | {r['calls']} | {r['avg_us']:.1f} | " if timed else
                f"| `{r['name']}` | {r['calls']} | - | ")
            lines[-1] += (f"{r['algorithmic_bytes'] / 1e6:.1f} | "
                          + (f"{r['achieved_GBs']:.0f} | {100 * r['frac']:.1f} % | " if timed
                             else "- | - | ")
                          + f"{mb(r, 'read_bytes')} | {num(r, 'read_over_algorithmic', '.3f')} | "
                          f"{mb(r, 'write_bytes')} | {num(r, 'traffic_over_algorithmic', '.3f')} |")
    c = sm["calibration"]
    if c["fetch_factor"] and c["write_factor"]:
        lines += ["", f"Counter calibration on the {c['copy_kib'] / 1024:.0f} MiB copy kernel: "
                  f"FETCH_SIZE x {c['fetch_factor']:.4f}, WRITE_SIZE x {c['write_factor']:.4f}."]
    lines += ["", "Commands (from the repository root):", ""]
    lines += [f"    {v}" for v in sm["commands"].values()]
    return "\n".join(lines) + "\n"


def merge(path):
    with open(path) as f:
        sm = json.load(f)
    try:
        with open(TRAFFIC) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}
    table["_about"] = ("HBM bytes per launch of the step kernels: rocprofv3 FETCH_SIZE and "
                       "WRITE_SIZE (KiB per dispatch, averaged over the dispatches of the "
                       "profiled bench.py run) times the counters' calibration on the plain copy "
                       "kernel; keyed workload -> precision -> segment height -> kernel; "
                       "written by python -m mepgl_amd.profiling merge")
    ent = table.setdefault(sm["workload"], {}).setdefault(sm["precision"], {})
    for k in [k for k in ent if not k.isdigit()]:
        del ent[k]
    rows = ent.setdefault(str(sm["segment_rows"]), {})
    for kid, recs in sm["kernels"].items():
        if len(recs) != 1 or "read_bytes" not in recs[0]:
            continue  # several classes: no single per-launch figure
        r = recs[0]
        rec = {"read": round(r["read_bytes"]), "algorithmic": r["algorithmic_bytes"],
               "read_over_algorithmic": round(r["read_over_algorithmic"], 4),
               "dispatches": r["dispatches"]}
        if "write_bytes" in r:
            rec.update({"bytes": round(r["bytes"]), "write": round(r["write_bytes"]),
                        "traffic_over_algorithmic": round(r["traffic_over_algorithmic"], 4)})
        if r["avg_us"]:
            rec["avg_us"] = round(r["avg_us"], 2)
        rows[kid] = rec
    rows["calibration"] = sm["calibration"]
    with open(TRAFFIC, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    print("merged", path, "into", TRAFFIC)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", default="vortex_entry_256")
    r.add_argument("--precision", default="mixed", choices=("mixed", "fp64"))
    r.add_argument("--rows", type=int, required=True)
    r.add_argument("--steps", type=int, default=100)
    r.add_argument("--warmup", type=int, default=10)
    r.add_argument("--timeout", type=int, default=240)
    r.add_argument("--passes", default="",
                   help="comma-separated subset of trace,fetch,write (default: all)")
    r.add_argument("--out", required=True)
    sm = sub.add_parser("summarize", help="recompute a summary from the outputs of `run`")
    for act in r._actions[1:]:
        sm._add_action(act)
    m = sub.add_parser("merge")
    m.add_argument("summary", nargs="+")
    a = ap.parse_args(argv)
    if a.cmd == "run":
        run(a)
    elif a.cmd == "summarize":
        summarize(a)
    else:
        for p in a.summary:
            merge(p)


if __name__ == "__main__":
    main()
