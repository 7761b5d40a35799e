"""Removes a range of images from a simulation's output (mepgl/cut.py:1-47).

    python -m mepgl_amd.cut -s SIMULATION -i A-B [--root DIR]

Images A .. B-1 of simulations/SIMULATION/output_data are removed and the
images after them renumbered, after a backup of output_data to
output_data_bak (replacing an older backup).  As in the reference, the
program exits with status 1 without touching anything when -s or -i is
missing.  Beyond the reference, the per-image arrays fenergy.npy, s.npy and
s_jb.npy are cut the same way, so output_data stays consistent for
post-processing and for a reload (-rc) of the shortened string.
"""
from __future__ import annotations

import argparse
import os
import shutil
import sys

import numpy as np

PER_IMAGE = ("fenergy", "s", "s_jb")


def cut(simulation: str, first: int, last: int, root: str = ".") -> int:
    """Removes images [first, last); returns the new image count."""
    sim = os.path.join(root, "simulations", simulation)
    out = os.path.join(sim, "output_data")
    if not os.path.isdir(out):
        raise FileNotFoundError(f"{out} does not exist")
    frames = sorted(int(d) for d in os.listdir(out)
                    if d.isdigit() and os.path.isdir(os.path.join(out, d)))
    F = max(frames) + 1 if frames else 0
    if not (0 <= first < last <= F):
        raise ValueError(f"interval {first}-{last} is not inside the {F} images")
    bak = os.path.join(sim, "output_data_bak")
    if os.path.exists(bak):
        shutil.rmtree(bak)
    shutil.copytree(out, bak)
    for n in range(first, last):
        shutil.rmtree(os.path.join(out, str(n)))
    for n in range(last, F):
        shutil.move(os.path.join(out, str(n)), os.path.join(out, str(n - (last - first))))
    for nm in PER_IMAGE:
        p = os.path.join(out, nm + ".npy")
        if os.path.exists(p):
            a = np.load(p, allow_pickle=False)
            if a.shape[:1] == (F,):
                np.save(p, np.concatenate([a[:first], a[last:]]))
    return F - (last - first)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Cut frames")
    ap.add_argument("-s", "--simulation", help="simulation to process")
    ap.add_argument("-i", "--interval", help="specify interval to cut")
    ap.add_argument("--root", default=".", help="directory holding simulations/")
    args = ap.parse_args(argv)
    if not args.simulation or not args.interval:
        return 1
    a, b = args.interval.split("-")
    first, last = int(a), int(b)
    print(f"Cut frames {first} to {last} in simulation {args.simulation}")
    cut(args.simulation, first, last, args.root)
    return 0


if __name__ == "__main__":
    sys.exit(main())
