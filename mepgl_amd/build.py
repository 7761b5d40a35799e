"""Build libmepgl.so (HIP kernels for gfx950 + host context + C ABI) in-tree.

    python -m mepgl_amd.build            # library
    python -m mepgl_amd.build --cli      # library and the mepgl CLI
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmepgl.so")
CLI = os.path.join(HERE, "mepgl")
COPY_CALIB = os.path.join(HERE, "copy_calib")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for every HIP compile of this project.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]

# The step kernels are split by storage type and component count so that they
# compile concurrently.
SOURCES = ["step_f64_2.hip", "step_f64_1.hip", "step_f32_2.hip", "step_f32_1.hip", "kernels.hip",
           "context.cpp", "gauge_local.hip", "noalloc.hip"]
CLI_SOURCES = ["cli.cpp", "npyio.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_library(verbose=False) -> str:
    objs, procs = [], []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    # The translation units compile concurrently.
    for src in SOURCES:
        obj = os.path.join(HERE, "_build", src + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError("compile failed: " + " ".join(cmd[:3]) + " ...")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_copy_calib(verbose=False) -> str:
    """Plain copy kernel used to calibrate rocprofv3 byte counters."""
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "copy_calib.hip"), "-o", COPY_CALIB])
    if verbose:
        print("built", COPY_CALIB)
    return COPY_CALIB


def build_cli(verbose=False) -> str:
    srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES]
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", *srcs, "-o", CLI, "-L", HERE, "-lmepgl",
          "-Wl,-rpath,$ORIGIN", "-lpthread"])
    if verbose:
        print("built", CLI)
    return CLI


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cli", action="store_true")
    args = ap.parse_args(argv)
    build_library(verbose=True)
    build_copy_calib(verbose=True)
    if args.cli:
        build_cli(verbose=True)


if __name__ == "__main__":
    main()
