"""The config.py / run.py workflow of the reference (mepgl/run.py:1-119) over
the mepgl program of this project.

    python -m mepgl_amd.run [-d] [-ni] [-rc] [--config config.py] [--root .]
                            [--precision fp64|mixed] [--no-post]

The config module is a user's config.py for the reference (config.py:17-132):
it defines simulation_name, the per-stage arrays F, iterations and modes,
default_relaxation_step_number, multicomponent, thin_film,
double_precision, dev_number, the grid (N, dx, x, y), the domains, the
Ginzburg-Landau parameter planes, q_1, q_2, eta, gamma, delta, and the
initial fields u_1, v_1, u_2, v_2, ax, ay as (F[0], N, N) arrays in the file
layout [i][j].  Its ``from mepgl_lib.builder import StringBuilder`` resolves
to mepgl_amd.builder, which builds the same arrays.

Steps, as run.py:
1. unless -ni or -rc: create simulations/<name>/, copy config.py (and
   batched_params.json when present) into it and write input_data
   (launcher_utils.py:195-266).  There is no header generation or compile
   step: sizes are runtime values of the program.  -d (debug build) is
   accepted and has nothing to do.
2. with -rc: output_data becomes input_data (run.py:107-112).
3. the stages: mepgl --mode modes[k] --Fin F[k] --Fout F[k+1]
   --iterations iterations[k], the output becoming the next input between
   stages (launcher_utils.py:151-191).
4. post-processing: the simulations/<name>/<name>.npz summary of post.py.

Differences, each deliberate: input files keep the config's precision
(float64 when double_precision, else float32; the reference writes float32
whatever double_precision says, run.py:51-79 vs launcher_utils.py:195-203);
double_precision selects the FP64 path, otherwise the mixed-precision path
(float fields, double reductions), which is more accurate than the
reference's all-float build; dx is written to input_data/dx.npy.
"""
from __future__ import annotations

import argparse
import importlib.util
import os
import shutil
import sys
import types

import numpy as np

from . import builder, launcher, npyio, observables

# Ginzburg-Landau parameter planes and the couplings of q.npy, in the order
# the reference writes them (launcher_utils.py:212-227).
PLANES = ("a_1", "b_1", "m_xx_1", "m_yy_1", "a_2", "b_2", "m_xx_2", "m_yy_2", "h")
COUPLINGS = ("q_1", "q_2", "eta", "gamma", "delta")


def _install_builder_alias():
    """Make ``from mepgl_lib.builder import StringBuilder`` import this
    project's builder (the reference's package is not needed)."""
    if "mepgl_lib.builder" in sys.modules:
        return
    pkg = types.ModuleType("mepgl_lib")
    pkg.__path__ = []  # a package, so that submodule imports resolve
    pkg.builder = builder
    sys.modules.setdefault("mepgl_lib", pkg)
    sys.modules["mepgl_lib.builder"] = builder


def load_config(path: str):
    """Import a config.py as a module (its directory first on sys.path)."""
    path = os.path.abspath(path)
    if not os.path.exists(path):
        raise FileNotFoundError(f"config module {path} not found")
    _install_builder_alias()
    sys.path.insert(0, os.path.dirname(path))
    try:
        spec = importlib.util.spec_from_file_location("config", path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
    finally:
        sys.path.pop(0)
    return mod


def _stage_arrays(cfg):
    F = [int(v) for v in np.atleast_1d(cfg.F)]
    it = [int(v) for v in np.atleast_1d(cfg.iterations)]
    modes = [str(v) for v in np.atleast_1d(cfg.modes)]
    if len(it) < len(F) or len(modes) < len(F):
        raise ValueError("config: iterations and modes need one entry per entry of F")
    return F, it, modes


def write_input(cfg, root: str = ".") -> str:
    """input_data of the config (launcher_utils.py:195-266), in the config's
    precision."""
    name = cfg.simulation_name
    dt = np.float64 if getattr(cfg, "double_precision", False) else np.float32
    F = int(np.atleast_1d(cfg.F)[0])
    d = os.path.join(npyio.sim_dir(name, root), "input_data")
    if os.path.exists(d):
        shutil.rmtree(d)
    os.makedirs(d)
    N = int(cfg.N)

    def plane(v):
        return np.broadcast_to(np.asarray(v, dtype=np.float64), (N, N))

    np.save(os.path.join(d, "x.npy"), np.asarray(cfg.x).astype(dt))
    np.save(os.path.join(d, "y.npy"), np.asarray(cfg.y).astype(dt))
    np.save(os.path.join(d, "dx.npy"), np.array([float(cfg.dx)], dtype=np.float64))
    np.save(os.path.join(d, "comp_domain.npy"), np.asarray(cfg.comp_domain).astype(np.int32))
    np.save(os.path.join(d, "sc_domain.npy"), np.asarray(cfg.sc_domain).astype(np.int32))
    for nm in PLANES:
        np.save(os.path.join(d, f"{nm}.npy"), plane(getattr(cfg, nm)).astype(dt))
    np.save(os.path.join(d, "q.npy"),
            np.array([float(np.mean(getattr(cfg, nm))) for nm in COUPLINGS]).astype(dt))
    mc = bool(cfg.multicomponent)
    fields = ("ax", "ay", "u_1", "v_1") + (("u_2", "v_2") if mc else ())
    for n in range(F):
        fd = os.path.join(d, str(n))
        os.makedirs(fd)
        for nm in fields:
            np.save(os.path.join(fd, f"{nm}.npy"), np.asarray(getattr(cfg, nm)[n]).astype(dt))
    return d


def program_args(cfg, precision: str | None = None) -> list:
    """mepgl options that carry what the reference compiles in
    (launcher_utils.py:14-124): relaxation steps, device, components,
    thin film, precision."""
    prec = precision or ("fp64" if getattr(cfg, "double_precision", False) else "mixed")
    args = ["--relaxation-steps", str(int(getattr(cfg, "default_relaxation_step_number", 25))),
            "--device", str(int(getattr(cfg, "dev_number", 0))),
            "--ncomp", "2" if cfg.multicomponent else "1", "--precision", prec,
            "--no-stdin"]
    if getattr(cfg, "thin_film", False):
        args.append("--no-self-field")
    return args


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Initialization.")
    ap.add_argument("-d", "--debug", help="debug build (nothing to build here)",
                    action="store_true")
    ap.add_argument("-ni", "--no-init", help="no initialization", action="store_true",
                    dest="no_init")
    ap.add_argument("-rc", "--reload-continue", help="reload and continue",
                    action="store_true", dest="rel_cont")
    ap.add_argument("--config", default="config.py", help="the config module (config.py)")
    ap.add_argument("--root", default=".", help="directory holding simulations/")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default=None,
                    help="override: fp64 or mixed (default from double_precision)")
    ap.add_argument("--no-post", action="store_true", help="skip the .npz summary")
    args = ap.parse_args(argv)

    cfg = load_config(args.config)
    name = cfg.simulation_name
    F, iterations, modes = _stage_arrays(cfg)
    sim = npyio.sim_dir(name, args.root)
    print(f"Simulation name: {name}")
    if not (args.no_init or args.rel_cont):
        os.makedirs(sim, exist_ok=True)
        shutil.copy(args.config, os.path.join(sim, "config.py"))
        bp = os.path.join(os.path.dirname(os.path.abspath(args.config)), "batched_params.json")
        if os.path.exists(bp):
            shutil.copy(bp, os.path.join(sim, "batched_params.json"))
        print("[*] Generating init files. ")
        print("Saving input data...    ", end="")
        write_input(cfg, args.root)
        print("done.")
        print("==> Initialization finished.")
        print("")
        print("[*] Grid size and precision are runtime values: nothing to compile.")
    if args.rel_cont:
        print("Reloading simulation data...   ", end="")
        npyio.reload(name, args.root)
        print("done.")
        print("")
    print("[*] Running simulation.", flush=True)
    launcher.launch_simulation(name, F, iterations, modes, root=args.root,
                               extra_args=program_args(cfg, args.precision))
    if not args.no_post:
        print("[*] Post processing.")
        path = observables.write_summary(name, args.root)
        e = np.load(path, allow_pickle=True)["fenergy"]
        b_n, b_e = observables.barriers(e)
        print(f"Summary written to {path}; barriers dF_n = {b_n:.8g}, dF_e = {b_e:.8g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
