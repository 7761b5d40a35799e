"""Runs the mepgl program over a sequence of stages, as the original launcher does.

``launch_simulation`` follows launcher_utils.py:151-191: stage k runs
``mepgl --simname S --mode modes[k] --Fin F[k] --Fout F[k+1]
--iterations iterations[k]`` and then makes the output the next input
(reload.py); the last stage keeps its image count.  The program is the
in-tree ``mepgl_amd/mepgl`` binary.
"""
from __future__ import annotations

import os
import subprocess
import sys
from typing import Sequence

from . import npyio

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "mepgl")


def executable() -> str:
    if not os.path.exists(EXE):
        raise FileNotFoundError(f"{EXE} is not built (python -m mepgl_amd.build --cli)")
    return EXE


def run_stage(simname: str, mode: str, F_in: int, F_out: int, iterations: int, root: str = ".",
              extra_args: Sequence[str] = (), stdin=subprocess.DEVNULL, capture: bool = False,
              timeout: float | None = None) -> subprocess.CompletedProcess:
    cmd = [executable(), "--simname", simname, "--mode", mode, "--Fin", str(F_in), "--Fout",
           str(F_out), "--iterations", str(iterations), *extra_args]
    r = subprocess.run(cmd, cwd=root, stdin=stdin, capture_output=capture, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        msg = r.stderr if capture else ""
        raise RuntimeError(f"mepgl exited with {r.returncode}: {' '.join(cmd)}\n{msg}")
    return r


def launch_simulation(simulation_name: str, F: Sequence[int], iterations: Sequence[int],
                      modes: Sequence[str], root: str = ".", extra_args: Sequence[str] = (),
                      capture: bool = False) -> list:
    """Stages of (mode, iterations) with image counts F[0] -> F[1] -> ...

    ``F`` has one entry more than the stages that resample; as in the
    original, the final stage runs with ``--Fin F[-1] --Fout F[-1]``.
    """
    P = len(F)
    if len(modes) < P or len(iterations) < P:
        raise ValueError("need one mode and one iteration count per entry of F")
    out = []
    for i in range(P - 1):
        out.append(run_stage(simulation_name, modes[i], F[i], F[i + 1], iterations[i], root,
                             extra_args, capture=capture))
        npyio.reload(simulation_name, root)
    out.append(run_stage(simulation_name, modes[-1], F[-1], F[-1], iterations[-1], root,
                         extra_args, capture=capture))
    return out


def main(argv=None) -> int:
    """``python -m mepgl_amd.launcher ARGS`` forwards ARGS to the program."""
    argv = sys.argv[1:] if argv is None else argv
    return subprocess.call([executable(), *argv])


if __name__ == "__main__":
    sys.exit(main())
