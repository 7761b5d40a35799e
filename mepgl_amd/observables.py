"""Observables of a relaxed string: the summary the reference's
post-processing writes (mepgl/post.py:125-152) and the energy barriers it
prints (post.py:188-189).

    python -m mepgl_amd.observables --simulation NAME [--root DIR]

reads ``simulations/NAME/output_data`` and writes ``simulations/NAME/NAME.npz``
with the reference's keys: ``fenergy``, ``s_gd`` (s.npy), ``s_ph``
(s_jb.npy), ``w_1``, ``w_2`` (boundary phase windings, None for one
component) and ``phi`` (flux sum(b) dx^2).
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np


def barriers(fenergy):
    """Energy barriers of a string, post.py:188-189: the entry barrier
    max(F - F_0) and the exit barrier max(F - F_last)."""
    e = np.asarray(fenergy, dtype=np.float64)
    return float(np.max(e - e[0])), float(np.max(e - e[-1]))


def precision_error(fenergy_mixed, fenergy_fp64):
    """Error of a mixed-precision string against the FP64 path: the
    per-image relative energy error and the relative errors of the entry and
    exit barriers (absolute where the FP64 value is zero)."""
    em = np.asarray(fenergy_mixed, dtype=np.float64)
    ed = np.asarray(fenergy_fp64, dtype=np.float64)

    def rel(x, ref):
        ref = np.asarray(ref, dtype=np.float64)
        return np.abs(np.asarray(x) - ref) / np.where(ref != 0.0, np.abs(ref), 1.0)

    bm, bd = barriers(em), barriers(ed)
    return {"fenergy_mixed": em, "fenergy_fp64": ed, "fenergy_rel_err": rel(em, ed),
            "barriers_mixed": np.array(bm), "barriers_fp64": np.array(bd),
            "barrier_rel_err": rel(bm, bd)}


def winding(u, v):
    """Phase winding around the grid boundary per image (gl_utils.py:10-41).

    `u`, `v` are (F, N, N) arrays in the file layout [n][i][j] (i along x).
    The boundary is walked along j = 0 with i increasing, then i = N-1 with
    j increasing, then j = N-1 with i decreasing, then i = 0 with j
    decreasing; each edge adds (u dv - v du) / |psi|^2 taken at its start.
    The sum is 2 pi times the winding number for a smooth phase (it is not
    divided by 2 pi, as in the reference).
    """
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)

    def edges(a):
        # Values along the closed boundary loop, in walking order.
        return np.concatenate([a[:, :-1, 0], a[:, -1, :-1], a[:, :0:-1, -1], a[:, 0, :0:-1]],
                              axis=1)

    bu, bv = edges(u), edges(v)
    nu = np.roll(bu, -1, axis=1)
    nv = np.roll(bv, -1, axis=1)
    return np.sum((bu * (nv - bv) - bv * (nu - bu)) / (bu * bu + bv * bv), axis=1)


def flux(b, dx):
    """Magnetic flux per image, post.py:134-136: sum(b) dx^2."""
    b = np.asarray(b, dtype=np.float64)
    return np.sum(b.reshape(b.shape[0], -1), axis=1) * dx * dx


def summary(output_dir):
    """The post.py summary of one output_data directory, as a dict."""
    from .npyio import read_dir

    s, extra = read_dir(output_dir)
    # File layout [n][i][j]: the transpose of the in-memory [n][j][i].
    u1 = np.swapaxes(s.psi_1[..., 0], 1, 2)
    v1 = np.swapaxes(s.psi_1[..., 1], 1, 2)
    x = np.load(os.path.join(output_dir, "x.npy"), allow_pickle=False).astype(np.float64)
    dx = float(x[1][0] - x[0][0])
    out = {
        "fenergy": extra.get("fenergy"),
        "s_gd": extra.get("s"),
        "s_ph": extra.get("s_jb"),
        "w_1": winding(u1, v1),
        "w_2": None,
        "phi": flux(extra["b"], dx) if "b" in extra else None,
    }
    if s.multicomponent:
        u2 = np.swapaxes(s.psi_2[..., 0], 1, 2)
        v2 = np.swapaxes(s.psi_2[..., 1], 1, 2)
        out["w_2"] = winding(u2, v2)
    # A mixed-precision run of the mepgl program writes its energies and the
    # FP64 path's as the columns of fp64_check.npy.
    chk = os.path.join(output_dir, "fp64_check.npy")
    if os.path.exists(chk):
        e = np.load(chk, allow_pickle=False)
        out.update(precision_error(e[:, 0], e[:, 1]))
    return out


def write_summary(simulation_name, root="."):
    """Writes simulations/<name>/<name>.npz (post.py:146-151)."""
    d = os.path.join(root, "simulations", simulation_name)
    out = summary(os.path.join(d, "output_data"))
    path = os.path.join(d, f"{simulation_name}.npz")
    np.savez(path, **out)
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description="Summary of a relaxed string.")
    ap.add_argument("-s", "--simulation", required=True)
    ap.add_argument("--root", default=".")
    args = ap.parse_args(argv)
    path = write_summary(args.simulation, args.root)
    out = summary(os.path.join(args.root, "simulations", args.simulation, "output_data"))
    e = out["fenergy"]
    if e is not None:
        dn, de = barriers(e)
        print(f"Delta F_n = {dn:.6}   Delta F_e = {de:.6}")
    if "fenergy_rel_err" in out:
        b = out["barrier_rel_err"]
        print(f"mixed vs FP64: energy rel. error <= {out['fenergy_rel_err'].max():.3e}   "
              f"barrier rel. errors {b[0]:.3e} (entry) {b[1]:.3e} (exit)")
    print(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
