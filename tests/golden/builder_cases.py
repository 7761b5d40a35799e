"""Cases of the StringBuilder vectors: small grids, seeded noise, vortex
paths with zero and non-zero windings, a disk domain, a non-square grid,
a phase difference and the three gauges.  `build` runs one case on a
StringBuilder class (the reference's when the vectors are written, the
project's in the test)."""
import numpy as np

CASES = [
    dict(F=3, Nx=12, Ny=12, x_lim=[-3.0, 3.0], y_lim=[-3.0, 3.0], multicomponent=False,
         B_z=0.6, seed=7, noise=0.1, psi_abs_1=1.0, psi_abs_2=1.0, disk=None, phase=0.0,
         v1=[[[4.0, 2.0, 0.0], [0.0, 0.0, 0.0], [-1, -1, -1]]], v2=[],
         gauges=["symmetric"]),
    dict(F=4, Nx=10, Ny=14, x_lim=[-2.0, 3.0], y_lim=[-3.0, 2.5], multicomponent=True,
         B_z=1.3, seed=11, noise=0.05, psi_abs_1=0.8, psi_abs_2=0.6, disk=2.2, phase=0.7,
         v1=[[[3.0, 2.0, 1.0, 0.0], [0.5, 0.4, 0.3, 0.2], [0, 1, 1, 2]]],
         v2=[[[-1.0, -0.5, 0.0, 0.5], [1.0, 1.0, 1.0, 1.0], [1, 1, 0, -1]],
             [[0.2, 0.2, 0.2, 0.2], [-1.5, -1.0, -0.5, 0.0], [-1, -1, -1, -1]]],
         gauges=["Landau_x", "Landau_y", "symmetric"]),
]


def build(StringBuilder, case):
    c = case
    b = StringBuilder(c["F"], Nx=c["Nx"], Ny=c["Ny"], x_lim=c["x_lim"], y_lim=c["y_lim"],
                      multicomponent=c["multicomponent"], B_z=c["B_z"])
    if c["disk"] is not None:
        R = c["disk"]
        b.set_sc_domain(lambda x, y: x ** 2 + y ** 2 < R ** 2)
    ph = c["phase"]
    if ph:
        b.add_phase_diff(lambda x, y: ph)
    for xv, yv, wv in c["v1"]:
        b.add_vortex(np.array(xv), np.array(yv), np.array(wv), 1)
    for xv, yv, wv in c["v2"]:
        b.add_vortex(np.array(xv), np.array(yv), np.array(wv), 2)
    comp, sc = b.get_domain_matrices()
    np.random.seed(c["seed"])
    u1, v1, u2, v2 = b.generate_matter_fields(psi_abs_1=c["psi_abs_1"],
                                              psi_abs_2=c["psi_abs_2"], noise=c["noise"])
    out = {"x": b.x, "y": b.y, "comp_domain": comp, "sc_domain": sc, "u_1": u1, "v_1": v1,
           "u_2": u2, "v_2": v2}
    for g in c["gauges"]:
        ax, ay = b.generate_vector_potential(gauge=g)
        out[f"ax_{g}"], out[f"ay_{g}"] = ax, ay
    return out
