"""Host mirror of the reference solver interface over libmepgl.

The methods keep the reference's names and argument meaning
(mepgl/src/nlcg.cuh:14-54, interpolation.cuh:6-24, loop() in
mepgl/src/mepgl.cu:38-391), but act on a context that owns the device
memory of one or many strings.  Where the reference exits on a CUDA error,
these raise MepglError.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _lib
from .problem import GLString

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Solver:
    """A set of strings resident on one GPU.

    precision: "fp64" (the reference's DOUBLE_PRECISION arithmetic) or
    "mixed" (float fields and stencils, double reductions and scalars).
    poly_exact / cubic_textbook switch off the reference's line-search
    quartic and spline-weight quirks.  rollback: loop() restores the state
    from the start of an outer iteration that produced a NaN or Inf.
    """

    def __init__(self, strings: Sequence[GLString], precision: str = "fp64",
                 poly_exact: bool = False, cubic_textbook: bool = False,
                 use_graph: bool = True, device: int = 0, segment_rows: int = 0,
                 gauge_align: bool = False, rollback: bool = False,
                 gauge_align_local: bool = False):
        self._L = _lib.load()
        self.strings: List[GLString] = [s.validate() for s in strings]
        self.precision = precision
        prec = {"fp64": _lib.FP64, "mixed": _lib.MIXED}[precision]
        flags = (_lib.POLY_EXACT if poly_exact else 0) | \
                (_lib.CUBIC_TEXTBOOK if cubic_textbook else 0) | \
                (0 if use_graph else _lib.NO_GRAPH) | \
                (_lib.GAUGE_ALIGN if gauge_align else 0) | \
                (_lib.ROLLBACK if rollback else 0) | \
                (_lib.GAUGE_ALIGN_LOCAL if gauge_align_local else 0)
        self._options = dict(poly_exact=poly_exact, cubic_textbook=cubic_textbook,
                             use_graph=use_graph, device=device, gauge_align=gauge_align,
                             gauge_align_local=gauge_align_local)
        self._keep = []
        descs = (_lib.StringDesc * len(self.strings))()
        for d, s in zip(descs, self.strings):
            self._fill_desc(d, s, segment_rows)
        self._ctx = C.c_void_p()
        rc = self._L.mepgl_create(descs, len(self.strings), prec, flags, device,
                                  C.byref(self._ctx))
        if rc != 0:
            msg = self._L.mepgl_last_error(self._ctx).decode() if self._ctx else "create failed"
            if self._ctx:
                self._L.mepgl_destroy(self._ctx)
                self._ctx = C.c_void_p()
            raise _lib.MepglError(rc, msg)
        self._keep = []
        self.F = [s.F for s in self.strings]
        for i, s in enumerate(self.strings):
            self.upload(i, s)

    # -- construction helpers ------------------------------------------------
    def _fill_desc(self, d, s: GLString, rows):
        N = s.N
        d.N, d.F, d.ncomp = N, s.F, s.ncomp
        d.no_self_field = int(bool(s.no_self_field))
        d.dx = float(s.dx)
        for name in ("a_1", "b_1", "m_xx_1", "m_yy_1", "a_2", "b_2", "m_xx_2", "m_yy_2", "h"):
            v = getattr(s, name)
            p = getattr(d, name)
            if np.isscalar(v):
                p.value, p.plane = float(v), None
            else:
                arr = np.ascontiguousarray(v, dtype=np.float64)
                self._keep.append(arr)
                p.value, p.plane = 0.0, _ptr(arr)
        for name in ("q_1", "q_2", "eta", "gamma", "delta"):
            setattr(d, name, float(getattr(s, name)))
        for name in ("comp_domain", "sc_domain"):
            arr = np.ascontiguousarray(getattr(s, name), dtype=np.int32)
            self._keep.append(arr)
            setattr(d, name, arr.ctypes.data_as(C.POINTER(C.c_int32)))
        d.segment_rows = int(rows)

    def _check(self, rc):
        if rc != 0:
            raise _lib.MepglError(rc, self._L.mepgl_last_error(self._ctx).decode())

    def close(self):
        if self._ctx:
            self._L.mepgl_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- fields ----------------------------------------------------------------
    def upload(self, i: int, s: GLString):
        a = np.ascontiguousarray(s.a, dtype=np.float64)
        p1 = np.ascontiguousarray(s.psi_1, dtype=np.float64)
        p2 = None if s.psi_2 is None or not s.multicomponent else \
            np.ascontiguousarray(s.psi_2, dtype=np.float64)
        self._check(self._L.mepgl_upload(self._ctx, i, _ptr(a), _ptr(p1), _ptr(p2)))

    def _shape(self, i):
        N = self.strings[i].N
        return (self.F[i], N, N, 2)

    def download(self, i: int):
        """Current fields (a, psi_1, psi_2) of string i, float64 (F, N, N, 2)."""
        a, p1, p2 = (np.empty(self._shape(i)) for _ in range(3))
        self._check(self._L.mepgl_download(self._ctx, i, _ptr(a), _ptr(p1), _ptr(p2)))
        return a, p1, p2

    def download_state(self, i: int, which: str):
        """Solver buffers: "g" (caller's g), "g_old", "d" (direction)."""
        k = {"g": 0, "g_old": 1, "d": 2}[which]
        a, p1, p2 = (np.empty(self._shape(i)) for _ in range(3))
        self._check(self._L.mepgl_download_state(self._ctx, i, k, _ptr(a), _ptr(p1), _ptr(p2)))
        return a, p1, p2

    def upload_state(self, i: int, which: str, a, p1, p2=None):
        """Restores a solver buffer ("g", "g_old" or "d") of string i."""
        k = {"g": 0, "g_old": 1, "d": 2}[which]
        a = np.ascontiguousarray(a, dtype=np.float64)
        p1 = np.ascontiguousarray(p1, dtype=np.float64)
        p2 = None if p2 is None else np.ascontiguousarray(p2, dtype=np.float64)
        self._check(self._L.mepgl_upload_state(self._ctx, i, k, _ptr(a), _ptr(p1), _ptr(p2)))

    def set_scalars(self, i: int, which: str, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (self.F[i],):
            raise ValueError(f"{which} needs {self.F[i]} values")
        self._check(self._L.mepgl_set_scalars(self._ctx, i, _lib.SCALARS[which], _ptr(v)))

    _STATE_SCALARS = ("g_norm2", "g_norm2_old", "s", "s_jb")

    def checkpoint(self) -> dict:
        """The complete relaxation state: fields, both gradient buffers, the
        direction and the per-image scalars they pair with, per string.
        restore() of it continues the relaxation bit for bit."""
        out = {}
        for i, s in enumerate(self.strings):
            mc = s.multicomponent
            for key, arrs in (("x", self.download(i)), ("g", self.download_state(i, "g")),
                              ("g_old", self.download_state(i, "g_old")),
                              ("d", self.download_state(i, "d"))):
                out[f"s{i}_{key}_a"], out[f"s{i}_{key}_p1"] = arrs[0], arrs[1]
                if mc:
                    out[f"s{i}_{key}_p2"] = arrs[2]
            for nm in self._STATE_SCALARS:
                out[f"s{i}_{nm}"] = self.scalars(i, nm)
        return out

    def restore(self, ckpt: dict):
        for i, s in enumerate(self.strings):
            mc = s.multicomponent

            def get(key):
                return (ckpt[f"s{i}_{key}_a"], ckpt[f"s{i}_{key}_p1"],
                        ckpt[f"s{i}_{key}_p2"] if mc else None)

            a, p1, p2 = get("x")
            if a.shape != self._shape(i):
                raise ValueError(f"checkpoint of string {i} has shape {a.shape}, "
                                 f"expected {self._shape(i)}")
            self._check(self._L.mepgl_upload(self._ctx, i, _ptr(np.ascontiguousarray(a)),
                                             _ptr(np.ascontiguousarray(p1)),
                                             _ptr(None if p2 is None else np.ascontiguousarray(p2))))
            for key in ("g", "g_old", "d"):
                self.upload_state(i, key, *get(key))
            for nm in self._STATE_SCALARS:
                self.set_scalars(i, nm, ckpt[f"s{i}_{nm}"])

    def save_checkpoint(self, path: str):
        np.savez(path, **self.checkpoint())

    def load_checkpoint(self, path: str):
        with np.load(path, allow_pickle=False) as z:
            self.restore({k: z[k] for k in z.files})

    def scalars(self, i: int, which: str) -> np.ndarray:
        n = self.F[i] * (4 if which == "poly" else 1)
        out = np.empty(n)
        self._check(self._L.mepgl_get_scalars(self._ctx, i, _lib.SCALARS[which], _ptr(out)))
        return out.reshape(self.F[i], 4) if which == "poly" else out

    # -- reference solver interface -------------------------------------------
    def computeFreeEnergy(self) -> List[np.ndarray]:
        """computeFreeEnergy (nlcg.cu:217-240): per-image energies of every string."""
        self._check(self._L.mepgl_free_energy(self._ctx))
        return [self.scalars(i, "energy") for i in range(len(self.strings))]

    def winding(self):
        """Boundary phase windings (compute_winding, gl_utils.py:10-41) of
        every image: per string (w_1, w_2), w_2 None for one component."""
        self._check(self._L.mepgl_winding(self._ctx))
        return [(self.scalars(i, "winding_1"),
                 self.scalars(i, "winding_2") if s.multicomponent else None)
                for i, s in enumerate(self.strings)]

    def gradient(self, i: int):
        """gradientPsi + gradientA (nlcg.cu:504-888) at the current fields."""
        a, p1, p2 = (np.empty(self._shape(i)) for _ in range(3))
        self._check(self._L.mepgl_gradient(self._ctx, i, _ptr(a), _ptr(p1), _ptr(p2)))
        return a, p1, p2

    def computeJB(self, i: int):
        """computeJB (nlcg.cu:367-380): b (F, N, N), j_1 and j_2 (F, N, N, 2)."""
        N = self.strings[i].N
        b = np.empty((self.F[i], N, N))
        j1, j2 = np.empty(self._shape(i)), np.empty(self._shape(i))
        self._check(self._L.mepgl_compute_jb(self._ctx, i, _ptr(b), _ptr(j1), _ptr(j2)))
        return b, j1, j2

    def nlcgSteps(self, iter_n: int, nlcg: bool = True, maxwell_solver: bool = False):
        """nlcgSteps (nlcg.cu:1462-1586) on every string."""
        self._check(self._L.mepgl_nlcg_steps(self._ctx, int(iter_n), int(bool(nlcg)),
                                             int(bool(maxwell_solver))))

    def arcLength(self):
        """arcLength + computeJB + arcLengthJB (mepgl.cu:193-201); returns
        (s, s_jb) per string."""
        self._check(self._L.mepgl_arc_length(self._ctx))
        return [(self.scalars(i, "s"), self.scalars(i, "s_jb")) for i in range(len(self.strings))]

    def redistributeLinear(self):
        """redistributeLinear (interpolation.cu:229-250) on the last s_jb."""
        self._check(self._L.mepgl_redistribute(self._ctx, 0))

    def redistributeCubic(self):
        """redistributeCubic (interpolation.cu:404-481) on the last s_jb."""
        self._check(self._L.mepgl_redistribute(self._ctx, 1))

    def resample(self, i: int, F_new: int):
        """End-of-run resampling to F_new images (mepgl.cu:699-766)."""
        self._check(self._L.mepgl_resample(self._ctx, i, int(F_new)))
        self.F[i] = int(F_new)

    def gauge_align(self):
        """Constant-phase alignment of neighbouring images; returns the angles
        applied to each string's images."""
        self._check(self._L.mepgl_gauge_align(self._ctx))
        return [self.scalars(i, "theta") for i in range(len(self.strings))]

    def gauge_align_local(self):
        """Local gauge alignment of neighbouring images (fields, and the psi
        parts of the direction and gradients); strings with a fixed A are
        left unchanged."""
        self._check(self._L.mepgl_gauge_align_local(self._ctx))

    def reset_interior_beta(self):
        """mepgl.cu:231-234."""
        self._check(self._L.mepgl_reset_interior_beta(self._ctx))

    def reset_solver(self):
        """Solver state at the start of loop() (mepgl.cu:124-140)."""
        self._check(self._L.mepgl_reset_solver(self._ctx))

    def loop(self, iterations: int, relaxation_steps: int, mode: str = "L", nlcg: bool = True):
        """loop() (mepgl.cu:148-306) without its printing."""
        self._check(self._L.mepgl_loop(self._ctx, int(iterations), int(relaxation_steps),
                                       mode.encode()[:1], int(bool(nlcg))))

    def fp64_check(self, iterations: int = 1, relaxation_steps: int = 25, mode: str = "L",
                   nlcg: bool = True) -> List[dict]:
        """Error of the strings' current state against the FP64 path, per
        string (observables.precision_error): the state is relaxed further in
        FP64 for `iterations` outer iterations of loop() and the energies and
        barriers are compared with this solver's.  On a relaxed string the
        difference is what float arithmetic left in it; on one still relaxing
        it also holds that further relaxation, an upper bound.  The mepgl
        program states the same after every mixed-precision run."""
        from . import observables

        e_mixed = self.computeFreeEnergy()
        strings = []
        for i, s in enumerate(self.strings):
            t = s.copy()
            t.F = self.F[i]
            t.a, t.psi_1, p2 = self.download(i)
            t.psi_2 = p2 if s.multicomponent else None
            strings.append(t)
        with Solver(strings, "fp64", **self._options) as ref:
            ref.loop(iterations, relaxation_steps, mode, nlcg)
            e64 = ref.computeFreeEnergy()
        return [observables.precision_error(m, d) for m, d in zip(e_mixed, e64)]

    # -- control, timing -------------------------------------------------------
    def set_control(self, key: int, value: int):
        self._check(self._L.mepgl_set_control(self._ctx, key, int(value)))

    def get_control(self, key: int) -> int:
        v = C.c_int(0)
        self._check(self._L.mepgl_get_control(self._ctx, key, C.byref(v)))
        return v.value

    def synchronize(self):
        self._check(self._L.mepgl_synchronize(self._ctx))

    def profile(self, enable: bool):
        self._check(self._L.mepgl_profile(self._ctx, int(bool(enable))))

    def profile_read(self, kernel: int):
        ms, cnt = C.c_double(), C.c_int()
        self._check(self._L.mepgl_profile_read(self._ctx, kernel, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def set_segment_rows(self, i: int, rows: int):
        self._check(self._L.mepgl_set_segment_rows(self._ctx, i, int(rows)))

    def autotune(self, steps: int = 5):
        """Pick each string class's segment height by timing five rounds per
        candidate of an outer iteration of mode L with `steps` NLCG steps
        (each round visits every candidate in turn; the fastest round counts);
        the solver state is left unchanged."""
        rows = (C.c_int * len(self.strings))()
        self._check(self._L.mepgl_autotune(self._ctx, int(steps), rows))
        return list(rows)

    def info(self, i: int = 0):
        rows, items, nbytes = C.c_int(), C.c_int(), C.c_longlong()
        self._check(self._L.mepgl_info(self._ctx, i, C.byref(rows), C.byref(items),
                                       C.byref(nbytes)))
        return {"tile_rows": rows.value, "work_items": items.value, "bytes": nbytes.value}
