"""The C ABI library builds, loads without a GPU, and exports every entry
point include/mepgl.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "mepgl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mepgl_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = header_functions()
    assert "mepgl_create" in names and "mepgl_loop" in names
    assert len(names) >= 20


def test_library_exports_all_declared_symbols():
    from mepgl_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmepgl.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == header_functions()


def header_defines():
    text = open(os.path.join(ROOT, "include", "mepgl.h")).read()
    return {k: int(v) for k, v in
            re.findall(r"#define (MEPGL_[A-Z_0-9]+)\s+\(?(-?\d+)\)?u?\b", text)}


def test_python_constants_match_the_header():
    """The ctypes binding's constants are the header's: error codes,
    precisions, flags, scalar kinds and control keys."""
    from mepgl_amd import _lib

    d = header_defines()
    pairs = {"OK": "MEPGL_OK", "ERR_ARG": "MEPGL_ERR_ARG", "ERR_HIP": "MEPGL_ERR_HIP",
             "ERR_NONFINITE": "MEPGL_ERR_NONFINITE", "ERR_STATE": "MEPGL_ERR_STATE",
             "ERR_NOMEM": "MEPGL_ERR_NOMEM", "FP64": "MEPGL_FP64", "MIXED": "MEPGL_MIXED",
             "POLY_EXACT": "MEPGL_POLY_EXACT", "CUBIC_TEXTBOOK": "MEPGL_CUBIC_TEXTBOOK",
             "NO_GRAPH": "MEPGL_NO_GRAPH", "GAUGE_ALIGN": "MEPGL_GAUGE_ALIGN",
             "ROLLBACK": "MEPGL_ROLLBACK", "GAUGE_ALIGN_LOCAL": "MEPGL_GAUGE_ALIGN_LOCAL"}
    for py, c in pairs.items():
        assert getattr(_lib, py) == d[c], (py, c)
    controls = {k[len("MEPGL_CONTROL_"):]: v for k, v in d.items()
                if k.startswith("MEPGL_CONTROL_")}
    assert len(controls) >= 6
    for name, v in controls.items():
        assert getattr(_lib, "CONTROL_" + name) == v, name
    for name, v in _lib.SCALARS.items():
        assert d["MEPGL_SCALAR_" + name.upper()] == v, name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from mepgl_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libmepgl.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus above 1 exits non-zero with a message
    (no GPU is touched, so this runs on the CPU)."""
    import json
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


@pytest.mark.parametrize("N,ncomp,precision", [(8200, 1, 0), (6700, 2, 0), (9500, 2, 1),
                                               (11700, 1, 1)])
def test_create_rejects_images_above_2GiB(N, ncomp, precision):
    """An image's field set must fit the 32-bit byte offsets of the step
    kernels' buffer accesses: mepgl_create rejects larger strings with
    MEPGL_ERR_ARG before touching the device (so this runs without a GPU)."""
    import numpy as np

    from mepgl_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmepgl.so not built (run __graft_entry__.build())")
    elem = 4 if precision == 1 else 8
    assert (N + 1) * (2 + 2 * ncomp) * ((N + 63) // 64 * 64) * elem > 2 ** 31 - 1
    L = _lib.load()
    d = _lib.StringDesc()
    d.N, d.F, d.ncomp, d.dx = N, 2, ncomp, 0.1
    for name in ("a_1", "a_2"):
        getattr(d, name).value = -1.0
    for name in ("b_1", "m_xx_1", "m_yy_1", "b_2", "m_xx_2", "m_yy_2"):
        getattr(d, name).value = 1.0
    dom = np.ones(4, dtype=np.int32)  # never read: the size check comes first
    d.comp_domain = dom.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    d.sc_domain = dom.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    ctx = ctypes.c_void_p()
    rc = L.mepgl_create(ctypes.byref(d), 1, precision, 0, 0, ctypes.byref(ctx))
    assert rc == _lib.ERR_ARG
    assert not ctx.value
