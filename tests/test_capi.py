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
