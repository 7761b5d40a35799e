"""The mepgl program's arguments and input handling (no GPU: --dry-run and
error paths), and the Python side of the on-disk layout.

The reference program takes --simname --mode --Fin --Fout --iterations
(mepgl.cu:394-418) and reads simulations/<name>/input_data (io.cu:10-126),
reinterpreting the bytes of every file as its compile-time `real`.  Here
float32 and float64 files are both accepted and converted, and a float
domain mask is an error.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import small_string

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mepgl_amd", "mepgl")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        from mepgl_amd import build as B

        if not os.path.exists(B.LIB):
            B.build_library()
        B.build_cli()
    return EXE


def run(exe, *args, cwd=None):
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, cwd=cwd,
                          stdin=subprocess.DEVNULL, timeout=60)


def dry(exe, root, name="t", F=3, *extra):
    r = run(exe, "--simname", name, "--mode", "L", "--Fin", F, "--iterations", 1, "--dry-run",
            "--root", root, *extra)
    info = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(" ")
        info[k] = v
    return r, info


def test_help(exe):
    r = run(exe, "--help")
    assert r.returncode == 0
    for flag in ("--simname", "--mode", "--Fin", "--Fout", "--iterations"):
        assert flag in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--mode", "L", "--Fin", "3", "--iterations", "1"], "--simname"),
    (["--simname", "x", "--mode", "Q", "--Fin", "3", "--iterations", "1"], "--mode"),
    (["--simname", "x", "--mode", "L", "--iterations", "1"], "--Fin"),
    (["--simname", "x", "--mode", "L", "--Fin", "3"], "--iterations"),
    (["--simname", "x", "--mode", "L", "--Fin", "three", "--iterations", "1"], "integer"),
    (["--simname", "x", "--mode", "L", "--Fin", "3", "--iterations", "1", "--bogus"], "bogus"),
    (["--simname", "x", "--mode", "L", "--Fin", "3", "--iterations", "1", "--precision", "fp16"],
     "precision"),
    (["--simname", "x", "--mode", "L", "--Fin", "3", "--iterations", "1", "--fp64-check", "all"],
     "--fp64-check"),
    (["--simname", "x", "--mode", "L", "--Fin", "3", "--iterations", "1",
      "--fp64-check-iterations", "0"], "--fp64-check-iterations"),
])
def test_bad_arguments(exe, args, msg):
    r = run(exe, *args)
    assert r.returncode != 0
    assert msg in r.stderr


@pytest.mark.parametrize("double_precision", [True, False])
@pytest.mark.parametrize("mc", [False, True])
def test_dry_run_reads_layout(exe, tmp_path, double_precision, mc):
    from mepgl_amd import npyio

    s = small_string(N=12, F=3, multicomponent=mc, seed=2, planes=True)
    npyio.write_input(s, "t", str(tmp_path), double_precision=double_precision)
    r, info = dry(exe, str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert int(info["N"]) == 12 and int(info["F"]) == 3
    assert int(info["ncomp"]) == (2 if mc else 1)
    assert float(info["dx"]) == s.dx  # dx.npy is float64 whatever the field dtype
    q = [float(v) for v in info["q"].split()]
    want = [s.q_1, s.q_2, s.eta, s.gamma, s.delta]
    if double_precision:
        assert q == want
    else:
        assert q == [float(np.float32(v)) for v in want]
    # a_1, m_xx_1 and h are planes; the other six parameters are uniform
    assert int(info["uniform_planes"]) == 6


def test_dx_from_mesh_when_no_dx_file(exe, tmp_path):
    from mepgl_amd import npyio

    s = small_string(N=10, F=2, seed=3)
    d = npyio.write_input(s, "t", str(tmp_path))
    os.remove(os.path.join(d, "dx.npy"))
    x = np.load(os.path.join(d, "x.npy"))
    r, info = dry(exe, str(tmp_path), "t", 2)
    assert r.returncode == 0, r.stderr
    assert float(info["dx"]) == float(x[1, 0] - x[0, 0])
    r, info = dry(exe, str(tmp_path), "t", 2, "--dx", "0.125")
    assert float(info["dx"]) == 0.125


def test_int64_and_bool_masks_accepted(exe, tmp_path):
    from mepgl_amd import npyio

    s = small_string(N=10, F=2, seed=4)
    d = npyio.write_input(s, "t", str(tmp_path))
    np.save(os.path.join(d, "comp_domain.npy"), np.asarray(s.comp_domain).T.astype(np.int64))
    np.save(os.path.join(d, "sc_domain.npy"), np.asarray(s.sc_domain).T.astype(bool))
    r, _ = dry(exe, str(tmp_path), "t", 2)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("damage", ["float_mask", "wrong_shape", "missing_frame", "short_q",
                                    "not_npy"])
def test_invalid_input_rejected(exe, tmp_path, damage):
    from mepgl_amd import npyio

    s = small_string(N=10, F=2, seed=5)
    d = npyio.write_input(s, "t", str(tmp_path))
    if damage == "float_mask":
        np.save(os.path.join(d, "sc_domain.npy"), np.ones((10, 10)))
    elif damage == "wrong_shape":
        np.save(os.path.join(d, "h.npy"), np.zeros((11, 11)))
    elif damage == "missing_frame":
        os.remove(os.path.join(d, "1", "u_1.npy"))
    elif damage == "short_q":
        np.save(os.path.join(d, "q.npy"), np.zeros(3))
    elif damage == "not_npy":
        with open(os.path.join(d, "b_1.npy"), "w") as f:
            f.write("not an array")
    r, _ = dry(exe, str(tmp_path), "t", 2)
    assert r.returncode != 0
    assert "mepgl:" in r.stderr


def test_missing_simulation(exe, tmp_path):
    r, _ = dry(exe, str(tmp_path), "nothing", 2)
    assert r.returncode != 0 and "not found" in r.stderr


@pytest.mark.parametrize("mc", [False, True])
def test_python_layout_round_trip(tmp_path, mc):
    from mepgl_amd import npyio

    s = small_string(N=9, F=3, multicomponent=mc, seed=6, planes=True)
    d = npyio.write_input(s, "t", str(tmp_path))
    t, extra = npyio.read_dir(d)
    assert extra == {}
    assert (t.N, t.F, t.multicomponent, t.dx) == (s.N, s.F, s.multicomponent, s.dx)
    np.testing.assert_array_equal(t.a, s.a)
    np.testing.assert_array_equal(t.psi_1, s.psi_1)
    if mc:
        np.testing.assert_array_equal(t.psi_2, s.psi_2)
    np.testing.assert_array_equal(t.comp_domain, s.comp_domain)
    np.testing.assert_array_equal(t.a_1, s.a_1)
    np.testing.assert_array_equal(t.h, s.h)
    assert t.q_1 == s.q_1 and t.delta == s.delta
    # The file arrays are the transpose of the in-memory planes (io.cu:23-30).
    u = np.load(os.path.join(d, "0", "u_1.npy"))
    np.testing.assert_array_equal(u, s.psi_1[0, :, :, 0].T)


def test_reload_copies_output_to_input(tmp_path):
    from mepgl_amd import npyio

    s = small_string(N=8, F=2, seed=7)
    d = npyio.write_input(s, "t", str(tmp_path))
    out = os.path.join(os.path.dirname(d), "output_data")
    os.makedirs(os.path.join(out, "0"))
    np.save(os.path.join(out, "marker.npy"), np.arange(3))
    npyio.reload("t", str(tmp_path))
    assert os.path.exists(os.path.join(d, "marker.npy"))
    assert not os.path.exists(os.path.join(d, "1"))
