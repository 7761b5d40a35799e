"""The mepgl program end to end on the GPU against the oracle: the files it
writes after a run of the relaxation loop (io.cu:128-291), energy.csv
(mepgl.cu:348-355), the resampling to --Fout images (mepgl.cu:699-766),
the stdin stop key (mepgl.cu:647-651) and the staged launcher
(launcher_utils.py:151-191).

Tolerance: fields per-image norm-wise relative 1e-10 (FP64), energies
relative 1e-11 after --iterations outer iterations.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import normwise, rel_err, small_string

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mepgl_amd", "mepgl")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cli(root, *args, stdin=subprocess.DEVNULL, timeout=120):
    if not os.path.exists(EXE):
        raise FileNotFoundError(f"{EXE} is not built")
    r = subprocess.run([EXE, *map(str, args)], cwd=root, capture_output=True, text=True,
                       stdin=stdin, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    return r


@pytest.mark.parametrize("mode", ["L", "C", "S"])
def test_cli_run_matches_oracle(tmp_path, oracle, mode):
    from mepgl_amd import npyio

    s = small_string(N=20, F=6, seed=21, multicomponent=True)
    say(f"[gpu] cli mode={mode} N=20 F=6")
    npyio.write_input(s, "t", str(tmp_path))
    r = cli(str(tmp_path), "--simname", "t", "--mode", mode, "--Fin", 6, "--Fout", 6,
            "--iterations", 3, "--relaxation-steps", 4)
    assert r.stdout.count("Iteration:") == 3
    assert "Saving results...  done." in r.stdout
    out, extra = npyio.read_dir(os.path.join(tmp_path, "simulations", "t", "output_data"))

    ref = oracle.OracleString(s)
    ref.loop(3, 4, mode)
    assert normwise(out.psi_1, ref.psi_1) < 1e-10
    assert normwise(out.psi_2, ref.psi_2) < 1e-10
    assert normwise(out.a, ref.a) < 1e-10
    assert rel_err(extra["fenergy"], ref.free_energy()) < 1e-11
    b, j1, j2 = ref.compute_jb()
    assert normwise(extra["b"], b) < 1e-10
    assert normwise(extra["j_1"], j1) < 1e-10
    assert normwise(extra["j_2"], j2) < 1e-10
    if mode != "S":
        # s and s_jb are those of the last reparametrisation
        assert rel_err(extra["s_jb"], ref.s_jb) < 1e-12
        assert rel_err(extra["s"], ref.s) < 1e-12
    # The post-processing summary of the same output (post.py:125-152).
    from mepgl_amd import observables

    z = np.load(observables.write_summary("t", str(tmp_path)), allow_pickle=True)
    np.testing.assert_array_equal(z["fenergy"], extra["fenergy"])
    np.testing.assert_allclose(z["phi"], extra["b"].reshape(6, -1).sum(axis=1) * s.dx ** 2,
                               rtol=1e-12)
    # energy.csv holds the last iteration's line-search energies
    with open(os.path.join(tmp_path, "energy.csv")) as f:
        csv = np.array([float(v) for v in f.read().split(",")])
    assert csv.shape == (6,)
    np.testing.assert_allclose(csv, ref.fenergy, rtol=1e-5)


def test_cli_resample_to_fout(tmp_path, oracle):
    from mepgl_amd import npyio

    s = small_string(N=20, F=5, seed=22, multicomponent=False)
    say("[gpu] cli resample 5 -> 9")
    npyio.write_input(s, "t", str(tmp_path))
    cli(str(tmp_path), "--simname", "t", "--mode", "L", "--Fin", 5, "--Fout", 9,
        "--iterations", 2, "--relaxation-steps", 3)
    out, extra = npyio.read_dir(os.path.join(tmp_path, "simulations", "t", "output_data"))
    assert out.F == 9 and extra["fenergy"].shape == (9,)
    ref = oracle.OracleString(s)
    ref.loop(2, 3, "L")
    ref.arc_lengths()
    ref.redistribute(F_new=9)
    assert normwise(out.psi_1, ref.psi_1) < 1e-10
    assert normwise(out.a, ref.a) < 1e-10
    assert rel_err(extra["fenergy"], ref.free_energy()) < 1e-11


def test_cli_mixed_and_float32_input(tmp_path, oracle):
    from mepgl_amd import npyio

    s = small_string(N=20, F=5, seed=23, multicomponent=True)
    say("[gpu] cli mixed precision, float32 files")
    npyio.write_input(s, "t", str(tmp_path), double_precision=False)
    cli(str(tmp_path), "--simname", "t", "--mode", "L", "--Fin", 5, "--iterations", 2,
        "--relaxation-steps", 3, "--precision", "mixed")
    d = os.path.join(tmp_path, "simulations", "t", "output_data")
    assert np.load(os.path.join(d, "0", "u_1.npy")).dtype == np.float32
    out, extra = npyio.read_dir(d)
    s32, _ = npyio.read_dir(os.path.join(tmp_path, "simulations", "t", "input_data"))
    ref = oracle.OracleString(s32)
    ref.loop(2, 3, "L")
    assert rel_err(extra["fenergy"], ref.free_energy()) < 1e-4


@pytest.mark.parametrize("check", ["continue", "shadow"])
def test_cli_mixed_fp64_check(tmp_path, oracle, check):
    """A mixed-precision run states its error against the FP64 path, printed
    and in output_data/fp64_check.npy.  Its first column is the run's final
    energies; its second is the oracle's, after (continue) two further
    iterations from the saved end state or (shadow) the run's iterations
    from the input, resampled to --Fout.  Energies relative 1e-11."""
    from mepgl_amd import npyio, observables

    s = small_string(N=20, F=5, seed=25, multicomponent=True)
    say(f"[gpu] cli mixed --fp64-check {check}")
    npyio.write_input(s, "t", str(tmp_path), double_precision=False)
    r = cli(str(tmp_path), "--simname", "t", "--mode", "L", "--Fin", 5, "--Fout", 7,
            "--iterations", 2, "--relaxation-steps", 3, "--precision", "mixed",
            "--fp64-check", check, "--fp64-check-iterations", 2)
    assert "Mixed precision against the FP64 path" in r.stdout
    assert "exit barrier" in r.stdout
    d = os.path.join(tmp_path, "simulations", "t", "output_data")
    chk = np.load(os.path.join(d, "fp64_check.npy"))
    assert chk.shape == (7, 2) and chk.dtype == np.float64
    out, extra = npyio.read_dir(d)
    np.testing.assert_allclose(chk[:, 0], extra["fenergy"], rtol=1e-7)  # fenergy.npy is float32
    if check == "continue":
        ref = oracle.OracleString(out)
        ref.loop(2, 3, "L")
    else:
        s32, _ = npyio.read_dir(os.path.join(tmp_path, "simulations", "t", "input_data"))
        ref = oracle.OracleString(s32)
        ref.loop(2, 3, "L")
        ref.arc_lengths()
        ref.redistribute(F_new=7)
    assert rel_err(chk[:, 1], ref.free_energy()) < 1e-11
    z = np.load(observables.write_summary("t", str(tmp_path)), allow_pickle=True)
    np.testing.assert_array_equal(z["fenergy_fp64"], chk[:, 1])
    assert z["barrier_rel_err"].shape == (2,)


def test_cli_fp64_check_off(tmp_path):
    from mepgl_amd import npyio

    s = small_string(N=16, F=4, seed=27)
    say("[gpu] cli mixed --fp64-check off")
    npyio.write_input(s, "t", str(tmp_path))
    r = cli(str(tmp_path), "--simname", "t", "--mode", "S", "--Fin", 4, "--iterations", 1,
            "--relaxation-steps", 2, "--precision", "mixed", "--fp64-check", "off")
    assert "FP64 path" not in r.stdout
    assert not os.path.exists(os.path.join(tmp_path, "simulations", "t", "output_data",
                                           "fp64_check.npy"))


def test_cli_stop_key(tmp_path):
    from mepgl_amd import npyio

    s = small_string(N=16, F=4, seed=24)
    say("[gpu] cli stop key")
    npyio.write_input(s, "t", str(tmp_path))
    # The key arrives while the first iterations run; the loop ends after
    # the iteration in progress and the results are still written.
    p = subprocess.run([EXE, "--simname", "t", "--mode", "S", "--Fin", "4", "--iterations",
                        "100000", "--relaxation-steps", "2"], cwd=str(tmp_path), input="q",
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "I am quitting! Please wait." in p.stdout
    n = p.stdout.count("Iteration:")
    assert 1 <= n < 100000
    assert os.path.exists(os.path.join(tmp_path, "simulations", "t", "output_data", "3", "b.npy"))


def test_launcher_stages(tmp_path, oracle):
    from mepgl_amd import launcher, npyio

    s = small_string(N=16, F=4, seed=25)
    say("[gpu] launcher stages 4 -> 7")
    npyio.write_input(s, "t", str(tmp_path))
    launcher.launch_simulation("t", [4, 7], [2, 2], ["L", "L"], root=str(tmp_path),
                               extra_args=["--relaxation-steps", "3", "--no-stdin"], capture=True)
    out, extra = npyio.read_dir(os.path.join(tmp_path, "simulations", "t", "output_data"))
    assert out.F == 7
    ref = oracle.OracleString(s)
    ref.loop(2, 3, "L")
    ref.arc_lengths()
    ref.redistribute(F_new=7)
    # The reloaded input is what the program wrote: planes and couplings
    # round-trip through float64 files exactly.
    ref2 = oracle.OracleString(npyio.read_dir(os.path.join(
        tmp_path, "simulations", "t", "input_data"))[0])
    np.testing.assert_array_equal(ref2.psi_1.shape, ref.psi_1.shape)
    assert normwise(ref2.psi_1, ref.psi_1) < 1e-10
    ref2.loop(2, 3, "L")
    assert normwise(out.psi_1, ref2.psi_1) < 1e-10


def test_cli_stdin_keys(tmp_path):
    """+ and - change the relaxation steps, C toggles NLCG, q stops
    (mepgl.cu:647-669).  As in the reference, one read takes up to 4 bytes
    and acts on the first, so the keys are sent one at a time once the loop
    runs; the output goes to a file so that the program never blocks."""
    import time

    from mepgl_amd import npyio

    s = small_string(N=16, F=4, seed=26)
    say("[gpu] cli stdin keys")
    npyio.write_input(s, "t", str(tmp_path))
    log = tmp_path / "out.txt"
    with open(log, "w") as fo:
        p = subprocess.Popen([EXE, "--simname", "t", "--mode", "S", "--Fin", "4", "--iterations",
                              "100000", "--relaxation-steps", "2"], cwd=str(tmp_path),
                             stdin=subprocess.PIPE, stdout=fo, stderr=subprocess.STDOUT,
                             text=True)
        try:
            t0 = time.time()
            while "Starting MEP-MGL." not in log.read_text():
                assert p.poll() is None and time.time() - t0 < 60, log.read_text()[-2000:]
                time.sleep(0.1)
            for key in ("+", "+", "-", "C", "q"):
                time.sleep(0.3)
                p.stdin.write(key)
                p.stdin.flush()
            p.stdin.close()
            p.wait(timeout=120)
        finally:
            if p.poll() is None:
                p.kill()
    out = log.read_text()
    assert p.returncode == 0, out[-2000:]
    i3 = out.index("relaxation_steps = 3")
    i4 = out.index("relaxation_steps = 4")
    assert i3 < i4 < out.rindex("relaxation_steps = 3")
    assert "NLCG set OFF" in out
    assert "I am quitting! Please wait." in out


def test_run_py_workflow_matches_oracle(tmp_path, oracle):
    """python -m mepgl_amd.run on a reference-style config.py: two stages
    (5 images, 2 outer iterations of mode L, resampled to 7; then 7 images,
    2 iterations), then a reload-and-continue run (-rc) of the second stage,
    against the oracle running the same sequence; and the .npz summary."""
    from test_run_workflow import write_config

    from mepgl_amd import npyio, observables

    cfgp = write_config(tmp_path, dp=True, mc=True)
    say("[gpu] run.py workflow: 2 stages, then -rc")
    r = subprocess.run([sys.executable, "-m", "mepgl_amd.run", "--config", cfgp, "--root",
                        str(tmp_path)], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    sim = os.path.join(tmp_path, "simulations", "wf")
    assert os.path.exists(os.path.join(sim, "config.py"))
    s0, _ = npyio.read_dir(os.path.join(sim, "input_data"))  # stage 2's input (reloaded)
    assert s0.F == 7
    out, extra = npyio.read_dir(os.path.join(sim, "output_data"))

    # The oracle: stage 1 from the config's initial guess.
    from mepgl_amd import run as runmod

    cfg = runmod.load_config(cfgp)
    init_dir = os.path.join(tmp_path, "init")
    os.makedirs(init_dir)
    runmod.write_input(cfg, init_dir)
    s_init, _ = npyio.read_dir(os.path.join(init_dir, "simulations", "wf", "input_data"))
    ref = oracle.OracleString(s_init)
    ref.loop(2, 4, "L")
    ref.arc_lengths()
    ref.redistribute(F_new=7)
    assert normwise(s0.psi_1, ref.psi_1) < 1e-10
    # Stage 2 restarts the solver from the reloaded files.
    ref2 = oracle.OracleString(s0)
    ref2.loop(2, 4, "L")
    assert normwise(out.psi_1, ref2.psi_1) < 1e-10
    assert normwise(out.psi_2, ref2.psi_2) < 1e-10
    assert normwise(out.a, ref2.a) < 1e-10
    assert rel_err(extra["fenergy"], ref2.free_energy()) < 1e-11
    z = np.load(os.path.join(sim, "wf.npz"), allow_pickle=True)
    np.testing.assert_array_equal(z["fenergy"], extra["fenergy"])
    b = observables.barriers(extra["fenergy"])
    assert f"{b[0]:.8g}" in r.stdout

    # -rc: the output becomes the input and every stage runs again from it.
    r = subprocess.run([sys.executable, "-m", "mepgl_amd.run", "-rc", "--no-post", "--config",
                        cfgp, "--root", str(tmp_path)], cwd=ROOT, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Reloading simulation data" in r.stdout
