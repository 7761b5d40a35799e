"""The config.py / run.py workflow on the CPU: a reference-style config
module (importing mepgl_lib.builder) loads through mepgl_amd.run, its input
files have the reference's names, layout and precision, the program options
carry its compile-time settings, and cut removes images as mepgl/cut.py
does.  The end-to-end run on the GPU is in test_gpu_cli.py."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CONFIG = '''
import numpy as np
from mepgl_lib.builder import StringBuilder

dev_number = 0
double_precision = {dp}
simulation_name = "wf"
F = np.array([5, 7])
iterations = np.array([2, 2])
modes = np.array(["L", "L"])
default_relaxation_step_number = 4
multicomponent = {mc}
thin_film = False
N = 20
L = 5.0
x_lim = [-L / 2, L / 2]
y_lim = [-L / 2, L / 2]
builder = StringBuilder(F[0], Nx=N, Ny=N, x_lim=x_lim, y_lim=y_lim, multicomponent=multicomponent)
x, y = builder.x, builder.y
dx = builder.x_ax[1] - builder.x_ax[0]
comp_domain, sc_domain = builder.get_domain_matrices()
a_1, b_1, m_1, q_1 = -1.0, 1.0, 1.0, 1.0
a_2, b_2, m_2, q_2 = -0.6, 0.8, 2.0, 0.7
m_xx_1 = m_yy_1 = m_1
m_xx_2, m_yy_2 = 0.5, m_2
eta, gamma, delta = -0.2, 0.3, 0.05
h_z = 0.8
ones = np.ones((N, N))
a_1 = a_1 * ones; b_1 = b_1 * ones; m_xx_1 = m_xx_1 * ones; m_yy_1 = m_yy_1 * ones
a_2 = a_2 * ones; b_2 = b_2 * ones; m_xx_2 = m_xx_2 * ones; m_yy_2 = m_yy_2 * ones
h = h_z * comp_domain
xv = np.linspace(L / 2 + 1, 0, F[0])
builder.add_vortex(xv, 0 * xv, np.ones(F[0]), 1)
np.random.seed(5)
u_1, v_1, u_2, v_2 = builder.generate_matter_fields(psi_abs_1=1.0, psi_abs_2=np.sqrt(0.75), noise=0.05)
builder.B_z = h_z
ax, ay = builder.generate_vector_potential(gauge="symmetric")
'''


def write_config(tmp_path, dp=True, mc=False):
    p = tmp_path / "config.py"
    p.write_text(CONFIG.format(dp=dp, mc=mc))
    return str(p)


@pytest.mark.parametrize("dp,mc", [(True, False), (False, True)])
def test_config_loads_and_writes_reference_layout(tmp_path, dp, mc):
    from mepgl_amd import npyio, run

    cfg = run.load_config(write_config(tmp_path, dp, mc))
    d = run.write_input(cfg, str(tmp_path))
    dt = np.float64 if dp else np.float32
    names = sorted(os.listdir(d))
    assert names[:5] == ["0", "1", "2", "3", "4"]
    for nm in ("x", "y", "a_1", "b_1", "m_xx_1", "m_yy_1", "a_2", "b_2", "m_xx_2", "m_yy_2",
               "h", "q"):
        assert np.load(os.path.join(d, nm + ".npy")).dtype == dt, nm
    assert np.load(os.path.join(d, "comp_domain.npy")).dtype == np.int32
    np.testing.assert_allclose(np.load(os.path.join(d, "q.npy")), [1.0, 0.7, -0.2, 0.3, 0.05],
                               rtol=1e-7)
    # The files hold the config's [i][j] arrays as they are; read back, the
    # string's planes are their transposes ([j][i]).
    s, _ = npyio.read_dir(d)
    assert s.F == 5 and s.N == 20 and s.multicomponent == mc
    np.testing.assert_allclose(s.psi_1[:, :, :, 0], np.transpose(cfg.u_1, (0, 2, 1)), rtol=1e-6)
    np.testing.assert_allclose(s.a[:, :, :, 1], np.transpose(cfg.ay, (0, 2, 1)), rtol=1e-6)
    np.testing.assert_allclose(s.dx, cfg.dx, rtol=1e-15)
    args = run.program_args(cfg)
    assert args[args.index("--relaxation-steps") + 1] == "4"
    assert args[args.index("--ncomp") + 1] == ("2" if mc else "1")
    assert args[args.index("--precision") + 1] == ("fp64" if dp else "mixed")


def _fake_output(tmp_path, F=6):
    from conftest import small_string

    from mepgl_amd import npyio

    s = small_string(N=8, F=F, seed=3)
    npyio.write_input(s, "c", str(tmp_path))
    sim = os.path.join(tmp_path, "simulations", "c")
    os.rename(os.path.join(sim, "input_data"), os.path.join(sim, "output_data"))
    out = os.path.join(sim, "output_data")
    for nm in ("fenergy", "s", "s_jb"):
        np.save(os.path.join(out, nm + ".npy"), np.arange(F, dtype=np.float64) + 10)
    return s, out


def test_cut_removes_and_renumbers(tmp_path):
    from mepgl_amd import cut

    s, out = _fake_output(tmp_path)
    assert cut.main(["-s", "c", "-i", "1-3", "--root", str(tmp_path)]) == 0
    frames = sorted(int(d) for d in os.listdir(out) if d.isdigit())
    assert frames == [0, 1, 2, 3]
    # old images 3, 4, 5 are now 1, 2, 3
    u = np.load(os.path.join(out, "2", "u_1.npy"))
    np.testing.assert_array_equal(u, s.psi_1[4, :, :, 0].T)
    np.testing.assert_array_equal(np.load(os.path.join(out, "fenergy.npy")), [10, 13, 14, 15])
    bak = os.path.join(os.path.dirname(out), "output_data_bak")
    assert sorted(int(d) for d in os.listdir(bak) if d.isdigit()) == list(range(6))


def test_cut_needs_both_arguments(tmp_path):
    _fake_output(tmp_path)
    r = subprocess.run([sys.executable, "-m", "mepgl_amd.cut", "-s", "c", "--root",
                        str(tmp_path)], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 1
    assert not os.path.exists(os.path.join(tmp_path, "simulations", "c", "output_data_bak"))
