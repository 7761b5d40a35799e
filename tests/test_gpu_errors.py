"""Error handling of the C ABI on the GPU: argument and state checks, and the
detection of non-finite energies.  Every case is decided by the host, or by
ordinary arithmetic on NaN data; no kernel is made to fault."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import small_string

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_NONFINITE = -1, -3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_nonfinite_image_is_reported():
    """A NaN in one image's order parameter makes that image's energy and
    gradient norm NaN; nlcgSteps returns MEPGL_ERR_NONFINITE naming it, and
    the other images stay finite."""
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    s = small_string(N=20, F=4, seed=60)
    s.psi_1[2, 10, 10, 0] = np.nan
    say("[gpu] non-finite image")
    with Solver([s]) as sv:
        with pytest.raises(_lib.MepglError) as e:
            sv.nlcgSteps(1)
        assert e.value.code == ERR_NONFINITE
        assert "image 2" in str(e.value)
        fen = sv.scalars(0, "fenergy")
    assert np.isnan(fen[2])
    assert np.all(np.isfinite(fen[[0, 1, 3]]))


def test_argument_errors():
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    s = small_string(N=20, F=4, seed=61)
    say("[gpu] argument errors")
    with Solver([s]) as sv:
        out = np.zeros(4)
        assert sv._L.mepgl_nlcg_steps(sv._ctx, -1, 1, 0) == ERR_ARG
        assert sv._L.mepgl_get_scalars(sv._ctx, 5, _lib.SCALARS["fenergy"], _ptr(out)) == ERR_ARG
        assert sv._L.mepgl_set_scalars(sv._ctx, 0, _lib.SCALARS["alpha"], _ptr(out)) == ERR_ARG
        with pytest.raises(_lib.MepglError) as e:
            sv.autotune(steps=0)
        assert e.value.code == ERR_ARG
        # The context still works after the rejected calls.
        sv.nlcgSteps(2)
        assert np.all(np.isfinite(sv.scalars(0, "fenergy")))


def test_reparametrisation_needs_enough_images():
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    say("[gpu] reparametrisation image counts")
    with Solver([small_string(N=20, F=3, seed=62)]) as sv:
        sv.arcLength()
        with pytest.raises(_lib.MepglError) as e:
            sv.redistributeCubic()
        assert e.value.code == ERR_ARG
        assert "4 images" in str(e.value)
    with Solver([small_string(N=20, F=1, seed=63)]) as sv:
        with pytest.raises(_lib.MepglError) as e:
            sv.arcLength()
        assert e.value.code == ERR_ARG
