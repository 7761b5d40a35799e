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


def _nan_direction(sv, image):
    """Replaces string 0's direction with one holding a NaN at one site of
    `image` (the owed update is applied first by the library)."""
    a, p1, p2 = sv.download_state(0, "d")
    p1 = p1.copy()
    p1[image, 7, 9, 1] = np.nan
    sv.upload_state(0, "d", a, p1, p2)


@pytest.mark.parametrize("use_graph", [True, False])
def test_loop_names_the_first_nonfinite_iteration(use_graph):
    """A NaN put into one image's direction between two loop calls makes the
    next call's first outer iteration non-finite: mepgl_loop stops with
    MEPGL_ERR_NONFINITE naming that iteration and the image, after at most
    one further iteration (the check of iteration k runs while k + 1 is
    queued)."""
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    s = small_string(N=20, F=4, seed=64)
    say(f"[gpu] non-finite iteration, graph={use_graph}")
    with Solver([s], use_graph=use_graph) as sv:
        sv.loop(2, 4, "S")
        _nan_direction(sv, 1)
        with pytest.raises(_lib.MepglError) as e:
            sv.loop(5, 4, "S")
        assert e.value.code == ERR_NONFINITE
        msg = str(e.value)
        assert "outer iteration 0" in msg and "image 1" in msg, msg
        # The flag was cleared with the report: a reset state runs clean.
        sv.upload(0, s)
        sv.reset_solver()
        sv.loop(1, 4, "S")


def test_rollback_restores_the_start_of_the_bad_iteration():
    """With rollback the failed call leaves the fields as they were at the
    start of the bad iteration; a clean run with rollback gives the same
    fields as one without it, bit for bit."""
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    s = small_string(N=20, F=4, seed=65)
    say("[gpu] rollback")
    out = []
    for rb in (False, True):
        with Solver([s], rollback=rb) as sv:
            sv.loop(3, 4, "L")
            out.append(sv.download(0))
    for k in range(2):
        np.testing.assert_array_equal(out[0][k], out[1][k])
    with Solver([s], rollback=True) as sv:
        sv.loop(2, 4, "L")
        sv.loop(1, 4, "S")  # leaves an owed update, applied by upload_state
        _nan_direction(sv, 2)
        before = sv.download(0)
        with pytest.raises(_lib.MepglError) as e:
            sv.loop(3, 4, "S")
        assert e.value.code == ERR_NONFINITE
        assert "rolled back" in str(e.value) and "outer iteration 0" in str(e.value)
        after = sv.download(0)
    for k in range(2):
        np.testing.assert_array_equal(after[k], before[k])


def test_gradient_norm_overflow_is_reported():
    """Finite fields whose gradient norm overflows to +Inf (|psi| = 1e60 at
    one site: the energy density is about 1e239, |g|^2 about 1e358) are
    reported; a NaN test alone would miss an infinite norm."""
    from mepgl_amd import _lib
    from mepgl_amd.solver import Solver

    s = small_string(N=20, F=3, seed=66)
    s.psi_1[1, 8, 8, 0] = 1e60
    say("[gpu] gradient norm overflow")
    with Solver([s]) as sv:
        with pytest.raises(_lib.MepglError) as e:
            sv.nlcgSteps(1)
        assert e.value.code == ERR_NONFINITE
        assert "image 1" in str(e.value)


def test_resample_keeps_the_other_strings_state():
    """Resampling string 0 of a two-string context restarts string 0 only:
    string 1 continues exactly as in a context of its own that went through
    the same calls (the resample applies string 1's owed update, as any
    download does)."""
    from mepgl_amd.solver import Solver

    s0 = small_string(N=20, F=5, seed=67)
    s1 = small_string(N=24, F=4, seed=68, multicomponent=True)
    say("[gpu] resample keeps the other strings")
    rows = 8
    with Solver([s0, s1], segment_rows=rows) as sv:
        sv.loop(2, 4, "L")
        g_old = sv.scalars(1, "g_norm2")
        sv.resample(0, 7)
        np.testing.assert_array_equal(sv.scalars(1, "g_norm2"), g_old)
        assert np.all(np.isinf(sv.scalars(0, "g_norm2")))
        sv.loop(1, 4, "S")
        batched = sv.download(1)
    with Solver([s1], segment_rows=rows) as sv:
        sv.loop(2, 4, "L")
        sv.download(0)  # applies the owed update, as the resample's arclength does
        sv.loop(1, 4, "S")
        alone = sv.download(0)
    for k in range(3):
        np.testing.assert_array_equal(batched[k], alone[k])
