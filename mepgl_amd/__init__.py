"""MI355X-native gauged string method (minimum-energy paths of Ginzburg-Landau
superconductors): HIP kernels for gfx950 behind a C ABI (libmepgl.so), with
a Python host mirror of the reference solver interface."""
from .problem import GLString  # noqa: F401
from . import configs  # noqa: F401

__all__ = ["GLString", "configs", "Solver"]


def __getattr__(name):
    if name == "Solver":
        from .solver import Solver

        return Solver
    raise AttributeError(name)
