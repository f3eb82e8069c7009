"""Drop-in replacement for the reference's compiled ``pycusaxs_cuda`` module
(pybind/cuda_bindings.cpp:172-195): ``run(...)`` with the same keyword
arguments, return value and exceptions, computed on the MI355X engine."""
from pycusaxs_amd.backend import run

__doc__ = "MI355X interface for configuring CuSAXS"
__all__ = ["run"]
