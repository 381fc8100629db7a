"""rrlbm: MI355X (gfx950) recursively-regularised lattice-Boltzmann solver.

Python host mirror of the C ABI in include/rrlbm.h; all computation happens in
librrlbm.so on the GPU.
"""
from ._lib import LIB_PATH, Diag, Params, RRLBMError, load
from .solver import Batch, Field, Solver, Storage, default_params, device_count, diag_flow_noalloc, lattice

__all__ = ["LIB_PATH", "Batch", "Diag", "Field", "Params", "RRLBMError", "Solver", "Storage", "default_params",
           "device_count", "diag_flow_noalloc", "lattice", "load"]


def version():
    return load().rrlbm_version().decode()
