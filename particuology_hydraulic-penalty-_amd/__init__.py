"""MI355X-native D3Q19 partially-saturated-cell IMB lattice-Boltzmann step.

A drop-in GPU implementation of ``src/lbm_dem/psc3d.py`` of
wangjianfttt/particuology_Hydraulic-penalty-: hand-written gfx950 HIP
kernels (libpsc_hip.so, C ABI in include/psc_hip.h) behind the reference's
Python class API.  The directory name is not a Python identifier, so import
it with ``importlib.import_module("particuology_hydraulic-penalty-_amd")``
(or its ``.psc3d`` submodule).
"""

from . import _lib
from .psc3d import (C, CS2, OPP, W, LazyDiag, Particle3D, PscD3Q19, SparsePscD3Q19, StaleDiagnostics, equilibrium,
                    guo_force)

__all__ = ["C", "CS2", "OPP", "W", "Particle3D", "PscD3Q19", "SparsePscD3Q19", "LazyDiag", "StaleDiagnostics",
           "equilibrium", "guo_force", "library_path", "load_library"]


def library_path() -> str:
    return _lib.LIB_PATH


def load_library():
    """Load libpsc_hip.so (raises PscError if it has not been built)."""
    return _lib.load()
