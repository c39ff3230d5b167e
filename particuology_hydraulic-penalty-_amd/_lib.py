"""ctypes binding of libpsc_hip.so (the C ABI declared in include/psc_hip.h).

The library is built in-tree by ``csrc/build.py``; there is no fallback: if
it is missing or cannot be loaded, every solver constructor raises.
"""

from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsc_hip.so")

PSC_OK = 0
PSC_F32_DEV = 1
PSC_F64_PLAIN = 2
PSC_PUSH = 0
PSC_PULL = 1
PSC_STEP_NONTEMPORAL = 1
PSC_NSUM = 13
PSC_ALPHA_FIELD = 0
PSC_BATCH_WG512 = 0x40000
PSC_BATCH_IMB_NT = 0x80000
PSC_STEP_LINEAR = 0x100000
PSC_RUN_INPLACE = 0x200000
PSC_STEP_PLANES64 = 0x400000
PSC_PENDING_CHUNKS = 1
PSC_PENDING_FINAL = 2

# Names and argument types of every exported entry point, in header order.
_P = ct.c_void_p
_I32 = ct.c_int32
_I64 = ct.c_int64
_D = ct.c_double
SIGNATURES = {
    "psc_version": ([], ct.c_int),
    "psc_last_error": ([], ct.c_char_p),
    "psc_lattice_constants": ([_P, _P, _P], ct.c_int),
    "psc_clear": ([_P, _I64, _P], ct.c_int),
    "psc_init_state": ([_P, _I32, _I64, _P], ct.c_int),
    "psc_equilibrium": ([_P, _P, _P, _P, _I64, _P, _P], ct.c_int),
    "psc_guo_force": ([_D, _D, _D, _P, _P, _P, _I64, _D, _P, _P], ct.c_int),
    "psc_voxelize": ([_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P, _P, _I32, ct.c_uint32,
                      _P], ct.c_int),
    "psc_step_batched": ([_P, _P, _P, _P, _I64, _P, _P, _I32, _I32, _D, _I32, _I32, _P], ct.c_int),
    "psc_step_batched_loads": ([_P, _P, _P, _P, _I64, _P, _P, _I32, _I32, _D, _I32, _I32, _P, _P, _P, _P, _P,
                                _I32, _P], ct.c_int),
    "psc_clear_entries": ([_P, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _I32, _P], ct.c_int),
    "psc_voxelize_owners": ([_P, _P, _I32, _P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P,
                             _P], ct.c_int),
    "psc_clear_entries_owners": ([_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _P], ct.c_int),
    "psc_step_owners": ([_P, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _I32, _I32, _D, _D, _D, _D, _I32, _I32, _I32,
                         _P], ct.c_int),
    "psc_step": ([_P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _D, _D, _D, _D, _I32, _I32, _I32, _P],
                 ct.c_int),
    "psc_particle_loads": ([_P, _P, _P, _P, _I32, _I32, _I32, _P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _P,
                            _P, _P], ct.c_int),
    "psc_moments": ([_P, _I32, _I32, _P, _P, _I32, _I32, _I32, _I32, _I32, _D, _D, _D, _P, _P, _P, _P], ct.c_int),
    "psc_moments_blocks": ([_I64], ct.c_int64),
    "psc_moments_noalloc": ([_P, _I32, _I32, _P, _P, _I32, _I32, _I32, _I32, _I32, _D, _D, _D, _P, _P, _P],
                            ct.c_int),
    "psc_export_f64": ([_P, _I32, _I32, _I32, _I32, _I32, _P, _P], ct.c_int),
    "psc_import_f64": ([_P, _I32, _I32, _I32, _I32, _I32, _P, _P], ct.c_int),
    "psc_run_steps": ([_P, _I32, _I32, _I32, _P, _P], ct.c_int),
    "psc_prepare_steps": ([_P, _I32, _I32, _P], ct.c_int),
    "psc_graph_cache_clear": ([], ct.c_int),
    "psc_run_steps_deferred": ([_P, _I32, _P, _P, _P], ct.c_int),
    "psc_flush_loads": ([_P, _I32, _P, _P], ct.c_int),
    "psc_copy_f32": ([_P, _P, _I64, _P], ct.c_int),
    "psc_voxelize_binary": ([_P, _P, _I32, _P, _I32, _I32, _I32, _P, _P], ct.c_int),
    "psc_bb_step": ([_P, _P, _P, _I32, _I32, _I32, _I32, _D, _D, _D, _D, _I32, _I32, _I32, _P], ct.c_int),
    "psc_smooth_alpha": ([_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P], ct.c_int),
    "psc_smooth_alpha_grad": ([_P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P, _P],
                              ct.c_int),
    "psc_adjoint_seed": ([_P, _P, _P, _P, _I32, _I32, _I32, _I32, _D, _D, _D, _D, _D, _P], ct.c_int),
    "psc_adjoint_step": ([_P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _D, _D, _D, _D, _I32, _I32, _I32, _P],
                         ct.c_int),
}


class PscStepPlan(ct.Structure):
    """Mirror of struct PscStepPlan (include/psc_hip.h)."""

    _fields_ = [
        ("state_a", _P), ("state_b", _P), ("count_sum", _P), ("us_num", _P), ("imb", _P), ("imb_alt", _P),
        ("storage", _I32), ("layout", _I32), ("nx", _I32), ("ny", _I32), ("nz", _I32), ("subsamples", _I32),
        ("tau", _D), ("fx", _D), ("fy", _D), ("fz", _D),
        ("block_z", _I32), ("block_y", _I32), ("flags", _I32),
        ("particles", _P), ("boxes", _P), ("chunks", _P), ("chunk_first", _P),
        ("nparticles", _I32), ("nchunks", _I32), ("chunk_cells", _I32),
        ("ent_cell", _P), ("ent_loc", _P), ("ent_cnt", _P), ("chunk_nnz", _P), ("partials", _P),
        ("partials_alt", _P), ("loads", _P), ("fuse_loads", _I32),
    ]


class PscLoadsTables(ct.Structure):
    """Mirror of struct PscLoadsTables (include/psc_hip.h)."""

    _fields_ = [("particles", _P), ("boxes", _P), ("chunks", _P), ("chunk_first", _P),
                ("nparticles", _I32), ("nchunks", _I32), ("chunk_cells", _I32),
                ("ent_cell", _P), ("ent_loc", _P), ("ent_cnt", _P), ("chunk_nnz", _P), ("count_sum", _P)]


class PscWindow(ct.Structure):
    """Mirror of struct PscWindow (include/psc_hip.h)."""

    _fields_ = [("state_off", _I64), ("cell_off", _I64), ("nx", _I32), ("ny", _I32), ("nz", _I32), ("pz", _I32),
                ("inv_pz", ct.c_float), ("inv_nypz", ct.c_float), ("fx", ct.c_float), ("fy", ct.c_float),
                ("fz", ct.c_float), ("span", _I32), ("ps", _I64)]


class PscError(RuntimeError):
    """A libpsc_hip call returned a non-zero status."""


_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes library with typed entry points."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PscError(
            f"{p} is missing: build it with `python particuology_hydraulic-penalty-_amd/csrc/build.py` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ct.CDLL(p)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != PSC_OK:
        msg = load().psc_last_error().decode(errors="replace")
        raise PscError(f"{what} failed with status {rc}: {msg}")


def call(name: str, *args) -> int:
    """Call an entry point and raise PscError on a non-zero status."""
    fn = getattr(load(), name)
    rc = fn(*args)
    check(rc, name)
    return rc
