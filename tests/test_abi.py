"""The C ABI library loads on the host and exports every declared entry
point; argument checking works without touching a GPU (CPU)."""

import ctypes as ct
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT, load_package

HEADER = os.path.join(ROOT, "include", "psc_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(psc_[a-z_0-9]+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    return load_package("_lib")


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 12
    handle = lib.load()
    for name in names:
        assert hasattr(handle, name), name
    assert set(names) == set(lib.SIGNATURES), "ctypes signature table out of sync with the header"


def test_lattice_constants_match_reference_order(lib):
    from oracle import psc_oracle as O
    c = np.zeros(57, dtype=np.int32)
    w = np.zeros(19)
    opp = np.zeros(19, dtype=np.int32)
    lib.call("psc_lattice_constants", c.ctypes.data, w.ctypes.data, opp.ctypes.data)
    assert np.array_equal(c.reshape(19, 3), O.C)
    assert np.array_equal(w, O.W)
    assert np.array_equal(opp, O.OPP)
    psc3d = load_package("psc3d")
    assert np.array_equal(psc3d.C, O.C) and np.array_equal(psc3d.W, O.W) and np.array_equal(psc3d.OPP, O.OPP)
    assert psc3d.CS2 == O.CS2


def test_argument_errors_are_reported(lib):
    h = lib.load()
    rc = h.psc_step(None, None, None, None, None, 1, 1, 8, 8, 8, 2, 0.8, 0.0, 0.0, 0.0, 0, 0, 0, None)
    assert rc == -1
    assert b"null" in h.psc_last_error()
    rc = h.psc_step(ct.c_void_p(16), ct.c_void_p(16), ct.c_void_p(64), None, None, 1, 1, 8, 8, 8, 2, 0.8, 0, 0, 0,
                    0, 0, 0, None)
    assert rc == -1 and b"alias" in h.psc_last_error()
    rc = h.psc_step(ct.c_void_p(16), ct.c_void_p(32), ct.c_void_p(64), None, None, 1, 1, 8, 8, 8, 2, 0.4, 0, 0, 0,
                    0, 0, 0, None)
    assert rc == -1 and b"tau" in h.psc_last_error()
    rc = h.psc_step(ct.c_void_p(16), ct.c_void_p(32), ct.c_void_p(64), None, None, 1, 1, 8, 8, 8, 2, 0.8, 0, 0, 0,
                    48, 3, 0, None)
    assert rc == -2
    rc = h.psc_voxelize(None, None, None, 0, 2048, 8, 8, 8, 0, 41, None, None, None, None, None, None, 0, 0, None)
    assert rc == -1 and b"subsamples" in h.psc_last_error()
    rc = h.psc_moments(None, 3, 0, None, None, 8, 8, 8, 0, 2, 0, 0, 0, None, None, None, None)
    assert rc == -1
    rc = h.psc_clear(None, 0, None)
    assert rc == 0
    with pytest.raises(lib.PscError):
        lib.call("psc_clear", None, 16, None)
    assert h.psc_moments_blocks(128 ** 3) >= 1
    assert h.psc_version() >= 10000


def test_fused_load_entry_points_check_arguments(lib):
    """psc_step_batched_loads and the fused form of psc_run_steps (also with
    in-place step pairs) reject bad arguments before touching a GPU."""
    h = lib.load()
    t = lib.PscLoadsTables()
    p16, p32, p48, p64 = (ct.c_void_p(v) for v in (16, 32, 48, 64))
    args = [p16, p32, p64, p48, 100, p64, p64, 1, 3, 0.8, 0, 0]
    rc = h.psc_step_batched_loads(*args, None, None, None, None, None, 1, None)
    assert rc == -1 and b"null" in h.psc_last_error()  # no tables
    rc = h.psc_step_batched_loads(*args, ct.byref(t), None, None, None, None, 0, None)
    assert rc == -1 and b"sizes" in h.psc_last_error()  # period 0
    t.nparticles, t.nchunks, t.chunk_cells = 2, 3, 2048
    rc = h.psc_step_batched_loads(*args, ct.byref(t), p48, p16, None, None, 5, None)
    assert rc == -1 and b"imb" in h.psc_last_error()  # imb_prev aliases imb
    rc = h.psc_step_batched_loads(*args, ct.byref(t), p16, p16, None, None, 5, None)
    assert rc == -1  # chunk tables missing
    rc = h.psc_step_batched_loads(*args[:10], 0x100000, 0, ct.byref(t), None, None, None, None, 5, None)
    assert rc == -1 and b"flags" in h.psc_last_error()
    plan = lib.PscStepPlan()
    plan.state_a, plan.state_b, plan.count_sum = 16, 32, 64
    plan.imb, plan.imb_alt, plan.partials, plan.partials_alt = 128, 128, 256, 512
    plan.storage, plan.layout, plan.nx, plan.ny, plan.nz, plan.subsamples = 1, 1, 8, 8, 8, 2
    plan.tau, plan.nparticles, plan.fuse_loads = 0.8, 2, 5
    par = ct.c_int32(0)
    rc = h.psc_run_steps(ct.byref(plan), 4, 0, 16, ct.byref(par), None)
    assert rc == -1 and b"alias" in h.psc_last_error()
    plan.imb_alt, plan.tau = 192, 0.5
    rc = h.psc_run_steps(ct.byref(plan), 4, 0, 16, ct.byref(par), None)
    assert rc == -1 and b"tau" in h.psc_last_error()
    plan.tau, plan.flags = 0.8, lib.PSC_RUN_INPLACE | 0x20000  # in place with two cells per thread
    rc = h.psc_run_steps(ct.byref(plan), 4, 0, 16, ct.byref(par), None)
    assert rc == -1 and b"one cell per thread" in h.psc_last_error()
    plan.flags, plan.fuse_loads = lib.PSC_RUN_INPLACE, 0  # in place outside the fused form
    rc = h.psc_run_steps(ct.byref(plan), 4, 0, 16, ct.byref(par), None)
    assert rc == -1 and b"fused form" in h.psc_last_error()
    rc = h.psc_step(p16, p32, p64, None, None, 1, 1, 8, 8, 8, 2, 0.8, 0.0, 0.0, 0.0, 64, 1, lib.PSC_RUN_INPLACE, None)
    assert rc == -1 and b"flags" in h.psc_last_error()  # a plan flag, not a psc_step flag
    rc = h.psc_step(p16, p32, p64, None, None, lib.PSC_F64_PLAIN, 1, 8, 8, 8, 2, 0.8, 0.0, 0.0, 0.0, 64, 1,
                    lib.PSC_STEP_PLANES64, None)
    assert rc == -1 and b"FP32" in h.psc_last_error()  # padded planes are an FP32 layout


def test_deferred_load_entry_points_check_arguments(lib):
    """psc_run_steps_deferred / psc_flush_loads take a fused plan, a parity of
    0 or 1 and pending bits of PSC_PENDING_*, and reject anything else before
    touching a GPU."""
    h = lib.load()
    plan = lib.PscStepPlan()
    plan.state_a, plan.state_b, plan.count_sum = 16, 32, 64
    plan.imb, plan.imb_alt, plan.partials, plan.partials_alt = 128, 192, 256, 512
    plan.storage, plan.layout, plan.nx, plan.ny, plan.nz, plan.subsamples = 1, 1, 8, 8, 8, 2
    plan.tau, plan.nparticles, plan.fuse_loads = 0.8, 2, 5
    par, pend = ct.c_int32(0), ct.c_int32(0)
    for fn, call in (("psc_run_steps_deferred", lambda: h.psc_run_steps_deferred(ct.byref(plan), 2, ct.byref(par),
                                                                                  ct.byref(pend), None)),
                     ("psc_flush_loads", lambda: h.psc_flush_loads(ct.byref(plan), par.value, ct.byref(pend), None))):
        par.value, pend.value = 2, 0
        assert call() == -1 and b"parity" in h.psc_last_error(), fn
        par.value, pend.value = 1, 4
        assert call() == -1 and b"pending" in h.psc_last_error(), fn
        pend.value = lib.PSC_PENDING_CHUNKS | lib.PSC_PENDING_FINAL
        plan.fuse_loads = 0  # loads not fused
        assert call() == -1 and b"fused plan" in h.psc_last_error(), fn
        plan.fuse_loads, plan.storage = 5, lib.PSC_F64_PLAIN
        assert call() == -1 and b"fused plan" in h.psc_last_error(), fn
        plan.storage, plan.imb_alt = 1, 128
        assert call() == -1 and b"alias" in h.psc_last_error(), fn
        plan.imb_alt, plan.tau = 192, 0.5
        assert call() == -1 and b"tau" in h.psc_last_error(), fn
        plan.tau, plan.flags = 0.8, lib.PSC_RUN_INPLACE | 0x20000  # in place with two cells per thread
        assert call() == -1 and b"one cell per thread" in h.psc_last_error(), fn
        plan.flags = 0
        assert par.value == 1 and pend.value == 3, fn  # a rejected call leaves its outputs alone
    assert h.psc_run_steps_deferred(ct.byref(plan), -1, ct.byref(par), ct.byref(pend), None) == -1
    assert h.psc_run_steps_deferred(None, 1, ct.byref(par), ct.byref(pend), None) == -1
    assert h.psc_run_steps_deferred(ct.byref(plan), 1, None, ct.byref(pend), None) == -1
    assert h.psc_flush_loads(ct.byref(plan), 0, None, None) == -1


def test_solver_refuses_to_run_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    psc3d = load_package("psc3d")
    lib = load_package("_lib")
    with pytest.raises(lib.PscError):
        psc3d.PscD3Q19(8, [psc3d.Particle3D(4, 4, 4, 2)])


def test_fp32_lattices_beyond_2_30_cells_are_rejected(lib):
    """The FP32 kernels address a direction plane with 32-bit byte offsets:
    psc_step and the native loop refuse FP32 lattices above 2^30 cells
    (1025 x 1024 x 1024 here) before any launch, and the facade refuses them
    before allocating; FP64 keeps element offsets (cell ids up to 2^31 - 1)."""
    h = lib.load()
    p16, p32, p64 = (ct.c_void_p(v) for v in (16, 32, 64))
    rc = h.psc_step(p16, p32, p64, None, None, lib.PSC_F32_DEV, 1, 1025, 1024, 1024, 2, 0.8, 0, 0, 0, 0, 0, 0, None)
    assert rc == -1 and b"2^30" in h.psc_last_error()
    rc = h.psc_step(p16, p32, p64, None, None, lib.PSC_F32_DEV, 1, 1024, 1024, 1040, 2, 0.8, 0, 0, 0, 0, 0, 0, None)
    assert rc == -1 and b"2^30" in h.psc_last_error()
    plan = lib.PscStepPlan()
    plan.state_a, plan.state_b, plan.count_sum = 16, 32, 64
    plan.imb, plan.imb_alt, plan.partials, plan.partials_alt = 128, 192, 256, 512
    plan.storage, plan.layout, plan.nx, plan.ny, plan.nz, plan.subsamples = lib.PSC_F32_DEV, 1, 1025, 1024, 1024, 2
    plan.tau, plan.nparticles, plan.fuse_loads = 0.8, 2, 5
    par = ct.c_int32(0)
    rc = h.psc_run_steps(ct.byref(plan), 4, 0, 16, ct.byref(par), None)
    assert rc == -1 and b"2^30" in h.psc_last_error()
    psc3d = load_package("psc3d")
    with pytest.raises(ValueError, match="FP32 lattice"):
        psc3d.PscD3Q19((1025, 1024, 1024), [], precision="fp32")
    with pytest.raises(ValueError, match="at most"):
        psc3d.PscD3Q19((2048, 1024, 1024), [], precision="fp64")
