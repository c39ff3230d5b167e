"""Plain bounce-back solver on the GPU (GPU): the voxelizers and run_lbm
against the reference's own outputs (tests/golden/bb_*.npz) bit for bit in
FP64, against the oracle on a larger window, and the FP32 path within
tolerance (relative 1e-4 in k, 1e-5 of the largest velocity in u_z)."""

import numpy as np
import pytest

from oracle import bounce_back as BB
from tests._golden import load
from tests.conftest import load_package

pytestmark = pytest.mark.gpu


def test_voxelizers_match_reference():
    lbm = load_package("lbm")
    g = load("bb_voxelize")
    print("[bb voxelize] golden", flush=True)
    solid, dx = lbm.voxelize(g["points"], g["center"], g["size"], float(g["radius"]), int(g["n"]))
    assert np.array_equal(solid, g["solid"]) and np.array_equal(dx, g["dx"])
    v = load("bb_voxelize_variable")
    assert np.array_equal(lbm.voxelize_variable(v["particles"], float(v["box_size"]), int(v["n"]),
                                                float(v["inflate"])), v["solid"])
    assert np.array_equal(lbm.voxelize_variable(v["particles"], float(v["box_size"]), int(v["n"])),
                          v["solid_no_inflate"])


@pytest.mark.parametrize("graph", [False, True])
def test_run_lbm_fp64_matches_reference(graph):
    lbm = load_package("lbm")
    g = load("bb_run_lbm")
    print(f"[bb run_lbm fp64] golden, graph={graph}", flush=True)
    s = lbm.BounceBackD3Q19(g["solid"], tau=float(g["tau"]), force_z=float(g["force_z"]), device="cuda:0")
    s.run(int(g["steps"]), graph=graph)
    out = s.result(return_fields=True)
    for k, v in out.items():
        assert np.array_equal(np.asarray(v), g["out_" + k]), k


def test_run_force_set_matches_reference():
    lbm = load_package("lbm")
    g = load("bb_force_set")
    print("[bb force set] golden", flush=True)
    res, slope, intercept, r2 = lbm.run_force_set(g["solid"], float(g["tau"]), float(g["force_z"]), int(g["steps"]),
                                                  True, device="cuda:0")
    assert slope == float(g["slope"]) and intercept == float(g["intercept"]) and r2 == float(g["r2"])
    assert res["permeability_lu"] == float(g["permeability"])


def test_run_lbm_larger_window_vs_oracle_and_fp32():
    lbm = load_package("lbm")
    rng = np.random.default_rng(3)
    n, steps = 24, 150
    pts = rng.uniform(0.0, 1.0, (30, 3))
    solid, _ = BB.voxelize(pts, np.array([0.5, 0.5, 0.5]), np.array([1.0, 1.0, 1.0]), 0.09, n)
    print(f"[bb {n}^3 x {steps}] oracle vs fp64 vs fp32, solid {solid.mean():.3f}", flush=True)
    ref = BB.run_lbm(solid, 0.8, 1.0e-5, steps, return_fields=True)
    out = lbm.run_lbm(solid, 0.8, 1.0e-5, steps, return_fields=True, device="cuda:0")
    for k, v in out.items():
        assert np.array_equal(np.asarray(v), np.asarray(ref[k])), k
    o32 = lbm.run_lbm(solid, 0.8, 1.0e-5, steps, return_fields=True, precision="fp32", device="cuda:0")
    assert abs(o32["permeability_lu"] - ref["permeability_lu"]) <= 1e-4 * abs(ref["permeability_lu"])
    assert np.max(np.abs(o32["uz"] - ref["uz"])) <= 1e-5 * np.max(np.abs(ref["uz"]))


def test_every_bounce_back_launch_gives_identical_results():
    """The FP32 bounce-back step's launch tuner may pick any candidate
    (workgroup shape, non-temporal access, occupancy cap): all compute the
    same bits, and tuning leaves the state unchanged."""
    import torch
    lbm = load_package("lbm")
    at = load_package("autotune")
    rng = np.random.default_rng(4)
    n = 40
    pts = rng.uniform(0.0, 1.0, (25, 3))
    solid, _ = BB.voxelize(pts, np.array([0.5, 0.5, 0.5]), np.array([1.0, 1.0, 1.0]), 0.1, n)
    s = lbm.BounceBackD3Q19(solid, tau=0.8, force_z=1.0e-5, precision="fp32", device="cuda:0")
    s.run(5)  # tunes first
    assert not s._autotune
    before = s._a.clone()
    s._tune()
    assert torch.equal(before, s._a)
    stream = torch.cuda.current_stream().cuda_stream
    ref = None
    for bz, by, nt, bpc, flat in at.candidates(n):
        if flat:
            continue
        s.block, s.flags = (bz, by), at.flags_of(nt, bpc, 0)
        s._launch(s._a, s._b, stream)
        out = s._b.cpu().numpy().copy()
        if ref is None:
            ref = out
        assert np.array_equal(out, ref), (bz, by, nt, bpc)


def test_fp64_bounce_back_nontemporal_gives_the_same_bits():
    """The FP64 bounce-back step's one launch flag, non-temporal state access,
    leaves every bit of the step unchanged (with and without a force, over
    the FP64 workgroup shapes the tuner tries)."""
    lbm = load_package("lbm")
    at = load_package("autotune")
    import torch
    rng = np.random.default_rng(6)
    n = 36
    pts = rng.uniform(0.0, 1.0, (20, 3))
    solid, _ = BB.voxelize(pts, np.array([0.5, 0.5, 0.5]), np.array([1.0, 1.0, 1.0]), 0.1, n)
    stream = torch.cuda.current_stream().cuda_stream
    for force in (0.0, 1.0e-5):
        print(f"[bb fp64 nt] force {force}", flush=True)
        s = lbm.BounceBackD3Q19(solid, tau=0.8, force_z=force, precision="fp64", device="cuda:0", block=(64, 4))
        s.run(3, graph=False)
        ref = None
        for bz, by in [(64, 4), (128, 2)] + list(at.FP64_SHAPES):
            for nt in (0, 1):
                s.block, s.flags = (bz, by), at.flags_of(nt, 0, 0)
                s._launch(s._a, s._b, stream)
                out = s._b.cpu().numpy().copy()
                if ref is None:
                    ref = out
                assert np.array_equal(out, ref), (force, bz, by, nt)
