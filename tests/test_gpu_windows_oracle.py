"""Batched pore windows against the oracle (GPU).

A reduced, seeded batch from the batched-window generator
(``workloads.pebble_windows``: periodic windows of edge 26-34 with pebbles
of radius 7 near solid fraction 0.6 and up to 60 throat fines of lognormal
radius clipped to [2, 5]) holds one window of each variant: skeleton,
passive fines and preloaded fines (fines touching and overlapping pebbles
and each other, so cells with sum(alpha) > 1 exercise the unclipped owner
weights and the clipped solid field, psc3d.py:175, :344, :368-371, :386).
The whole batch runs in the persistent launch with the loads fused and
graph replays, exactly as the benchmarked configuration does, and each
window is compared with the oracle's sparse layout in periodic-geometry mode
(psc3d.py:284-407 with boxes wrapped) under the FP32 tolerances: f within
1e-4 of the signal max|f_ref - w|, loads within 1e-3 of the window's
largest, k within 1e-5 relative.
"""

import numpy as np
import pytest

from oracle import psc_oracle as O
from tests.conftest import load_package

pytestmark = pytest.mark.gpu

TAU, S, F, STEPS = 0.8, 3, (2.0e-6, 0.0, 1.0e-5), 40


def test_batched_generator_windows_match_oracle():
    wl = load_package("workloads")
    batch = load_package("batch")
    wins = wl.pebble_windows(seed=6, count=3, lmin=26, lmax=34, radius=7.0, max_fines=60)
    assert [v for _, _, v in wins] == ["skeleton", "passive", "preloaded"]
    print("[windows vs oracle] " + ", ".join(f"{v} L={L} P={len(t)}" for L, t, v in wins), flush=True)
    B = batch.BatchedPscD3Q19([(L, t) for L, t, _ in wins], tau=TAU, subsamples=S, body_force=F, device="cuda:0")
    B.run(STEPS)
    k_batch = B.permeability(axis=2)
    loads = B.loads()
    for w, (L, t, variant) in enumerate(wins):
        o = O.OracleSolver(L, [O.Sphere(*row) for row in t], tau=TAU, subsamples=S, body_force=F,
                           static_geometry=True, layout="sparse", periodic_geometry=True)
        for _ in range(STEPS):
            o.collide_stream_psc()
        if variant == "preloaded":
            assert np.max(o.alpha_sum) > 1.0 and np.count_nonzero(o.alpha_sum > 1.0) > 100
        # the batch stores post-collision deviations (pull layout): f_i(x) = g_i(x - c_i) + w_i
        g = B.window_state(w).astype(np.float64)
        f = np.stack([np.roll(g[i], tuple(int(c) for c in O.C[i]), axis=(0, 1, 2)) for i in range(19)])
        f += O.W[:, None, None, None]
        signal = np.max(np.abs(o.f - O.W[:, None, None, None]))
        err = np.max(np.abs(f - o.f))
        ref = np.concatenate([o.hydro_forces, o.hydro_torques], axis=1)
        scale = np.max(np.abs(o.hydro_forces))
        el = np.max(np.abs(loads[w][:, 0:3] - o.hydro_forces)) / scale
        et = np.max(np.abs(loads[w][:, 3:6] - o.hydro_torques)) / (scale * L)
        k_ref = O.permeability(o.solid, o.uz, L, TAU, F[2])
        ek = abs(k_batch[w] - k_ref) / abs(k_ref)
        print(f"  {variant}: f err {err / signal:.2e} of the signal {signal:.2e}; force err {el:.2e}, torque err "
              f"{et:.2e}; k {k_batch[w]:.6e} vs {k_ref:.6e} ({ek:.2e})", flush=True)
        assert ref.shape == loads[w].shape
        assert err <= 1e-4 * signal
        assert el <= 1e-3 and et <= 1e-3
        assert ek <= 1e-5
