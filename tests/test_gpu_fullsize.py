"""The batched-window, large moving-window and gradient cases at their full
sizes (GPU), checked through properties that hold at any size.

Momentum budget of one step.  With B = a/(a + tau(1 - a) + 1e-14) and
u_eff = (j + F)/rho, the post-collision populations of psc3d.py:220-232 carry
the momentum
    sum_i c_i f'_i = j + (1 - B) F/tau + B (rho u_s - j + F) + (1 - 1/(2 tau)) F
(the BGK term moves j to rho u_eff = j + F, the solid term sum_i c_i Omega_s,i
= rho u_s - j + F, Guo's first moment is (1 - 1/(2 tau)) F), and the solid
term's momentum is exactly imb = B (rho u_s - j + F) (psc3d.py:234-236).
Streaming permutes populations over the periodic lattice, so over a whole
lattice
    sum_x j(t+1) - sum_x j(t) = sum_x imb + F sum_x ((1 - B)/tau + 1 - 1/(2 tau)),
while mass is unchanged (sum_i Omega_s,i = sum_i Guo_i = 0).  The owner
weights of each cell add up to one, so sum_p F_p = -sum_x imb
(accumulate_particle_loads, psc3d.py:255-281, 375-407: the reference's
momentum_residual).  A dropped or sign-flipped term, a wrong owner weight or
a lost cell breaks one of these by O(1) of the driving; FP32 rounding moves
them by ~1e-8 of it (the bars below are 30-100 times the observed values).

* Batched windows (64 periodic windows, L ~ U{48..160}, pebbles r = 16 and
  0-400 throat fines in skeleton / passive / preloaded variants, tau 0.8,
  F = 1e-6 z, s = 3, FP32): the budget, mass and load closure per window
  after warm-up steps in the tuned persistent launch with fused loads.
* Large moving window (periodic 512^3, 2000 spheres r ~ U[6, 18] moving and
  re-voxelized every step, FP32 owner table): the budget (F = 0), mass and
  closure; and the per-particle loads against the FP64 kernel on the same
  trajectory.
* Permeability gradient (periodic 96^3 window, 40 pebbles r = 14 + 300
  fines, smooth solid fraction, 2000 steps, sqrt(T) checkpointing): the FP32
  adjoint gradient against FP64 central differences (h = 1e-3 cells) on the
  largest components, bar 1e-4 (the accuracy target of this case is 1e-3),
  and k(FP32) against k(FP64) within 1e-5.
"""

import importlib

import numpy as np
import pytest

from tests.conftest import load_package

pytestmark = pytest.mark.gpu

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
              [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], dtype=np.float64)


def _plane_sums(torch, state, off, ps, ncell):
    """FP64 sums over the first ncell cells of each of the 19 direction planes
    of ps cells starting at element off."""
    v = state[off: off + 19 * ps].view(19, ps)[:, :ncell]
    return np.array([float(v[i].double().sum()) for i in range(19)])


def _forcing_factor(torch, csum, s3, tau):
    """sum_x ((1 - B)/tau + 1 - 1/(2 tau)) over the cells of csum (FP64)."""
    a = torch.clamp(csum.double() / s3, max=1.0)
    B = a / (a + tau * (1.0 - a) + 1e-14)
    return float(((1.0 - B) / tau).sum()) + csum.numel() * (1.0 - 0.5 / tau)


def _imb_sum(torch, imb, csum):
    """sum of imb (3 interleaved values per cell) over the cells with solid."""
    m = (csum > 0).unsqueeze(1)
    return (imb.view(-1, 3).double() * m).sum(dim=0).cpu().numpy()


def test_batched_windows_full_size_momentum_budget():
    torch = importlib.import_module("torch")
    batch = load_package("batch")
    workloads = load_package("workloads")
    tau, s, F = 0.8, 3, np.array([0.0, 0.0, 1.0e-6])
    print("[windows full size] generating 64 windows", flush=True)
    wins = workloads.pebble_windows(seed=0)
    B = batch.BatchedPscD3Q19([(L, t) for L, t, _ in wins], tau=tau, subsamples=s, body_force=tuple(F),
                              device="cuda:0")
    assert int(B.nlive.sum()) > 80_000_000
    B.run(24)  # tuned persistent launch, loads fused, graph replays
    torch.cuda.synchronize()
    print(f"  {B.nwin} windows, {int(B.nlive.sum())} cells, launch {B._flags:#x} {B._mode}", flush=True)

    def sums():
        return [_plane_sums(torch, B._a, int(B.state_off[w]), int(B.plane[w]), int(B.ncell[w]))
                for w in range(B.nwin)]

    before = sums()
    B.step()  # one step writing B._imb, then its loads
    after = sums()
    loads = B.loads()
    worst_j = worst_m = worst_c = 0.0
    for w in range(B.nwin):
        o, n = int(B.cell_off[w]), int(B.ncell[w])
        csum = B._csum[o: o + n]
        imb = _imb_sum(torch, B._imb[3 * o: 3 * (o + n)], csum)
        drive = F * _forcing_factor(torch, csum, s ** 3, tau)
        dj = C.T @ (after[w] - before[w])
        scale = float(B.nlive[w]) * np.abs(F).max()  # the body force on the window per step
        worst_j = max(worst_j, float(np.abs(dj - (imb + drive)).max()) / scale)
        worst_m = max(worst_m, abs(float(after[w].sum() - before[w].sum())) / scale)
        worst_c = max(worst_c, float(np.abs(loads[w][:, :3].sum(axis=0) + imb).max()) / scale)
    print(f"  momentum budget {worst_j:.2e}, mass {worst_m:.2e}, load closure {worst_c:.2e} "
          "(of the window's body force)", flush=True)
    # observed 3.7e-8, 2.5e-9 and 1.1e-15
    assert worst_j <= 1e-6
    assert worst_m <= 1e-7
    assert worst_c <= 1e-12


def test_moving_window_full_size_momentum_budget_and_fp64_loads(psc3d):
    torch = importlib.import_module("torch")
    workloads = load_package("workloads")
    n, tau, s, warm = 512, 0.8, 3, 3
    base = workloads.moving_window(seed=0, n=n)

    def solver(precision):
        return psc3d.SparsePscD3Q19(n, base.copy(), tau=tau, subsamples=s, static_geometry=False,
                                    precision=precision, periodic_geometry=True, device="cuda:0")

    print(f"[moving full size] {n}^3, {len(base)} spheres, FP32 owner table", flush=True)
    s32 = solver("fp32")
    table = base.copy()
    for _ in range(warm):
        workloads.advance_centres(table, n)
        s32.particles = table
        s32.run(1)
    assert s32._geom_mode == "owners"
    N = n ** 3
    ps = int(s32._a.numel() // 19)
    before = _plane_sums(torch, s32._a, 0, ps, N)
    workloads.advance_centres(table, n)
    s32.particles = table
    s32.run(1)
    torch.cuda.synchronize()
    after = _plane_sums(torch, s32._a, 0, ps, N)
    imb = _imb_sum(torch, s32._imb, s32._csum)
    loads32 = s32._loads_np().copy()
    scale = float(np.abs(loads32[:, :3]).sum())  # total load magnitude, F = 0
    dj = C.T @ (after - before)
    ej = float(np.abs(dj - imb).max()) / scale
    em = abs(float(after.sum() - before.sum())) / scale
    ec = float(np.abs(loads32[:, :3].sum(axis=0) + imb).max()) / scale
    print(f"  momentum budget {ej:.2e}, mass {em:.2e}, load closure {ec:.2e} (of sum |F_p|)", flush=True)
    assert ej <= 1e-8 and em <= 1e-7 and ec <= 1e-12  # observed 1.1e-10, 6.2e-10, 1.3e-17
    del s32
    torch.cuda.empty_cache()

    print("  FP64 run of the same trajectory", flush=True)
    s64 = solver("fp64")
    table = base.copy()
    for _ in range(warm + 1):
        workloads.advance_centres(table, n)
        s64.particles = table
        s64.run(1)
    loads64 = s64._loads_np().copy()
    del s64
    torch.cuda.empty_cache()
    Fs = np.linalg.norm(loads64[:, :3], axis=1)
    eF = np.linalg.norm(loads32[:, :3] - loads64[:, :3], axis=1) / np.maximum(Fs, 1e-3 * Fs.max())
    Ts = np.linalg.norm(loads64[:, 3:], axis=1)
    eT = np.linalg.norm(loads32[:, 3:] - loads64[:, 3:], axis=1) / np.maximum(Ts, 1e-3 * Ts.max())
    print(f"  per-particle FP32 vs FP64: force max {eF.max():.2e} median {np.median(eF):.2e}, "
          f"torque max {eT.max():.2e}", flush=True)
    assert eF.max() <= 1e-4 and eT.max() <= 1e-4  # observed 2.4e-6 and 5.8e-6


def test_gradient_full_size_matches_central_differences():
    adj = load_package("adjoint")
    workloads = load_package("workloads")
    n, T, F = 96, 2000, (0.0, 0.0, 1.0e-6)
    tab = workloads.adjoint_window(seed=0)
    print(f"[gradient full size] {n}^3, {len(tab)} particles, {T} steps", flush=True)
    p32 = adj.PermeabilityAdjoint(n, tau=0.8, body_force=F, steps=T, precision="fp32", device="cuda:0")
    k32, g32 = p32.value_and_grad(tab)
    p64 = adj.PermeabilityAdjoint(n, tau=0.8, body_force=F, steps=T, precision="fp64", device="cuda:0")
    k64 = p64.forward(tab)
    assert abs(k32 - k64) / abs(k64) <= 1e-5
    g32 = np.asarray(g32)
    worst = 0.0
    for flat in np.argsort(-np.abs(g32).ravel())[:3]:
        pi, j = divmod(int(flat), 4)
        h = 1.0e-3
        tp, tm = tab.copy(), tab.copy()
        tp[pi, j] += h
        tm[pi, j] -= h
        fd = (p64.forward(tp) - p64.forward(tm)) / (2 * h)
        err = abs(float(g32[pi, j]) - fd) / abs(fd)
        print(f"  particle {pi} d/d{'xyzr'[j]}: adjoint {g32[pi, j]:.9e} FD {fd:.9e} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    print(f"  k32 {k32:.9e} k64 {k64:.9e}", flush=True)
    assert worst <= 1e-4  # observed <= 1.5e-6
