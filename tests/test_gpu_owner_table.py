"""The owner-table form of the FP32 moving step (GPU).

Moving spheres re-voxelized every step, FP32 pull storage: the voxelizer
records, per covered cell, the one particle whose sphere covers it (or marks
the cell shared) and accumulates the solid-velocity numerator
sum_p alpha_p (U + omega x r)_p at shared cells only; the step kernel forms a
single owner's U + omega x r itself (psc3d.py:181-206 / :347-373 for the
solid velocity, :208-246 for the step).  Checked against the oracle's
periodic sparse cell lists (the owners and the shared numerators), against
the plane form (owner_table=False) and the FP64 oracle over moving steps
(state to 1e-4 of the signal max|f_ref - w|, loads to 1e-3 of the largest),
and through the solid_velocity readout.
"""

import numpy as np
import pytest

from oracle import psc_oracle as O

pytestmark = pytest.mark.gpu


def _table(rng, n, count, rmin, rmax, umax):
    t = np.zeros((count, 10))
    t[:, 0:3] = rng.uniform(0, n, (count, 3))
    t[:, 3] = rng.uniform(rmin, rmax, count)
    t[:, 4:7] = rng.uniform(-umax, umax, (count, 3))
    t[:, 7:10] = rng.uniform(-umax, umax, (count, 3)) / t[:, 3:4]
    return t


def _oracle_owners(o, n):
    """Owner count, last owner and the numerator sum_p alpha_p v_p per cell
    from the oracle's sparse lists (unwrapped node positions)."""
    owners = np.zeros(n ** 3, dtype=np.int64)
    last = np.full(n ** 3, -1, dtype=np.int64)
    num = np.zeros((3, n ** 3))
    for ip, (p, ids, a, pos) in enumerate(zip(o.particles, o.node_ids, o.alpha_vals, o.node_pos)):
        if len(ids) == 0:
            continue
        owners[ids] += 1
        last[ids] = ip
        rx, ry, rz = pos[0] + 0.5 - p.x, pos[1] + 0.5 - p.y, pos[2] + 0.5 - p.z
        num[0, ids] += a * (p.vx + p.wy * rz - p.wz * ry)
        num[1, ids] += a * (p.vy + p.wz * rx - p.wx * rz)
        num[2, ids] += a * (p.vz + p.wx * ry - p.wy * rx)
    return owners, last, num


def test_owner_table_matches_oracle_cell_lists(psc3d):
    rng = np.random.default_rng(31)
    n = 24
    t = _table(rng, n, 9, 2.5, 6.0, 0.3)   # overlapping spheres, boxes wrapping the periodic lattice
    print(f"[owner table] {n}^3, {len(t)} spheres", flush=True)
    sol = psc3d.SparsePscD3Q19(n, t.copy(), subsamples=3, periodic_geometry=True, precision="fp32",
                               device="cuda:0")
    sol.rebuild_geometry()
    assert sol._geom_mode == "owners"
    o = O.OracleSolver(n, [O.Sphere(*r) for r in t], subsamples=3, layout="sparse", periodic_geometry=True)
    o.rebuild_geometry()
    owners, last, num = _oracle_owners(o, n)
    csum = sol._csum.cpu().numpy()
    own = sol._own.cpu().numpy().astype(np.int64)
    single, shared = owners == 1, owners > 1
    print(f"  covered {np.count_nonzero(owners)}, shared {np.count_nonzero(shared)}", flush=True)
    assert np.count_nonzero(shared) > 50 and np.count_nonzero(single) > 1000
    assert np.array_equal(csum > 0, owners > 0)
    assert np.array_equal(own[single], last[single])
    assert np.all(own[shared] == -1)   # 0xffffffff read as int32
    us = sol._us.cpu().numpy().astype(np.float64).reshape(3, -1)
    # FP32 terms (each rounded to 2^-24 relative) summed with FP32 atomics: a few ulps of the largest
    # term, absolute, where terms of opposite sign cancel
    tol = 4 * 2.0 ** -24 * np.max(np.abs(num[:, shared]))
    np.testing.assert_allclose(us[:, shared], num[:, shared], rtol=2e-6, atol=tol)
    assert np.all(us[:, ~shared] == 0.0)
    # the readout voxelizes the full planes: the oracle's solid velocity
    for got, ref in zip(sol.solid_velocity(), o.solid_velocity()):
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-9)


def test_owner_table_step_matches_planes_and_oracle(psc3d):
    rng = np.random.default_rng(32)
    n, steps = 24, 12
    t = _table(rng, n, 9, 2.5, 6.0, 2e-3)
    kw = dict(tau=0.8, subsamples=3, body_force_z=1e-6, periodic_geometry=True, precision="fp32", device="cuda:0")
    ta, tb = t.copy(), t.copy()
    A = psc3d.SparsePscD3Q19(n, ta, **kw)
    B = psc3d.SparsePscD3Q19(n, tb, owner_table=False, **kw)
    parts = [O.Sphere(*row) for row in t]
    o = O.OracleSolver(n, parts, tau=0.8, subsamples=3, body_force=(0, 0, 1e-6), layout="sparse",
                       periodic_geometry=True)
    print(f"[owner step] {n}^3, {len(t)} moving spheres, {steps} steps", flush=True)
    for _ in range(steps):
        for tt in (ta, tb):
            tt[:, 0:3] = np.mod(tt[:, 0:3] + tt[:, 4:7], n)
        for p, row in zip(parts, ta):
            p.x, p.y, p.z = row[0], row[1], row[2]
        A.particles = ta
        B.particles = tb
        A.collide_stream_psc()
        B.collide_stream_psc()
        o.collide_stream_psc()
    assert A._geom_mode == "owners" and B._geom_mode == "planes"
    signal = np.max(np.abs(o.f - O.W[:, None, None, None]))
    ea, eb = np.max(np.abs(A.f - o.f)), np.max(np.abs(B.f - o.f))
    eab = np.max(np.abs(A.f - B.f))
    scale = np.max(np.abs(o.hydro_forces))
    fa = np.max(np.abs(A.hydro_forces - o.hydro_forces)) / scale
    fab = np.max(np.abs(A.hydro_forces - B.hydro_forces)) / scale
    print(f"  signal {signal:.2e}: owners vs oracle {ea / signal:.2e}, planes vs oracle {eb / signal:.2e}, "
          f"owners vs planes {eab / signal:.2e}; forces owners vs oracle {fa:.2e}, vs planes {fab:.2e}", flush=True)
    assert ea <= 1e-4 * signal and eb <= 1e-4 * signal
    assert eab <= 1e-5 * signal
    assert fa <= 1e-3 and fab <= 1e-4
    for ua, ub in zip(A.solid_velocity(), B.solid_velocity()):
        np.testing.assert_allclose(ua, ub, rtol=1e-6, atol=1e-10)


def test_owner_table_loads_as_accurate_as_planes(psc3d):
    """Per-particle FP32 loads against the FP64 solver after 600 steps of the
    moving-window workload at 128^3: the owner table is as accurate as the
    plane form.  The velocity is rounded to FP32 once per cell; rounding U,
    omega or the centre to FP32 once per particle instead gives every cell of
    a sphere the same error, which adds up in its load (about ten times the
    plane form's error on the particles with small loads)."""
    from tests.conftest import load_package
    wl = load_package("workloads")
    n, steps = 128, 600
    base = wl.moving_window(seed=0, n=n, count=int(round(2000 * (n / 512) ** 3)))
    print(f"[owner accuracy] {n}^3, {len(base)} spheres, {steps} steps", flush=True)

    def loads(precision, owner):
        t = base.copy()
        s = psc3d.SparsePscD3Q19(n, t, tau=0.8, subsamples=3, precision=precision, periodic_geometry=True,
                                 device="cuda:0", owner_table=owner)
        for _ in range(steps):
            wl.advance_centres(t, n)
            s.particles = t
            s.run(1)
        return s._loads_np()[:, 0:3].copy()

    F64 = loads("fp64", None)
    mag = np.linalg.norm(F64, axis=1)
    floor = np.maximum(mag, 1e-3 * mag.max())
    err = {}
    for owner in (True, False):
        e = np.linalg.norm(loads("fp32", owner) - F64, axis=1) / floor
        err[owner] = (float(e.max()), float(np.median(e)))
    print(f"  force error vs FP64 (max, median): owners {err[True]}, planes {err[False]}", flush=True)
    assert err[True][0] <= 2.0 * err[False][0] + 1e-9
    assert err[True][1] <= 2.0 * err[False][1] + 1e-10


def test_owner_table_step_same_bits_tiled_and_flat(psc3d):
    """The moving owner-table step gives the same bits with every cell
    mapping the tuner may pick ((z, y) tiles, two cells per thread, the flat
    mapping over the whole lattice), on a lattice whose row length (37 cells)
    splits 64-byte granules."""
    import torch
    from tests.conftest import load_package
    lib = load_package("_lib")
    rng = np.random.default_rng(33)
    n = 37
    t = _table(rng, n, 8, 2.5, 6.0, 2e-3)
    sol = psc3d.SparsePscD3Q19(n, t, tau=0.8, subsamples=3, periodic_geometry=True, precision="fp32",
                               device="cuda:0", block=(64, 4))
    for _ in range(3):
        t[:, 0:3] = np.mod(t[:, 0:3] + t[:, 4:7], n)
        sol.particles = t
        sol.collide_stream_psc()
    assert sol._geom_mode == "owners"
    h = torch.cuda.current_stream().cuda_stream
    ref = None
    for bz, by, flags in [(64, 4, 0), (64, 2, 0x20000), (256, 1, 0x10401), (64, 1, lib.PSC_STEP_LINEAR),
                          (256, 1, lib.PSC_STEP_LINEAR | 0x401), (128, 2, lib.PSC_STEP_LINEAR)]:
        lib.call("psc_step_owners", sol._a.data_ptr(), sol._b.data_ptr(), sol._csum.data_ptr(),
                 sol._own.data_ptr(), sol._us.data_ptr(), sol._motion.data_ptr(), len(t), sol._imb.data_ptr(), n, n,
                 n, sol._s, sol.tau, 0.0, 0.0, 1e-6, bz, by, flags | sol._state_flags, h)
        out = (sol._b.cpu().numpy().copy(), sol._imb.cpu().numpy()[sol._csum.cpu().numpy().repeat(3) != 0].copy())
        if ref is None:
            ref = out
        assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), (bz, by, hex(flags))


def test_owner_ids_past_the_motion_records_read_the_planes(psc3d):
    """An owner id at or past the motion-record count (an owner table that
    does not belong to the records) is not looked up: such a cell takes its
    velocity from the numerator planes.  With a count of 0 every cell does,
    so the owner-table step equals psc_step on the same planes, bit for bit."""
    import torch
    from tests.conftest import load_package
    lib = load_package("_lib")
    rng = np.random.default_rng(34)
    n = 32
    t = _table(rng, n, 8, 2.5, 6.0, 2e-3)
    sol = psc3d.SparsePscD3Q19(n, t, tau=0.8, subsamples=3, periodic_geometry=True, precision="fp32",
                               device="cuda:0", block=(64, 4))
    sol.collide_stream_psc()
    assert sol._geom_mode == "owners"
    h = torch.cuda.current_stream().cuda_stream
    lib.call("psc_step_owners", sol._a.data_ptr(), sol._b.data_ptr(), sol._csum.data_ptr(), sol._own.data_ptr(),
             sol._us.data_ptr(), sol._motion.data_ptr(), 0, sol._imb.data_ptr(), n, n, n, sol._s, sol.tau, 0.0, 0.0,
             1e-6, 64, 4, 0, h)
    got = sol._b.cpu().numpy().copy()
    lib.call("psc_step", sol._a.data_ptr(), sol._b.data_ptr(), sol._csum.data_ptr(), sol._us.data_ptr(),
             sol._imb.data_ptr(), lib.PSC_F32_DEV, lib.PSC_PULL, n, n, n, sol._s, sol.tau, 0.0, 0.0, 1e-6, 64, 4, 0, h)
    assert np.array_equal(got, sol._b.cpu().numpy())
    lib.call("psc_step_owners", sol._a.data_ptr(), sol._b.data_ptr(), sol._csum.data_ptr(), sol._own.data_ptr(),
             sol._us.data_ptr(), sol._motion.data_ptr(), len(t), sol._imb.data_ptr(), n, n, n, sol._s, sol.tau, 0.0,
             0.0, 1e-6, 64, 4, 0, h)
    assert not np.array_equal(got, sol._b.cpu().numpy())  # with the records the single owners move
