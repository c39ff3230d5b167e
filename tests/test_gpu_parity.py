"""HIP path against the oracle and the reference's golden fixtures (GPU).

Tolerances: the FP64 plain kernel follows the reference's operation order,
so f is bit-identical to the reference's; per-particle loads and global sums
use tree reductions and agree to 1e-10 relative.  The FP32 deviation kernel
stores g = f - w with a 2^-24 relative step, so its error in f is measured
against the signal itself, max|f_ref - w| (2.8e-8 to 4.3e-5 on the golden
cases): it must stay below 1e-4 of it (a kernel that left every deviation at
zero would miss by the whole signal), and the loads agree to 1e-3 relative
over <= 100 steps.
"""

import numpy as np
import pytest

from oracle import psc_oracle as O
from tests import _golden as G
from tests.conftest import load_package

pytestmark = pytest.mark.gpu


def _solver(psc3d, g, precision, layout="pull", **kw):
    cls = psc3d.PscD3Q19 if str(g["layout"]) == "dense" else psc3d.SparsePscD3Q19
    parts = [psc3d.Particle3D(*row) for row in g["particles"]]
    s = cls(int(g["n"]), parts, tau=float(g["tau"]), subsamples=int(g["subsamples"]),
            body_force=tuple(float(v) for v in g["force"]), static_geometry=bool(g["static"]),
            precision=precision, stream_layout=layout, device="cuda:0", **kw)
    return s, parts


def _run(s, parts, name, steps):
    F, T, D = [], [], []
    for k in range(steps):
        if name == "sparse16":
            G.move_sparse16(parts, k)
        d = s.collide_stream_psc()
        F.append(s.hydro_forces.copy())
        T.append(s.hydro_torques.copy())
        D.append([d[key] for key in ("solid_volume_lu", "mean_uz", "momentum_residual", "max_force", "max_torque")])
    return np.array(F), np.array(T), np.array(D)


@pytest.mark.parametrize("name", ["smoke12", "moving20", "spin16", "sparse16"])
@pytest.mark.parametrize("layout", ["pull", "push"])
def test_fp64_matches_reference_trajectory(psc3d, name, layout):
    g = G.load(name)
    s, parts = _solver(psc3d, g, "fp64", layout)
    F, T, D = _run(s, parts, name, int(g["steps"]))
    print(f"[{name}/{layout}] running fp64", flush=True)
    # overlap cells with s^3 not a power of two may differ by one ulp (count
    # sum divided once instead of per-particle fractions added)
    np.testing.assert_allclose(s.solid, g["solid"], rtol=0, atol=4e-16)
    if "f" in g:
        err = np.max(np.abs(s.f - g["f"]))
    else:
        err = np.max(np.abs(s.f.ravel()[g["f_idx"]] - g["f_val"]))
    print(f"  max|f - f_ref| = {err:.3e}")
    assert err == 0.0  # bit-exact: same operation order as the reference
    scale = max(np.max(np.abs(g["forces"])), 1e-30)
    ef = np.max(np.abs(F - g["forces"])) / scale
    # torques of symmetric cases are round-off; scale by force * lever arm
    lever = float(np.max(g["particles"][:, 3])) + 2.0
    et = np.max(np.abs(T - g["torques"])) / max(np.max(np.abs(g["torques"])), scale * lever)
    print(f"  rel force err {ef:.3e}, torque err {et:.3e}")
    assert ef <= 1e-10 and et <= 1e-9
    gd = g["diag"]
    np.testing.assert_allclose(D[:, 0], gd[:, 0], rtol=1e-13)
    np.testing.assert_allclose(D[:, 1], gd[:, 1], rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(D[:, 3], gd[:, 3], rtol=1e-10)
    assert np.all(D[:, 2] <= 1e-12 * max(scale, 1e-12) + 1e-17)
    uz = s.uz
    np.testing.assert_allclose(uz, g["uz"], rtol=1e-12, atol=1e-18)
    imb = np.stack([s.imb_mx, s.imb_my, s.imb_mz])
    np.testing.assert_allclose(imb, g["imb"], rtol=1e-11, atol=1e-18)


@pytest.mark.parametrize("name", ["smoke12", "moving20", "spin16", "sparse16"])
@pytest.mark.parametrize("layout", ["pull", "push"])
def test_fp32_deviation_close_to_reference(psc3d, name, layout):
    g = G.load(name)
    s, parts = _solver(psc3d, g, "fp32", layout)
    print(f"[{name}/{layout}] running fp32", flush=True)
    F, T, D = _run(s, parts, name, int(g["steps"]))
    n = int(g["n"])
    if "f" in g:
        err = np.max(np.abs(s.f - g["f"]))
        signal = np.max(np.abs(g["f"] - O.W[:, None, None, None]))
    else:
        err = np.max(np.abs(s.f.ravel()[g["f_idx"]] - g["f_val"]))
        signal = np.max(np.abs(g["f_val"] - O.W[np.asarray(g["f_idx"]) // n ** 3]))
    scale = max(np.max(np.abs(g["forces"])), 1e-30)
    ef = np.max(np.abs(F[-1] - g["forces"][-1])) / scale
    tol = 1e-4 * signal
    print(f"  fp32 max|f - f_ref| = {err:.3e} = {err / signal:.2e} of the signal max|f_ref - w| = {signal:.3e}, "
          f"rel force err {ef:.3e}")
    assert signal > tol  # the rest state f = w misses by the whole signal
    assert err <= tol
    assert ef <= 1e-3


def test_sparse_geometry_lists(psc3d):
    g = G.load("sparse16")
    s, parts = _solver(psc3d, g, "fp64")
    s.rebuild_geometry()
    ids = np.concatenate(s.node_ids)
    vals = np.concatenate(s.alpha_vals)
    assert np.array_equal(ids, g["ids"])
    assert np.array_equal(vals, g["vals"])
    assert [len(a) for a in s.node_ids] == list(g["ids_len"])


def test_perm24_trace(psc3d):
    g = G.load("perm24")
    for precision, tol in (("fp64", 1e-11), ("fp32", 1e-4)):
        s, parts = _solver(psc3d, g, precision)
        ks = []
        for step in range(int(g["steps"])):
            s.collide_stream_psc()
            if (step + 1) % 20 == 0:
                ks.append(s.permeability(axis=2))
        ks = np.array(ks)
        rel = np.max(np.abs(ks - g["k_trace"]) / np.abs(g["k_trace"]))
        print(f"  perm24 {precision}: max rel k err {rel:.3e}")
        assert rel <= tol


def test_voxelizer_counts_random(psc3d):
    rng = np.random.default_rng(5)
    n = 40
    for s_ in (1, 2, 3, 4, 5, 7):
        parts = [psc3d.Particle3D(*rng.uniform(-2, n + 2, 3), rng.uniform(0.3, 6.0)) for _ in range(12)]
        sol = psc3d.SparsePscD3Q19(n, parts, subsamples=s_, static_geometry=True, device="cuda:0")
        sol.rebuild_geometry()
        oparts = [O.Sphere(p.x, p.y, p.z, p.r) for p in parts]
        o = O.OracleSolver(n, oparts, subsamples=s_, layout="sparse")
        o.rebuild_geometry()
        for a, b in zip(sol.node_ids, o.node_ids):
            assert np.array_equal(a, b)
        for a, b in zip(sol.alpha_vals, o.alpha_vals):
            assert np.array_equal(a, b)
        ref_cnt = np.zeros(n ** 3)
        for ids, vals in zip(o.node_ids, o.alpha_vals):
            np.add.at(ref_cnt, ids, np.rint(vals * s_ ** 3))
        got = np.rint(sol.alpha_sum_flat * s_ ** 3)
        assert np.array_equal(got, ref_cnt)


@pytest.mark.parametrize("precision,block,fuse,pipeline,inplace", [
    ("fp64", (64, 4), None, None, None), ("fp64", (64, 4), None, False, None),
    ("fp32", (64, 4), None, None, None), ("fp32", (64, 2), None, None, None), ("fp32", (64, 1), None, None, None),
    ("fp32", (96, 2), None, None, None), ("fp32", (64, 4), False, True, None), ("fp32", (64, 2), False, True, None),
    ("fp32", (64, 4), False, False, None), ("fp32", (64, 4), True, None, True), ("fp32", (64, 1), True, None, True)])
def test_native_runner_and_graph_replays_match_python_steps(psc3d, precision, block, fuse, pipeline, inplace):
    """psc_run_steps (direct launches and hipGraph replays, with eager reads in
    between the replays; load reduction fused into the next step's kernel, in
    line, or pipelined on the side stream; fused steps also in place in pairs
    on one state buffer, with odd counts starting with an ordinary step; with
    64/128/192/256-thread workgroups) gives the same state and loads as
    stepping one call at a time."""
    g = G.load("perm24")
    a, _ = _solver(psc3d, g, precision, block=block)
    b, _ = _solver(psc3d, g, precision, block=block)
    c, _ = _solver(psc3d, g, precision, block=block)
    a.defer_loads = False  # the reference: loads reduced after every step
    b.pipeline_loads = c.pipeline_loads = pipeline
    b.fuse_loads = c.fuse_loads = fuse
    b.inplace = c.inplace = inplace
    for chunk in (16, 16, 5, 32, 1, 2):
        for _ in range(chunk):
            a.collide_stream_psc()
        b.run(chunk, graph=True, graph_steps=8)
        c.run(chunk, graph=False)
        fa = a.f
        assert np.array_equal(fa, b.f)
        assert np.array_equal(fa, c.f)
        assert np.array_equal(a.hydro_forces, b.hydro_forces)
        assert np.array_equal(a.hydro_torques, b.hydro_torques)
        assert np.array_equal(a.hydro_forces, c.hydro_forces)
        assert np.array_equal(a.hydro_torques, c.hydro_torques)
        assert np.array_equal(np.stack([a.imb_mx, a.imb_my, a.imb_mz]), np.stack([b.imb_mx, b.imb_my, b.imb_mz]))
        assert np.array_equal(np.stack([a.imb_mx, a.imb_my, a.imb_mz]), np.stack([c.imb_mx, c.imb_my, c.imb_mz]))
    assert a._step_count == b._step_count == 72


@pytest.mark.parametrize("n,block,flags,period,inplace", [
    (64, (128, 1), 0x10601, 1, False), (64, (192, 1), 0x10601, 5, False), (64, (64, 4), 0x20001, 3, False),
    (64, (96, 2), 0x00301, 2, False), (64, (128, 1), 0x10601, 1001, False), (64, (256, 1), 0x100601, 5, False),
    (52, (128, 1), 0x100301, 3, False), (52, (64, 1), 0x100000, 1, False), (52, (64, 2), 0x10601, 5, False),
    (64, (128, 1), 0x10601, 5, True), (64, (64, 4), 0x00000, 1, True), (52, (256, 1), 0x100601, 5, True),
    (52, (64, 2), 0x10001, 3, True)])
def test_fused_loads_match_separate_reduction(psc3d, n, block, flags, period, inplace):
    """An n^3 lattice with 40 spheres (overlapping; 256-cell chunks, so
    hundreds of chunks and several per particle): the step kernels that also
    reduce earlier steps' loads (with the tuner's flag forms: non-temporal,
    occupancy cap, two cells per thread, flat cell mapping, also on rows of 52
    cells; load workgroups all ahead of the cells, interleaved, or spread over
    more than the grid; steps in place in pairs) give the state, imb and loads
    of separate launches, bit for bit."""
    rng = np.random.default_rng(7)
    parts = [psc3d.Particle3D(*rng.uniform(0, n, 3), rng.uniform(3.0, 7.5)) for _ in range(40)]

    def make():
        s = psc3d.SparsePscD3Q19(n, [psc3d.Particle3D(p.x, p.y, p.z, p.r) for p in parts], tau=0.8, subsamples=3,
                                 body_force=(1e-5, 0.0, 2e-5), static_geometry=True, precision="fp32",
                                 device="cuda:0", block=block, chunk_cells=256)
        s._flags = flags
        return s

    a, b = make(), make()
    a.defer_loads = False  # the reference: loads reduced after every step
    b.fuse_loads, b.fuse_period, b.inplace = True, period, inplace
    print(f"[fused] n {n} block {block} flags {flags:#x} period {period} inplace {inplace}", flush=True)
    for chunk in (1, 3, 24, 9):
        for _ in range(chunk):
            a.collide_stream_psc()
        b.run(chunk, graph=True, graph_steps=8)  # the first call builds the geometry, stepping in Python
        assert np.array_equal(a.f, b.f)
        assert np.array_equal(a.hydro_forces, b.hydro_forces)
        assert np.array_equal(a.hydro_torques, b.hydro_torques)
        assert np.array_equal(np.stack([a.imb_mx, a.imb_my, a.imb_mz]), np.stack([b.imb_mx, b.imb_my, b.imb_mz]))
    assert np.abs(a.hydro_forces).max() > 0
    assert int(a._tables.chunks.shape[0]) > 100 and int(np.diff(a._tables.chunk_first).max()) > 4


@pytest.mark.parametrize("shape,block,flags", [((20, 28, 36), (64, 1), 0x00000), ((36, 20, 52), (128, 1), 0x100601),
                                               ((24, 16, 3), (64, 4), 0x00301), ((16, 40, 24), (256, 1), 0x100000)])
def test_inplace_pairs_on_non_cubic_lattices(psc3d, shape, block, flags):
    """In-place step pairs index x, y and z separately (gather-scatter to
    x + c_i, periodic wrap, a 3-cell axis where z - 1 and z + 2 coincide):
    on non-cubic lattices the fused in-place loop gives the state, imb and
    loads of Python stepping, bit for bit, also with single steps taken
    between native runs."""
    rng = np.random.default_rng(3)
    nx, ny, nz = shape
    parts = [psc3d.Particle3D(rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz), rng.uniform(2.0, 5.0))
             for _ in range(10)]

    def make():
        s = psc3d.SparsePscD3Q19(shape, [psc3d.Particle3D(p.x, p.y, p.z, p.r) for p in parts], tau=0.8,
                                 subsamples=3, body_force=(1e-5, -2e-5, 3e-5), static_geometry=True,
                                 precision="fp32", device="cuda:0", block=block, chunk_cells=256)
        s._flags = flags
        return s

    a, b = make(), make()
    a.defer_loads = False  # the reference: loads reduced after every step
    b.fuse_loads, b.inplace = True, True
    print(f"[inplace] shape {shape} block {block} flags {flags:#x}", flush=True)
    for chunk in (2, 3, 18, 7, 1):
        for _ in range(chunk):
            a.collide_stream_psc()
        b.run(chunk, graph=True, graph_steps=8)
        assert np.array_equal(a.f, b.f)
        assert np.array_equal(a.hydro_forces, b.hydro_forces)
        assert np.array_equal(a.hydro_torques, b.hydro_torques)
        assert np.array_equal(np.stack([a.imb_mx, a.imb_my, a.imb_mz]), np.stack([b.imb_mx, b.imb_my, b.imb_mz]))
        # a single step between native runs continues from the pull form the loop left
        da, db = a.collide_stream_psc(), b.collide_stream_psc()
        assert np.array_equal(a.f, b.f) and da["mean_uz"] == db["mean_uz"]
    assert b.loads_mode == "fused" and np.abs(a.hydro_forces).max() > 0


def test_fused_tuning_leaves_the_solver_unchanged(psc3d, tmp_path, monkeypatch):
    """Timing the launch candidates in the fused native loop steps scratch
    copies: the solver's state and loads are untouched, the tuned run gives
    the bits of a fixed-launch run, and the choice is cached under its own key."""
    import torch
    monkeypatch.setenv("PSC_HIP_TUNE_CACHE", str(tmp_path / "tune.json"))
    at = load_package("autotune")
    at._MEM.clear()
    rng = np.random.default_rng(11)
    parts = [(*rng.uniform(0, 40, 3), rng.uniform(3.0, 6.0)) for _ in range(12)]

    def make(block):
        return psc3d.SparsePscD3Q19(40, [psc3d.Particle3D(*p) for p in parts], tau=0.8, subsamples=3,
                                    body_force=(1e-5, 0.0, 0.0), static_geometry=True, precision="fp32",
                                    device="cuda:0", block=block)

    a, b = make((64, 4)), make("auto")
    a.run(3)  # builds the geometry, stepping in Python
    b.run(3)  # the same; b tunes in the fused loop on its first step
    assert b._load_modes()[0] and not b._autotune
    assert any(k.endswith("|fused") for k in at._MEM)
    print(f"[tune] fused choice {b._block} {b._flags:#x}", flush=True)
    at._MEM.clear()
    (tmp_path / "tune.json").unlink()
    b._autotune = True
    b.hydro_forces  # the loads of b's last step, which it left pending
    before = (b._a.clone(), b._b.clone(), b._loads.clone())
    b._tune()  # both caches emptied: times every candidate again
    torch.cuda.synchronize()
    assert torch.equal(before[0], b._a) and torch.equal(before[2], b._loads)
    a.run(40)
    b.run(40)
    assert b.loads_mode == "fused"
    assert np.array_equal(a.f, b.f)
    assert np.array_equal(a.hydro_forces, b.hydro_forces)
    assert np.array_equal(a.hydro_torques, b.hydro_torques)


def test_every_launch_configuration_gives_identical_results(psc3d):
    """The autotuner may pick any candidate: all must compute the same bits."""
    import torch
    at = load_package("autotune")
    lib = load_package("_lib")
    g = G.load("perm24")
    s, _ = _solver(psc3d, g, "fp32", block=(64, 4))
    s.run(7)
    ref = None
    nx, ny, nz = s._shape
    h = torch.cuda.current_stream().cuda_stream
    flat = [(bz, 1, nt, 0, True) for bz in at.FLAT for nt in (0, 1)] + [(64, 4, 1, 3, True)]
    for bz, by, nt, bpc, lin in at.candidates(nz) + [(32, 8, 1, 3, False), (16, 16, 0, 0, False)] + flat:
        for cpt in ((1,) if lin else (1, 2)):
            lib.call("psc_step", s._a.data_ptr(), s._b.data_ptr(), s._csum.data_ptr(), None, s._imb.data_ptr(),
                     1, 1, nx, ny, nz, s._s, s.tau, 0.0, 0.0, float(g["force"][2]), bz, by,
                     at.flags_of(nt, bpc, cpt, lin), h)
            out = s._b.cpu().numpy().copy()
            if ref is None:
                ref = out
            assert np.array_equal(out, ref), (bz, by, nt, bpc, cpt, lin)


def test_every_fp64_launch_gives_identical_results(psc3d):
    """The tuner also picks the FP64 step's workgroup shape and whether it
    streams non-temporally: every launch it may pick computes the same bits
    (the reference's operation order per cell, no cross-cell arithmetic)."""
    at = load_package("autotune")
    lib = load_package("_lib")
    import torch
    g = G.load("perm24")
    s, _ = _solver(psc3d, g, "fp64", block=(64, 4))
    s.run(7)
    nx, ny, nz = s._shape
    h = torch.cuda.current_stream().cuda_stream
    ref = None
    for bz, by in at.SHAPES + at.FP64_SHAPES:
        for flags in (0, lib.PSC_STEP_NONTEMPORAL):
            lib.call("psc_step", s._a.data_ptr(), s._b.data_ptr(), s._csum.data_ptr(), None, s._imb.data_ptr(),
                     lib.PSC_F64_PLAIN, 1, nx, ny, nz, s._s, s.tau, 0.0, 0.0, float(g["force"][2]), bz, by, flags, h)
            out = s._b.cpu().numpy().copy()
            if ref is None:
                ref = out
            assert np.array_equal(out, ref), (bz, by, flags)


def _moving_table(rng, n, count, rmin, rmax, umax=2e-3):
    t = np.zeros((count, 10))
    t[:, 0:3] = rng.uniform(0, n, (count, 3))
    t[:, 3] = rng.uniform(rmin, rmax, count)
    t[:, 4:7] = rng.uniform(-umax, umax, (count, 3))
    t[:, 7:10] = rng.uniform(-umax, umax, (count, 3)) / t[:, 3:4]
    return t


def test_moving_periodic_fp64_matches_oracle(psc3d):
    """Moving, overlapping spheres on a periodic lattice (boxes wrap), centres
    advanced every step: FP64 kernel vs the oracle's periodic sparse mode.
    Overlap cells accumulate the solid-velocity numerator with atomics, so f
    agrees to round-off rather than bit for bit."""
    rng = np.random.default_rng(21)
    n, steps = 24, 8
    t = _moving_table(rng, n, 7, 2.0, 5.0)
    tg = t.copy()
    sol = psc3d.SparsePscD3Q19(n, tg, tau=0.9, subsamples=3, body_force_z=1e-6, periodic_geometry=True,
                               precision="fp64", device="cuda:0")
    parts = [O.Sphere(*row) for row in t]
    o = O.OracleSolver(n, parts, tau=0.9, subsamples=3, body_force=(0, 0, 1e-6), layout="sparse",
                       periodic_geometry=True)
    for _ in range(steps):
        tg[:, 0:3] = np.mod(tg[:, 0:3] + tg[:, 4:7], n)
        for p, row in zip(parts, tg):
            p.x, p.y, p.z = row[0], row[1], row[2]
        sol.particles = tg
        sol.collide_stream_psc()
        o.collide_stream_psc()
    np.testing.assert_allclose(sol.solid, o.solid, rtol=0, atol=1e-15)
    assert np.max(np.abs(sol.f - o.f)) < 1e-14
    scale = np.max(np.abs(o.hydro_forces))
    assert np.max(np.abs(sol.hydro_forces - o.hydro_forces)) <= 1e-10 * scale
    assert np.max(np.abs(sol.hydro_torques - o.hydro_torques)) <= 1e-9 * scale * 8


@pytest.mark.parametrize("n", [40, 37])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_incremental_clear_leaves_no_stale_counts(psc3d, n, precision):
    """After many re-voxelizations of moving spheres, count_sum and the
    solid-velocity numerator equal those of a fresh voxelization.  40^3 takes
    the clear by whole 16-cell groups (16 | N), 37^3 the clear by cells."""
    rng = np.random.default_rng(22)
    print(f"[incremental clear] {n}^3 {precision}", flush=True)
    t = _moving_table(rng, n, 25, 2.0, 6.0, umax=0.3)   # large steps: boxes move by cells
    a = psc3d.SparsePscD3Q19(n, t, subsamples=2, periodic_geometry=True, precision=precision, device="cuda:0")
    for _ in range(12):
        t[:, 0:3] = np.mod(t[:, 0:3] + t[:, 4:7], n)
        a.particles = t
        a.collide_stream_psc()
    b = psc3d.SparsePscD3Q19(n, t.copy(), subsamples=2, periodic_geometry=True, precision=precision, device="cuda:0")
    b.rebuild_geometry()
    csum = a._csum.cpu().numpy()
    assert np.array_equal(csum, b._csum.cpu().numpy())
    assert np.count_nonzero(csum) > 0.05 * n ** 3
    # FP32/FP64 atomics add a cell's terms in any order: a few ulps of the largest value, absolute
    ub = b._us.cpu().numpy()
    ulp = 2.0 ** -24 if precision == "fp32" else 2.0 ** -53
    np.testing.assert_allclose(a._us.cpu().numpy(), ub, rtol=1e-6, atol=4 * ulp * np.max(np.abs(ub)))
    assert a._geom_mode == b._geom_mode == ("owners" if precision == "fp32" else "planes")
    if a._geom_mode == "owners":   # owner slots are meaningful where count_sum > 0
        covered = csum != 0
        assert np.array_equal(a._own.cpu().numpy()[covered], b._own.cpu().numpy()[covered])


def test_moving_fp32_loads_track_fp64(psc3d):
    rng = np.random.default_rng(23)
    n, steps = 64, 30
    t = _moving_table(rng, n, 40, 3.0, 7.0, umax=1e-3)
    res = {}
    for prec in ("fp32", "fp64"):
        tt = t.copy()
        s = psc3d.SparsePscD3Q19(n, tt, tau=0.8, subsamples=3, periodic_geometry=True, precision=prec,
                                 device="cuda:0")
        for _ in range(steps):
            tt[:, 0:3] = np.mod(tt[:, 0:3] + tt[:, 4:7], n)
            s.particles = tt
            s.run(1)
        res[prec] = (s.hydro_forces.copy(), s.hydro_torques.copy())
    F32, T32 = res["fp32"]
    F64, T64 = res["fp64"]
    scale = np.max(np.linalg.norm(F64, axis=1))
    err = np.max(np.linalg.norm(F32 - F64, axis=1)) / scale
    print(f"  moving fp32 vs fp64 max force err (rel to max |F|) {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("launch,period", [("auto", 5), ((1 | 0x40000, 0, "rows"), 1), ((0x40000, 64, "spans"), 3),
                                           ((1, 0, "spans"), 2), ((0x200, 0, "rows"), 10001),
                                           ((1 | 0x80000, 0, "spans"), 5), ((0x40000 | 0x80000, 0, "rows"), 3)])
def test_batched_fused_loads_match_separate_reduction(launch, period):
    """Batched windows with the loads reduced inside the step launches
    (psc_step_batched_loads: 256/512-thread workgroups, persistent grid,
    non-temporal imb stores, placement periods) give the state and loads of the side-stream form, bit
    for bit, over runs of 1, 2, 9, 17 and 3 steps (graph replays, direct
    launches and the pending loads at the end of each run)."""
    batch = load_package("batch")
    print(f"[batched fused] launch {launch} period {period}", flush=True)
    rng = np.random.default_rng(5)
    wins = []
    for L in (24, 40, 31):
        t = np.zeros((7, 10))
        t[:, 0:3] = rng.uniform(0, L, (7, 3))
        t[:, 3] = rng.uniform(2.5, 6.0, 7)
        wins.append((L, t))
    kw = dict(tau=0.8, subsamples=3, body_force=(1e-6, 0, 2e-6), device="cuda:0", launch=launch, chunk_cells=256)
    A = batch.BatchedPscD3Q19(wins, **kw)
    B = batch.BatchedPscD3Q19(wins, **kw)
    A.fuse_loads = False
    B.fuse_period = period
    for chunk in (1, 2, 9, 17, 3):
        A.run(chunk)
        B.run(chunk)
        for w in range(len(wins)):
            assert np.array_equal(A.window_state(w), B.window_state(w)), (chunk, w)
        la, lb = A.loads(), B.loads()
        for w in range(len(wins)):
            assert np.array_equal(la[w], lb[w]), (chunk, w)
    assert B.C > 3 * B.P  # several chunks per particle
    assert max(np.abs(v).max() for v in B.loads()) > 0


@pytest.mark.parametrize("launch", ["auto", (1, 0, "spans"), (1 | 0x40000, 0, "rows"), (0x200, 0, "rows"),
                                    (0x40000, 64, "spans"), (1 | 0x80000, 0, "spans")])
@pytest.mark.parametrize("row_align", [1, 32])
def test_batched_windows_match_single_solvers(psc3d, launch, row_align):
    """Every window of a batch (uneven sizes, every workgroup size and item
    mode, persistent grid, graph replays, unpadded rows and rows padded to 32
    cells) equals a standalone FP32 solver of that window bit for bit, state
    and loads, and the per-window permeability matches."""
    batch = load_package("batch")
    print(f"[batched windows] launch {launch} row_align {row_align}", flush=True)
    rng = np.random.default_rng(31)
    wins = []
    for L in (20, 33, 48, 27):
        t = np.zeros((4, 10))
        t[:, 0:3] = rng.uniform(0, L, (4, 3))
        t[:, 3] = rng.uniform(2.0, 4.5, 4)
        wins.append((L, t))
    B = batch.BatchedPscD3Q19(wins, tau=0.8, subsamples=3, body_force=(0, 0, 2e-6), device="cuda:0", launch=launch,
                              row_align=row_align)
    assert B.pitch == [-(-L // row_align) * row_align for L, _ in wins]
    B.run(13)
    singles = []
    for L, t in wins:
        s = psc3d.SparsePscD3Q19(L, t, tau=0.8, subsamples=3, body_force=(0, 0, 2e-6), static_geometry=True,
                                 periodic_geometry=True, precision="fp32", device="cuda:0")
        s.run(13, graph=False)
        singles.append(s)
    loads = B.loads()
    k = B.permeability()
    for w, s in enumerate(singles):
        assert np.array_equal(B.window_state(w), s._compact_state().cpu().numpy().reshape(B.window_state(w).shape)), w
        assert np.array_equal(loads[w], s._loads_np()), w
        assert k[w] == pytest.approx(s.permeability(axis=2), rel=1e-12)


def test_module_equilibrium_and_guo_force_match_oracle(psc3d):
    """The module functions equilibrium / guo_force (psc3d.py:97-114) run on
    the GPU in the reference's operation order: bit-identical to the oracle's
    restatement (itself pinned by the golden trajectories and the moment
    identities), also for broadcast scalars and a z-only force."""
    rng = np.random.default_rng(41)
    shape = (5, 6, 7)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = 0.03 * rng.standard_normal((3,) + shape)
    got = psc3d.equilibrium(rho, *u)
    assert got.shape == (19,) + shape
    assert np.array_equal(got, O.equilibrium(rho, *u))
    assert np.array_equal(psc3d.equilibrium(rho, 0.0, 0.0, u[2]), O.equilibrium(rho, 0.0 * rho, 0.0 * rho, u[2]))
    for F in ((3e-6, -1e-6, 2e-6), (0.0, 0.0, 1e-7)):
        g = psc3d.guo_force(*F, *u, 0.72)
        assert np.array_equal(g, O.guo_force(*F, *u, 0.72)), F
    # the facade is what the reference's init uses: f = equilibrium(1, 0, 0, 0) = W
    w = psc3d.equilibrium(np.ones((2, 2, 2)), 0.0, 0.0, 0.0)
    assert np.array_equal(w[:, 0, 0, 0], psc3d.W)


def test_fluid_statistics_readout_matches_oracle(psc3d):
    """PscD3Q19.fluid_statistics(axis) -- (intrinsic, superficial, porosity)
    of run_precision_validation.py:142-147 from FP64 GPU sums -- against the
    oracle's readout of the oracle's own perm24 run, on every axis (the z
    readout carries the flow, x and y are round-off), and the permeability
    and mean-u_z readouts of the same state."""
    g = G.load("perm24")
    s, _ = _solver(psc3d, g, "fp64")
    parts = [O.Sphere(*row) for row in g["particles"]]
    n = int(g["n"])
    o = O.OracleSolver(n, parts, tau=float(g["tau"]), subsamples=int(g["subsamples"]),
                       body_force=tuple(float(v) for v in g["force"]), static_geometry=True, layout="sparse")
    steps = 120
    s.run(steps)
    for _ in range(steps):
        o.collide_stream_psc()
    fields = (o.ux, o.uy, o.uz)
    for axis in range(3):
        got = s.fluid_statistics(axis=axis)
        ref = O.fluid_statistics(o.solid, fields[axis], n)
        print(f"  axis {axis}: gpu {got} oracle {ref}", flush=True)
        scale = abs(O.fluid_statistics(o.solid, o.uz, n)[0])
        for a, b in zip(got[:2], ref[:2]):
            assert abs(a - b) <= 1e-12 * scale
        assert got[2] == pytest.approx(ref[2], rel=1e-14)
    assert s.fluid_statistics(axis=2)[0] == pytest.approx(O.fluid_weighted_mean(o.solid, o.uz), rel=1e-12)
    assert s.permeability(axis=2) == pytest.approx(O.permeability(o.solid, o.uz, n, o.tau, o.body_force[2]),
                                                   rel=1e-12)
    assert s.mean_uz() == pytest.approx(float(np.mean(o.uz)), rel=1e-12)


@pytest.mark.parametrize("launch", ["auto", (1, 0, "spans"), (1 | 0x40000, 0, "rows")])
def test_batched_degenerate_windows_match_single_solvers(psc3d, launch):
    """Windows of 1, 2 and 3 cells per edge (a cell whose upstream and
    downstream neighbours coincide or are itself; fluid only, since a periodic
    sphere box with its one-cell margin must fit in the window) batched with
    5- and 6-cell windows whose spheres' boxes fill the window or wrap two
    faces equal standalone FP32 solvers of those windows bit for bit, state
    and loads, over graph-replayed fused steps."""
    batch = load_package("batch")
    print(f"[batched degenerate windows] launch {launch}", flush=True)

    def table(rows):
        t = np.zeros((len(rows), 10))
        for i, r in enumerate(rows):
            t[i, :len(r)] = r
        return t

    wins = [(1, table([])),
            (2, table([])),
            (3, table([])),
            (5, table([(2.5, 2.5, 2.5, 0.9)])),
            (6, table([(0.3, 4.6, 2.5, 1.2), (3.5, 1.5, 4.8, 0.8)]))]
    kw = dict(tau=0.8, subsamples=3, body_force=(1e-6, -2e-6, 3e-6), device="cuda:0")
    B = batch.BatchedPscD3Q19(wins, launch=launch, **kw)
    B.run(13)
    loads = B.loads()
    for w, (L, t) in enumerate(wins):
        s = psc3d.SparsePscD3Q19(L, t, static_geometry=True, periodic_geometry=True, precision="fp32", **kw)
        s.run(13, graph=False)
        assert np.array_equal(B.window_state(w), s._compact_state().cpu().numpy().reshape(B.window_state(w).shape)), w
        assert np.array_equal(loads[w], s._loads_np()), w
    assert loads[0].size == 0 and np.abs(loads[3]).max() > 0 and np.abs(loads[4]).max() > 0
