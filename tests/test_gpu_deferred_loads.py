"""collide_stream_psc with the particle loads left pending
(psc_run_steps_deferred): each step is one launch that also reduces the
loads the steps before it still owe, and the last step's loads are reduced
at their first read (psc_flush_loads).  Where the steps run in place in pairs,
a step is held until the next call and the two launch together; a read
launches a held step on its own first.  Against a solver that steps and
reduces the loads after every step as the reference does (psc3d.py:208-281;
defer_loads=False), every read -- forces, torques, the diag dict, imb, f,
rho/u, the FP64 sums -- gives the same bits: reads after every step, after
long gaps of even and odd length, across native runs, explicit reductions,
rebuilds, f writes and checkpoints, on plain and padded (N % 64 != 0) planes
and with the tuner's flag forms."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,block,flags,period,defer,inplace", [
    (40, (64, 4), 0x00000, 1, None, None), (40, (64, 4), 0x00000, 1, None, True),
    (52, (128, 1), 0x10601, 3, True, False), (52, (128, 1), 0x10601, 3, True, True),
    (40, (256, 1), 0x100601, 5, True, True), (37, (64, 4), 0x00000, 1, True, False),
    (37, (64, 2), 0x00000, 1, True, True)])
def test_deferred_loads_match_reduction_after_every_step(psc3d, n, block, flags, period, defer, inplace):
    rng = np.random.default_rng(5)
    spheres = [(*rng.uniform(0, n, 3), rng.uniform(2.5, 6.0)) for _ in range(14)]

    def make(defer_loads):
        s = psc3d.SparsePscD3Q19(n, [psc3d.Particle3D(*p) for p in spheres], tau=0.8, subsamples=3,
                                 body_force=(1e-5, -1e-5, 2e-5), static_geometry=True, periodic_geometry=True,
                                 precision="fp32", device="cuda:0", block=block, chunk_cells=256)
        s._flags = flags
        s.fuse_period = period
        s.defer_loads = defer_loads
        s.inplace = inplace
        return s

    a, b = make(False), make(defer)
    print(f"[deferred] n {n} block {block} flags {flags:#x} period {period} defer {defer} inplace {inplace}",
          flush=True)
    pairs = bool(inplace)

    def step(k):
        for _ in range(k):
            a.collide_stream_psc()
            b.collide_stream_psc()

    def same_loads():
        assert np.array_equal(a.hydro_forces, b.hydro_forces)
        assert np.array_equal(a.hydro_torques, b.hydro_torques)

    def same_state():
        assert np.array_equal(a.f, b.f)
        assert np.array_equal(np.stack([a.imb_mx, a.imb_my, a.imb_mz]), np.stack([b.imb_mx, b.imb_my, b.imb_mz]))

    for _ in range(3):  # reads after every step, the diag dict first
        da, db = a.collide_stream_psc(), b.collide_stream_psc()
        assert da == db
        same_loads()
    assert b._pending is None and b._held is None  # the reads launched and reduced everything
    step(7)  # a gap without reads
    assert (b._held is not None) == pairs  # an odd count holds its last step
    same_state()
    assert b._held is None and b._pending is not None  # f and imb launch it, and need no loads
    same_loads()
    assert a.momentum_residual == b.momentum_residual
    step(8)  # an even gap: every step launched in pairs
    assert b._held is None and b._pending[2] == 3
    same_loads()
    assert np.array_equal(a.uz, b.uz) and np.array_equal(a.rho, b.rho)
    step(3)
    a.run(5)  # a native run starts where the held step leaves the state
    b.run(5)
    same_loads()
    step(1)
    a.accumulate_particle_loads()
    b.accumulate_particle_loads()
    assert b._pending is None and b._held is None
    same_loads()
    same_state()
    step(3)
    a.rebuild_geometry()  # the owed loads are those of the geometry being replaced
    b.rebuild_geometry()
    same_loads()
    step(2)
    assert a.readout_sums()["u"].tolist() == b.readout_sums()["u"].tolist()
    assert a.permeability(axis=2) == b.permeability(axis=2)
    step(1)
    ck = b.state_dict()  # a checkpoint launches the held step
    assert np.array_equal(ck["f"], a.f)
    step(5)
    a.load_state_dict(ck)  # both back to the checkpoint; b holds a step that the write follows
    b.load_state_dict(ck)
    step(3)
    da, db = a.collide_stream_psc(), b.collide_stream_psc()
    assert da["max_torque"] == db["max_torque"] and da["momentum_residual"] == db["momentum_residual"]
    same_state()
    same_loads()
    assert a._step_count == b._step_count == 37  # 33 at the checkpoint, then 4
    assert np.abs(a.hydro_forces).max() > 0
    assert int(a._tables.chunks.shape[0]) > 20


@pytest.mark.parametrize("shape,block,flags", [((20, 28, 36), (64, 1), 0x00000), ((36, 20, 52), (128, 1), 0x100601),
                                               ((24, 16, 3), (64, 4), 0x00301)])
def test_held_pairs_on_non_cubic_lattices(psc3d, shape, block, flags):
    """Held in-place pairs index x, y and z separately (a 3-cell axis where
    z - 1 and z + 2 coincide): odd and even runs of calls, with native runs
    and reads between them, give the bits of a step followed by its loads."""
    rng = np.random.default_rng(9)
    nx, ny, nz = shape
    spheres = [(rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz), rng.uniform(2.0, 5.0)) for _ in range(8)]

    def make(defer):
        s = psc3d.SparsePscD3Q19(shape, [psc3d.Particle3D(*p) for p in spheres], tau=0.8, subsamples=3,
                                 body_force=(1e-5, -2e-5, 3e-5), static_geometry=True, precision="fp32",
                                 device="cuda:0", block=block, chunk_cells=256)
        s._flags = flags
        s.defer_loads, s.inplace = defer, True
        return s

    a, b = make(False), make(True)
    print(f"[held] shape {shape} block {block} flags {flags:#x}", flush=True)
    for calls, native in ((5, 2), (4, 3), (1, 0), (6, 1), (3, 4)):
        for _ in range(calls):
            a.collide_stream_psc()
            b.collide_stream_psc()
        assert (b._held is not None) == bool(calls & 1) or calls == 0
        if native:
            a.run(native)
            b.run(native)
        assert np.array_equal(a.f, b.f)
        assert np.array_equal(a.hydro_forces, b.hydro_forces)
        assert np.array_equal(a.hydro_torques, b.hydro_torques)
        assert np.array_equal(np.stack([a.imb_mx, a.imb_my, a.imb_mz]), np.stack([b.imb_mx, b.imb_my, b.imb_mz]))
    assert np.abs(a.hydro_forces).max() > 0


def test_steps_that_cannot_carry_the_loads_reduce_them_at_once(psc3d):
    """FP64 storage, dynamic geometry, or no particles: nothing is deferred."""
    spheres = [psc3d.Particle3D(8.3, 7.6, 8.1, 3.2)]
    for kw in (dict(precision="fp64", static_geometry=True), dict(precision="fp32", static_geometry=False)):
        s = psc3d.SparsePscD3Q19(16, spheres, tau=0.8, subsamples=2, body_force=(0, 0, 1e-5), device="cuda:0",
                                 **kw)
        s.defer_loads = True
        s.collide_stream_psc()
        s.collide_stream_psc()
        assert s._pending is None, kw
    s = psc3d.SparsePscD3Q19(16, [], tau=0.8, subsamples=2, body_force=(0, 0, 1e-5), static_geometry=True,
                             precision="fp32", device="cuda:0")
    s.defer_loads = True
    s.collide_stream_psc()
    assert s._pending is None and s.hydro_forces.shape == (0, 3)
