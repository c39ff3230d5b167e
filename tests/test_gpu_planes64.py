"""FP32 state planes padded to 64 cells (PSC_STEP_PLANES64), GPU.

On a lattice whose cell count N is not a multiple of 64 the FP32 solver
keeps each of the 19 direction planes at N rounded up to 64 cells, so that
every plane starts on a 256-byte boundary.  The layout moves addresses only:
the steps (the native static loop with its fused loads, in-place step
pairs, moving spheres with the owner table and with numerator planes), the
moments and the f import must give the bits of the same calls on a state of
unpadded N-cell planes.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N37 = 37  # 37^3 = 50653 cells = 791 * 64 + 29


def _spheres(rng, n, count, umax):
    t = np.zeros((count, 10))
    t[:, 0:3] = rng.uniform(0, n, (count, 3))
    t[:, 3] = rng.uniform(2.5, 5.5, count)
    t[:, 4:7] = rng.uniform(-umax, umax, (count, 3))
    t[:, 7:10] = rng.uniform(-umax, umax, (count, 3)) / t[:, 3:4]
    return t


def _lib():
    from tests.conftest import load_package
    return load_package("_lib")


def _static(psc3d, seed):
    rng = np.random.default_rng(seed)
    s = psc3d.SparsePscD3Q19(N37, _spheres(rng, N37, 6, 0.0), tau=0.8, subsamples=3, body_force=(0, 0, 2e-6),
                             static_geometry=True, periodic_geometry=True, precision="fp32", device="cuda:0",
                             block=(64, 4))
    lib = _lib()
    assert s._state_flags == lib.PSC_STEP_PLANES64
    assert s._ps % 64 == 0 and s._N < s._ps < s._N + 64
    assert s._a.numel() == 19 * s._ps
    return s


@pytest.mark.parametrize("inplace", [False, True])
def test_static_run_equals_unpadded_steps(psc3d, inplace):
    """Nine steps of the native loop (fused loads, graphs; in place in pairs
    or not) on padded planes against nine psc_step calls on N-cell planes."""
    import torch
    lib = _lib()
    print(f"[planes64] static {N37}^3 inplace={inplace}", flush=True)
    s = _static(psc3d, 41)
    s.inplace = inplace
    s.rebuild_geometry()
    a = s._compact_state().clone()
    b = torch.empty_like(a)
    imb = torch.zeros_like(s._imb)
    s.run(9)
    assert s.loads_mode == "fused"
    h = torch.cuda.current_stream().cuda_stream
    fx, fy, fz = s.body_force
    for _ in range(9):
        lib.call("psc_step", a.data_ptr(), b.data_ptr(), s._csum.data_ptr(), None, imb.data_ptr(), lib.PSC_F32_DEV,
                 lib.PSC_PULL, N37, N37, N37, s._s, s.tau, fx, fy, fz, 64, 4, 0, h)
        a, b = b, a
    assert torch.equal(a, s._compact_state())
    assert torch.equal(imb, s._imb)
    loads = s._loads_np().copy()
    s._imb = imb  # the loads of the unpadded run's last imb
    s.accumulate_particle_loads()
    assert np.array_equal(loads, s._loads_np())


@pytest.mark.parametrize("owner_table", [True, False])
def test_moving_step_equals_unpadded(psc3d, owner_table):
    """A moving step (owner table, or numerator planes) on padded state
    planes against the same call on N-cell planes: the solid velocity keeps
    its N-cell planes in both."""
    import torch
    lib = _lib()
    print(f"[planes64] moving {N37}^3 owner_table={owner_table}", flush=True)
    rng = np.random.default_rng(42)
    t = _spheres(rng, N37, 8, 2e-3)
    s = psc3d.SparsePscD3Q19(N37, t, tau=0.8, subsamples=3, periodic_geometry=True, precision="fp32",
                             device="cuda:0", block=(64, 4), owner_table=owner_table)
    assert s._state_flags == lib.PSC_STEP_PLANES64
    for _ in range(3):
        t[:, 0:3] = np.mod(t[:, 0:3] + t[:, 4:7], N37)
        s.particles = t
        s.collide_stream_psc()
    assert s._geom_mode == ("owners" if owner_table else "planes")
    a = s._compact_state().clone()
    b = torch.empty_like(a)
    imb = torch.zeros_like(s._imb)
    h = torch.cuda.current_stream().cuda_stream
    if owner_table:
        lib.call("psc_step_owners", a.data_ptr(), b.data_ptr(), s._csum.data_ptr(), s._own.data_ptr(),
                 s._us.data_ptr(), s._motion.data_ptr(), len(t), imb.data_ptr(), N37, N37, N37, s._s, s.tau, 0.0, 0.0,
                 0.0, 64, 4, 0, h)
    else:
        lib.call("psc_step", a.data_ptr(), b.data_ptr(), s._csum.data_ptr(), s._us.data_ptr(), imb.data_ptr(),
                 lib.PSC_F32_DEV, lib.PSC_PULL, N37, N37, N37, s._s, s.tau, 0.0, 0.0, 0.0, 64, 4, 0, h)
    csum = s._csum.cpu().numpy()
    assert (csum != 0).sum() > 0
    s.collide_stream_psc()  # same particles: the same geometry, re-voxelized
    assert torch.equal(b, s._compact_state())
    solid = csum.repeat(3) != 0
    assert np.array_equal(imb.cpu().numpy()[solid], s._imb.cpu().numpy()[solid])


def test_moments_and_f_import_of_padded_planes(psc3d):
    """psc_moments reads padded planes with pitch_z = nz: the sums and fields
    of an unpadded copy, bit for bit; f written through the solver lands in
    the padded planes as psc_import_f64 writes N-cell planes."""
    import torch
    lib = _lib()
    print(f"[planes64] moments {N37}^3", flush=True)
    s = _static(psc3d, 43)
    s.run(6)
    sums = s.readout_sums()
    fields = np.array([s.rho, s.ux, s.uy, s.uz])
    c = s._compact_state().clone()
    h = torch.cuda.current_stream().cuda_stream
    rho_u = torch.empty(4 * s._N, dtype=torch.float64, device="cuda:0")
    out = torch.empty(lib.PSC_NSUM, dtype=torch.float64, device="cuda:0")
    lib.call("psc_moments", c.data_ptr(), lib.PSC_F32_DEV, lib.PSC_PULL, s._csum.data_ptr(), s._imb.data_ptr(), N37,
             N37, N37, 0, s._s, *s.body_force, rho_u.data_ptr(), s._mom_partials.data_ptr(), out.data_ptr(), h)
    ref = out.cpu().numpy()
    assert np.array_equal(ref[2:5], sums["u"]) and np.array_equal(ref[5:8], sums["fluid_u"])
    assert ref[11] == sums["rho"] and ref[0] == sums["solid"] and np.array_equal(ref[8:11], sums["imb"])
    assert np.array_equal(rho_u.cpu().numpy().reshape(fields.shape), fields)
    # f import: a perturbed f through the solver against psc_import_f64 on N-cell planes
    f = s.f.copy()
    f[:, 3:9, :, 5] *= 1.0 + 1e-3
    s.f = f
    want = torch.empty_like(c)
    src = torch.from_numpy(np.ascontiguousarray(f.reshape(-1))).to("cuda:0")
    lib.call("psc_import_f64", src.data_ptr(), lib.PSC_F32_DEV, lib.PSC_PULL, N37, N37, N37, want.data_ptr(), h)
    torch.cuda.synchronize()
    assert torch.equal(want, s._compact_state())
    assert not torch.equal(want, c)
