"""psc_moments_noalloc (GPU): the readout sums without a caller workspace,
from a library-owned device array, equal psc_moments' bit for bit -- on two
calls in a row, the second on a smaller lattice of the other storage -- and
reject bad arguments as psc_moments does."""

import numpy as np
import pytest

from tests.conftest import load_package

pytestmark = pytest.mark.gpu


def _solver(psc3d, n, precision, seed):
    rng = np.random.default_rng(seed)
    parts = [psc3d.Particle3D(*rng.uniform(0, n, 3), rng.uniform(2.0, 5.0)) for _ in range(6)]
    s = psc3d.SparsePscD3Q19(n, parts, tau=0.8, subsamples=3, body_force=(1e-6, -2e-6, 3e-6), static_geometry=True,
                             precision=precision, device="cuda:0")
    s.run(7)
    return s


def _both(lib, s):
    """(sums, fields) of psc_moments and of psc_moments_noalloc on one state."""
    import torch
    nx, ny, nz = s._shape
    h = torch.cuda.current_stream().cuda_stream
    out = []
    for name in ("psc_moments", "psc_moments_noalloc"):
        sums = torch.full((lib.PSC_NSUM,), float("nan"), dtype=torch.float64, device="cuda:0")
        rho_u = torch.full((4 * s._N,), float("nan"), dtype=torch.float64, device="cuda:0")
        args = [s._a.data_ptr(), s._storage, s._layout, s._csum.data_ptr(), s._imb.data_ptr(), nx, ny, nz, 0, s._s,
                *s.body_force, rho_u.data_ptr()]
        if name == "psc_moments":
            args.append(s._mom_partials.data_ptr())
        lib.call(name, *args, sums.data_ptr(), h)
        out.append((sums.cpu().numpy(), rho_u.cpu().numpy()))
    return out


def test_moments_noalloc_matches_moments():
    import ctypes as ct
    psc3d = load_package("psc3d")
    lib = load_package("_lib")
    for n, precision, seed in ((48, "fp32", 1), (20, "fp64", 2)):  # the second call is smaller
        print(f"[noalloc] {n}^3 {precision}", flush=True)
        s = _solver(psc3d, n, precision, seed)
        (s1, f1), (s2, f2) = _both(lib, s)
        assert np.array_equal(s1, s2), (s1, s2)
        assert np.array_equal(f1, f2)
        assert s1[1] > 0 and s1[12] == 0 and np.all(np.isfinite(s1))
    h = lib.load()
    rc = h.psc_moments_noalloc(ct.c_void_p(16), 1, 1, None, None, 8, 8, 8, 0, 2, 0, 0, 0, None, ct.c_void_p(64), None)
    assert rc == -1 and b"psc_moments_noalloc: null" in h.psc_last_error()
