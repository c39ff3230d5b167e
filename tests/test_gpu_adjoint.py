"""Permeability gradients on the GPU against the FP64 references (GPU).

Checks, all through the C ABI (adjoint module):
* psc_smooth_alpha against oracle.smooth_alpha (FP64 to 1e-13);
* the FP64 step with a PSC_ALPHA_FIELD solid fraction against the oracle's
  OracleSolver with the same field: bit-identical populations;
* the FP64 adjoint gradient against torch autograd of the FP64 reference
  model (tests/_torch_ref.py, itself pinned to the oracle and to finite
  differences on the CPU): 1e-9 relative to the largest component;
* the FP32 adjoint against the FP64 adjoint and against FP64 central finite
  differences (step 1e-3 cells): 1e-3 relative to the largest component
  (the stated adjoint tolerance);
* independence of the checkpoint interval (bit-identical gradients) and the
  torch.autograd.Function wrapper.
"""

import numpy as np
import pytest
import torch

from oracle import psc_oracle as O
from tests import _torch_ref as R
from tests.conftest import load_package

pytestmark = pytest.mark.gpu


def _pack(seed, n, P, rlo=2.0, rhi=4.0):
    rng = np.random.default_rng(seed)
    t = np.zeros((P, 10))
    t[:, 0:3] = rng.uniform(0, n, (P, 3))
    t[:, 3] = rng.uniform(rlo, rhi, P)
    return t


def test_smooth_alpha_matches_oracle():
    adj = load_package("adjoint")
    n = 24
    tab = _pack(0, n, 6, 2.0, 5.0)
    tab[0, 0:3] = (0.3, 23.6, 12.0)  # wraps over two faces
    tab[1, 0:4] = (10.0, 10.0, 10.0, 4.0)
    tab[2, 0:4] = (13.5, 10.4, 9.2, 3.0)  # overlaps particle 1
    print("[smooth_alpha] 24^3, 6 particles", flush=True)
    geo = adj.SmoothGeometry(tab, (n, n, n), "cuda:0")
    geo.build(torch.cuda.current_stream().cuda_stream)
    a64 = geo.alpha64.cpu().numpy().reshape(n, n, n)
    a32 = geo.alpha32.cpu().numpy().reshape(n, n, n)
    ref = O.smooth_alpha(tab, (n, n, n))
    assert np.max(np.abs(a64 - ref)) < 1e-13
    assert np.array_equal(a32, a64.astype(np.float32))
    assert np.array_equal(a64 == 1.0, ref == 1.0)
    # fixed-point log sums round terms below ~4e-16 to zero, never the reverse
    assert np.all(a64[ref == 0.0] == 0.0) and np.max(ref[a64 == 0.0]) < 1e-15


@pytest.mark.parametrize("force", [(0.0, 0.0, 1.0e-5), (0.0, 2.0e-5, 0.0)])
def test_field_step_fp64_bit_exact_vs_oracle(force):
    lib = load_package("_lib")
    n, T, tau = 14, 12, 0.8
    alpha = O.smooth_alpha(_pack(2, n, 3), (n, n, n))
    print(f"[field step fp64] {n}^3, force {force}", flush=True)
    N = n ** 3
    a = torch.from_numpy(alpha.ravel().copy()).to("cuda:0")
    st = [torch.empty(19 * N, dtype=torch.float64, device="cuda:0") for _ in range(2)]
    f = torch.empty(19 * N, dtype=torch.float64, device="cuda:0")
    s = torch.cuda.current_stream().cuda_stream
    lib.call("psc_init_state", st[0].data_ptr(), lib.PSC_F64_PLAIN, N, s)
    for t in range(T):
        lib.call("psc_step", st[t % 2].data_ptr(), st[(t + 1) % 2].data_ptr(), a.data_ptr(), None, None,
                 lib.PSC_F64_PLAIN, lib.PSC_PULL, n, n, n, lib.PSC_ALPHA_FIELD, tau, *force, 0, 0, 0, s)
    lib.call("psc_export_f64", st[T % 2].data_ptr(), lib.PSC_F64_PLAIN, lib.PSC_PULL, n, n, n, f.data_ptr(), s)
    ora = O.OracleSolver(n, [], tau=tau, body_force=force, static_geometry=True, solid_field=alpha)
    for _ in range(T):
        ora.collide_stream_psc()
    assert np.array_equal(f.cpu().numpy().reshape(19, n, n, n), ora.f)


def test_adjoint_fp64_matches_torch_autograd():
    adj = load_package("adjoint")
    n, T, tau, F = 12, 20, 0.8, (0.0, 0.0, 2.0e-5)
    tab = _pack(4, n, 3, 1.8, 3.2)
    print(f"[adjoint fp64 vs torch] {n}^3, T={T}", flush=True)
    prob = adj.PermeabilityAdjoint(n, tau=tau, body_force=F, steps=T, precision="fp64", checkpoint_every=5,
                                   device="cuda:0")
    k, g = prob.value_and_grad(tab)
    k_ref, g_ref = R.value_and_grad(tab, (n, n, n), tau, F, T)
    assert abs(k - k_ref) <= 1e-12 * abs(k_ref)
    scale = np.max(np.abs(g_ref))
    err = np.max(np.abs(g - g_ref)) / scale
    print(f"  k={k:.6e} max|dk|={scale:.3e} rel err {err:.2e}")
    assert err < 1e-9


def test_adjoint_fp32_vs_fp64_and_central_differences():
    adj = load_package("adjoint")
    n, T, tau, F = 32, 300, 0.8, (0.0, 0.0, 1.0e-6)
    tab = _pack(5, n, 8, 2.0, 4.0)
    print(f"[adjoint fp32 vs fp64 vs FD] {n}^3, T={T}", flush=True)
    p64 = adj.PermeabilityAdjoint(n, tau=tau, body_force=F, steps=T, precision="fp64", device="cuda:0")
    p32 = adj.PermeabilityAdjoint(n, tau=tau, body_force=F, steps=T, precision="fp32", device="cuda:0")
    k64, g64 = p64.value_and_grad(tab)
    k32, g32 = p32.value_and_grad(tab)
    scale = np.max(np.abs(g64))
    e32 = np.max(np.abs(g32 - g64)) / scale
    print(f"  k64={k64:.9e} k32={k32:.9e} fp32 grad rel err {e32:.2e}")
    assert abs(k32 - k64) <= 1e-5 * abs(k64)
    assert e32 < 1e-3
    # FP64 central differences on the 6 largest components
    h = 1e-3
    idx = np.argsort(-np.abs(g64).ravel())[:6]
    for flat in idx:
        p, j = divmod(int(flat), 4)
        tp, tm = tab.copy(), tab.copy()
        tp[p, j] += h
        tm[p, j] -= h
        fd = (p64.forward(tp) - p64.forward(tm)) / (2 * h)
        print(f"  p{p} j{j}: fd={fd:.6e} adj64={g64[p, j]:.6e} adj32={g32[p, j]:.6e}")
        assert abs(fd - g64[p, j]) <= 1e-4 * scale
        assert abs(fd - g32[p, j]) <= 1e-3 * scale


def test_checkpoint_interval_does_not_change_the_gradient():
    adj = load_package("adjoint")
    n, T = 16, 40
    tab = _pack(6, n, 4)
    print(f"[checkpoint intervals] {n}^3, T={T}", flush=True)
    out = []
    for K in (1, 7, 40, None):
        p = adj.PermeabilityAdjoint(n, tau=0.8, body_force=(1e-6, 0.0, 0.0), steps=T, precision="fp32",
                                    checkpoint_every=K, device="cuda:0")
        out.append(p.value_and_grad(tab))
    for k, g in out[1:]:
        assert k == out[0][0]
        assert np.array_equal(g, out[0][1])


def test_autograd_function_matches_backward():
    adj = load_package("adjoint")
    n, T = 16, 30
    tab = _pack(7, n, 3)
    print(f"[autograd.Function] {n}^3, T={T}", flush=True)
    prob = adj.PermeabilityAdjoint(n, tau=0.8, body_force=(0.0, 0.0, 1e-6), steps=T, precision="fp32",
                                   device="cuda:0")
    c = torch.tensor(tab[:, 0:3], dtype=torch.float64, requires_grad=True)
    r = torch.tensor(tab[:, 3], dtype=torch.float64, requires_grad=True)
    k = adj.permeability(c, r, prob)
    k.backward()
    g = prob.backward()
    assert np.array_equal(c.grad.numpy(), g[:, 0:3])
    assert np.array_equal(r.grad.numpy(), g[:, 3])
    # a second forward before backward is refused
    k2 = adj.permeability(c, r, prob)
    prob.forward(tab)
    with pytest.raises(RuntimeError):
        k2.backward()
