"""Differentiable FP64 torch restatement of the smooth-geometry permeability
(test helper, CPU).

Forward model, the same as the oracle's OracleSolver with a solid_field
(oracle/psc_oracle.py, which follows src/lbm_dem/psc3d.py:208-246) written in
the pull form: s_{t+1}(x) = collide(f(x)), f_i(x) = s_t[i](x - c_i), s_0 = w;
k = nu sum (1 - alpha) (e.u) / (N |F|) with u = (j + F/2)/rho of the
post-stream populations.  Autograd through this model gives the reference
gradient the HIP adjoint is checked against.
"""

from __future__ import annotations

import numpy as np
import torch

from oracle.psc_oracle import C, OPP, W


def smooth_alpha(centres, radii, shape):
    """alpha = 1 - prod_p (1 - S(r_p + 1/2 - |x - c_p|)), minimum image."""
    nx, ny, nz = shape
    dt = centres.dtype
    axes = [torch.arange(m, dtype=dt) + 0.5 for m in (nx, ny, nz)]
    X, Y, Z = torch.meshgrid(*axes, indexing="ij")
    prod = torch.ones(shape, dtype=dt)
    for p in range(centres.shape[0]):
        d2 = 0.0
        for a, (G, m) in enumerate(zip((X, Y, Z), (nx, ny, nz))):
            dd = G - centres[p, a]
            dd = dd - m * torch.round(dd.detach() / m)
            d2 = d2 + dd * dd
        t = radii[p] + 0.5 - torch.sqrt(d2)
        tc = torch.clamp(t, 0.0, 1.0)
        S = tc * tc * tc * (10.0 + tc * (-15.0 + 6.0 * tc))
        prod = prod * (1.0 - S)
    return 1.0 - prod


def gather(s):
    """Pull streaming: f_i(x) = s_i(x - c_i)."""
    return torch.stack([torch.roll(s[i], shifts=tuple(int(v) for v in C[i]), dims=(0, 1, 2)) for i in range(19)])


def collide(f, alpha, tau, F):
    c = torch.tensor(C, dtype=f.dtype)
    w = torch.tensor(W, dtype=f.dtype).view(19, 1, 1, 1)
    Fv = torch.tensor(F, dtype=f.dtype)
    rho = f.sum(0)
    j = torch.einsum("ia,ixyz->axyz", c, f)
    u = (j + Fv.view(3, 1, 1, 1)) / rho
    cu = torch.einsum("ia,axyz->ixyz", c, u)
    usq = (u * u).sum(0)
    feq = w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    feqs = w * rho
    obgk = (feq - f) / tau
    opp = torch.tensor(OPP)
    os_ = f[opp] - f + feqs - feq[opp]
    B = alpha / ((alpha + tau * (1.0 - alpha)) + 1.0e-14)
    cF = (c * Fv).sum(1).view(19, 1, 1, 1)
    uF = (u * Fv.view(3, 1, 1, 1)).sum(0)
    guo = (1.0 - 0.5 / tau) * w * (3.0 * (cF - uF) + 9.0 * cu * cF)
    return f + (1.0 - B) * obgk + B * os_ + guo


def permeability(alpha, tau, F, steps):
    """k after ``steps`` pull steps from rest for a given alpha field."""
    shape = alpha.shape
    dt = alpha.dtype
    s = torch.tensor(W, dtype=dt).view(19, 1, 1, 1).expand(19, *shape).clone()
    for _ in range(steps):
        s = collide(gather(s), alpha, tau, F)
    f = gather(s)
    c = torch.tensor(C, dtype=dt)
    Fv = torch.tensor(F, dtype=dt)
    rho = f.sum(0)
    j = torch.einsum("ia,ixyz->axyz", c, f)
    u = (j + 0.5 * Fv.view(3, 1, 1, 1)) / rho
    fm = float(np.sqrt(np.sum(np.square(F))))
    e = Fv / fm
    eu = (u * e.view(3, 1, 1, 1)).sum(0)
    nu = (tau - 0.5) / 3.0
    N = float(np.prod(shape))
    return nu * ((1.0 - alpha) * eu).sum() / (N * fm)


def k_of_particles(centres, radii, shape, tau, F, steps):
    return permeability(smooth_alpha(centres, radii, shape), tau, F, steps)


def value_and_grad(table, shape, tau, F, steps):
    """k and (P, 4) dk/d(x, y, z, r) by autograd in FP64."""
    c = torch.tensor(np.asarray(table)[:, 0:3], dtype=torch.float64, requires_grad=True)
    r = torch.tensor(np.asarray(table)[:, 3], dtype=torch.float64, requires_grad=True)
    k = k_of_particles(c, r, shape, tau, F, steps)
    gc, gr = torch.autograd.grad(k, [c, r])
    g = np.zeros((c.shape[0], 4))
    g[:, 0:3] = gc.numpy()
    g[:, 3] = gr.numpy()
    return float(k.detach()), g
