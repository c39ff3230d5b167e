"""Smooth solid fraction and the FP64 torch reference gradient, on the CPU.

Pins, without a GPU:
* ``oracle.smooth_alpha`` by closed forms: interior / exterior / transition
  cells of one sphere, the periodic image, the union rule of two spheres and
  the sphere volume;
* the torch reference forward (tests/_torch_ref.py) against the oracle's
  OracleSolver stepping the same solid-fraction field (the oracle is pinned to
  the reference's psc3d.py by the golden fixtures);
* the torch reference gradient against central finite differences in FP64.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import psc_oracle as O
from tests import _torch_ref as R


def _table(rows):
    t = np.zeros((len(rows), 10))
    t[:, 0:4] = np.asarray(rows, dtype=np.float64)
    return t


def test_smoothstep_shape():
    t = np.array([-0.3, 0.0, 0.25, 0.5, 0.75, 1.0, 1.7])
    S = O.smoothstep5(t)
    assert S[0] == 0.0 and S[1] == 0.0 and S[-2] == 1.0 and S[-1] == 1.0
    assert S[3] == 0.5
    assert np.allclose(O.smoothstep5(t) + O.smoothstep5(1.0 - t), 1.0, rtol=0, atol=1e-15)
    # flat ends (C1 and C2 joins): one-sided difference quotients vanish to O(h^2)
    h = 1e-4
    assert O.smoothstep5(np.array([h]))[0] / h < 1e-6
    assert (1.0 - O.smoothstep5(np.array([1.0 - h]))[0]) / h < 1e-6


def test_smooth_alpha_one_sphere_closed_form():
    n = 16
    c = np.array([7.2, 8.1, 6.7])
    r = 3.3
    a = O.smooth_alpha(_table([[*c, r]]), (n, n, n))
    g = np.indices((n, n, n)).astype(np.float64) + 0.5
    d = np.sqrt(((g - c.reshape(3, 1, 1, 1)) ** 2).sum(0))
    assert np.all(a[d <= r - 0.5] == 1.0)
    assert np.all(a[d >= r + 0.5] == 0.0)
    mid = (d > r - 0.5) & (d < r + 0.5)
    t = r + 0.5 - d[mid]
    assert np.allclose(a[mid], t ** 3 * (10 - 15 * t + 6 * t * t), rtol=1e-14, atol=1e-15)


def test_smooth_alpha_periodic_image_and_union():
    n = 12
    a1 = O.smooth_alpha(_table([[0.4, 5.0, 11.8, 2.6]]), (n, n, n))
    a2 = O.smooth_alpha(_table([[0.4 + n, 5.0 - n, 11.8 - n, 2.6]]), (n, n, n))
    assert np.allclose(a1, a2, rtol=0, atol=1e-12)  # images differ only by rounding of x - c
    assert a1[0, 5, 0] > 0 and a1[11, 5, 11] > 0  # the sphere wraps over the corner
    # union of two overlapping spheres: 1 - (1 - a)(1 - b)
    pa, pb = [5.0, 6.0, 6.0, 2.4], [7.1, 6.3, 5.8, 2.2]
    A = O.smooth_alpha(_table([pa]), (n, n, n))
    B = O.smooth_alpha(_table([pb]), (n, n, n))
    AB = O.smooth_alpha(_table([pa, pb]), (n, n, n))
    assert np.allclose(AB, 1.0 - (1.0 - A) * (1.0 - B), rtol=0, atol=1e-14)
    assert np.all((AB >= 0.0) & (AB <= 1.0))


@pytest.mark.parametrize("r", [4.0, 6.5])
def test_smooth_alpha_sphere_volume(r):
    """sum alpha = 4/3 pi r^3 + 8 pi r I + O(lattice), I = int_{-1/2}^{1/2}
    s (S(1/2 - s) - H(-s)) ds the interface-moment of the smoothstep."""
    n = 32
    c = np.array([15.37, 16.21, 15.83])
    a = O.smooth_alpha(_table([[*c, r]]), (n, n, n))
    s = np.linspace(-0.5, 0.5, 200001)
    I = np.trapezoid(s * (O.smoothstep5(0.5 - s) - (s < 0)), s)
    expect = 4.0 / 3.0 * math.pi * r ** 3 + 8.0 * math.pi * r * I
    assert abs(a.sum() - expect) / expect < 3e-3


def _pack(seed, n, P, rlo=2.0, rhi=3.5):
    rng = np.random.default_rng(seed)
    return _table([[*rng.uniform(0, n, 3), rng.uniform(rlo, rhi)] for _ in range(P)])


@pytest.mark.parametrize("force", [(0.0, 0.0, 2.0e-5), (1.5e-5, 0.0, 0.0)])
def test_torch_reference_matches_oracle(force):
    n, T, tau = 10, 15, 0.8
    tab = _pack(1, n, 2)
    alpha = O.smooth_alpha(tab, (n, n, n))
    ora = O.OracleSolver(n, [], tau=tau, body_force=force, static_geometry=True, solid_field=alpha)
    for _ in range(T):
        ora.collide_stream_psc()
    axis = int(np.argmax(np.abs(force)))
    u = (ora.ux, ora.uy, ora.uz)[axis]
    k_ora = O.permeability(ora.solid, u, n, tau, abs(force[axis]))
    k_ref = float(R.permeability(torch.tensor(alpha), tau, force, T))
    assert abs(k_ref - k_ora) <= 1e-12 * abs(k_ora)
    # the populations too: the torch state is post-collision (pull form)
    s = torch.tensor(O.W).view(19, 1, 1, 1).expand(19, n, n, n).clone()
    at = torch.tensor(alpha)
    for _ in range(T):
        s = R.collide(R.gather(s), at, tau, force)
    f = R.gather(s).numpy()
    assert np.max(np.abs(f - ora.f)) < 1e-14


def test_torch_reference_gradient_matches_central_differences():
    n, T, tau, F = 8, 8, 0.8, (0.0, 0.0, 2.0e-5)
    tab = _pack(3, n, 2, 1.6, 2.6)
    k0, g = R.value_and_grad(tab, (n, n, n), tau, F, T)
    assert k0 > 0
    h = 1e-4
    fd = np.zeros_like(g)
    for p in range(tab.shape[0]):
        for j in range(4):
            tp, tm = tab.copy(), tab.copy()
            tp[p, j] += h
            tm[p, j] -= h
            kp = float(R.k_of_particles(torch.tensor(tp[:, :3]), torch.tensor(tp[:, 3]), (n, n, n), tau, F, T))
            km = float(R.k_of_particles(torch.tensor(tm[:, :3]), torch.tensor(tm[:, 3]), (n, n, n), tau, F, T))
            fd[p, j] = (kp - km) / (2 * h)
    scale = np.max(np.abs(g))
    assert scale > 0
    assert np.max(np.abs(fd - g)) < 1e-6 * scale
