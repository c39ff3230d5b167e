"""FP64 NumPy oracle of the plain bounce-back D3Q19 solver and its binary
voxelizers.

TEST INFRASTRUCTURE ONLY, like psc_oracle.py: imported by ``tests/`` (and
nothing in the product package).  It restates, in the same arithmetic and
operation order,

* ``voxelize`` of scripts/run_3d_window_lbm.py:73-92 (equal radii around a
  window centre) and ``voxelize_variable`` of
  scripts/run_local_fines_lbm_permeability.py:39-59 (per-particle radii, fines
  optionally inflated): a cell is solid when its centre lies in a sphere,
  tested over the sphere's clamped, inclusive cell box;
* ``run_lbm`` of scripts/run_3d_window_lbm.py:95-146: BGK with the Guo force
  along z (velocity with one half-force shift, zeroed in solid cells), push
  streaming, then the solid-node reflection loop, which assigns
  f[i] <- f[opp(i)] one direction after the other in place, so each opposite
  pair ends with both members equal to the even member's incoming value
  (a reflection of the even member, not a swap; kept as in the reference);
* ``run_force_set`` of scripts/run_local_fines_lbm_permeability.py:62-79
  (Darcy fit over three forces).

Pinned by tests/golden/bb_*.npz (written from the reference by
tests/golden/make_golden_bb.py) and by closed forms in
tests/test_bounce_back_cpu.py.
"""

from __future__ import annotations

import numpy as np

from .psc_oracle import C, CS2, OPP, Q, equilibrium, guo_force


def _box(lo_edge, hi_edge, dx, n):
    """Inclusive cell range [floor(lo/dx), ceil(hi/dx)] clamped to [0, n-1]."""
    return max(0, int(np.floor(lo_edge / dx))), min(n - 1, int(np.ceil(hi_edge / dx)))


def _mark(solid, axes, box, x, y, z, r):
    (i0, i1), (j0, j1), (k0, k1) = box
    ax = axes[0][i0:i1 + 1][:, None, None]
    ay = axes[1][j0:j1 + 1][None, :, None]
    az = axes[2][k0:k1 + 1][None, None, :]
    inside = (ax - x) ** 2 + (ay - y) ** 2 + (az - z) ** 2 <= r * r
    solid[i0:i1 + 1, j0:j1 + 1, k0:k1 + 1] |= inside


def voxelize(points, center, size, radius, n):
    """(solid bool (n,n,n), dx (3,)) of equal spheres in the window
    [center - size/2, center + size/2]."""
    lo = center - 0.5 * size
    hi = center + 0.5 * size
    dx = size / n
    axes = [lo[a] + (np.arange(n) + 0.5) * dx[a] for a in range(3)]
    solid = np.zeros((n, n, n), dtype=bool)
    margin = radius + np.linalg.norm(dx)
    keep = np.all((points >= lo - margin) & (points <= hi + margin), axis=1)
    for x, y, z in points[keep]:
        box = [_box(c - radius - lo[a], c + radius - lo[a], dx[a], n) for a, c in enumerate((x, y, z))]
        _mark(solid, axes, box, x, y, z, radius)
    return solid, dx


def voxelize_variable(particles, box_size, n, fine_radius_inflate_dx=0.0):
    """solid bool (n,n,n) of rows (id, type, x, y, z, r) in [0, box_size]^3;
    type-2 radii grow by fine_radius_inflate_dx cells."""
    dx = box_size / n
    axes = [(np.arange(n) + 0.5) * dx for _ in range(3)]
    solid = np.zeros((n, n, n), dtype=bool)
    for _, ptype, x, y, z, radius in particles:
        if int(ptype) == 2:
            radius = radius + fine_radius_inflate_dx * dx
        margin = radius + np.sqrt(3.0) * dx
        if min(x, y, z) < -margin or max(x, y, z) > box_size + margin:
            continue
        box = [_box(c - radius, c + radius, dx, n) for c in (x, y, z)]
        _mark(solid, axes, box, x, y, z, radius)
    return solid


def _moments(f, force_z):
    rho = np.sum(f, axis=0)
    ux = np.sum(f * C[:, 0, None, None, None], axis=0) / rho
    uy = np.sum(f * C[:, 1, None, None, None], axis=0) / rho
    uz = (np.sum(f * C[:, 2, None, None, None], axis=0) + 0.5 * force_z) / rho
    return rho, ux, uy, uz


def bb_step(f, solid, tau, force_z):
    """One step: collide (u zero in solids), push-stream, reflect at solids."""
    rho, ux, uy, uz = _moments(f, force_z)
    ux[solid] = 0.0
    uy[solid] = 0.0
    uz[solid] = 0.0
    feq = equilibrium(rho, ux, uy, uz)
    fpost = f - (f - feq) / tau + guo_force(0.0, 0.0, force_z, ux, uy, uz, tau)
    out = np.empty_like(f)
    for i in range(Q):
        out[i] = np.roll(fpost[i], shift=tuple(int(v) for v in C[i]), axis=(0, 1, 2))
    for i in range(Q):  # in place, in direction order
        out[i, solid] = out[OPP[i], solid].copy()
    return out


def run_lbm(solid, tau, force_z, steps, return_fields=False):
    """The reference's run_lbm result dictionary after ``steps`` steps from rest."""
    shape = solid.shape
    fluid = ~solid
    one = np.ones(shape)
    zero = np.zeros(shape)
    f = equilibrium(one, zero, zero, zero)
    nu = CS2 * (tau - 0.5)
    for _ in range(steps):
        f = bb_step(f, solid, tau, force_z)
    rho, ux, uy, uz = _moments(f, force_z)
    uz[solid] = 0.0
    out = {
        "mean_uz_fluid": float(np.mean(uz[fluid])),
        "superficial_uz": float(np.mean(uz)),
        "max_uz": float(np.max(uz[fluid])) if np.any(fluid) else 0.0,
        "porosity_voxel": float(np.mean(fluid)),
        "speed_slice": uz[:, :, shape[2] // 2].copy(),
        "solid_slice": solid[:, :, shape[2] // 2].copy(),
    }
    out["permeability_lu"] = nu * out["superficial_uz"] / max(force_z, 1e-30)
    if return_fields:
        ux[solid] = 0.0
        uy[solid] = 0.0
        out.update({"ux": ux.copy(), "uy": uy.copy(), "uz": uz.copy()})
    return out


def run_force_set(solid, tau, force_z, steps, force_sweep):
    """(result at the middle force, Darcy slope, intercept, R^2)."""
    forces = [0.5 * force_z, force_z, 2.0 * force_z] if force_sweep else [force_z]
    results = [run_lbm(solid, tau, fz, steps) for fz in forces]
    fa = np.array(forces, dtype=float)
    sup = np.array([r["superficial_uz"] for r in results], dtype=float)
    if len(fa) > 1:
        slope, intercept = np.polyfit(fa, sup, 1)
        pred = slope * fa + intercept
        r2 = 1.0 - float(np.sum((sup - pred) ** 2)) / max(float(np.sum((sup - sup.mean()) ** 2)), 1.0e-30)
    else:
        slope, intercept, r2 = float(sup[0] / max(fa[0], 1.0e-30)), 0.0, float("nan")
    return results[len(results) // 2], float(slope), float(intercept), float(r2)
