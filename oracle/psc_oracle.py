"""FP64 NumPy oracle for the D3Q19 partially-saturated-cell IMB step.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and there only to check or
to time the CPU restatement.  The product package never imports it and has no
CPU fallback.

It restates ``src/lbm_dem/psc3d.py`` of the reference (wangjianfttt/
particuology_Hydraulic-penalty-) in FP64 with the same arithmetic, operation
order and quirks, so that on the reference's own inputs it reproduces the
reference bit for bit (pinned by ``tests/golden/*.npz``, written by
``tests/golden/make_golden.py`` from the reference itself).  Two documented
generalisations keep the reference behaviour as their default:

* ``body_force``: a 3-vector.  With ``(0, 0, Fz)`` every operation reduces to
  the reference's z-only force (psc3d.py:126,214,219,231,246), because the
  extra half-force terms are only formed for non-zero components.
* ``periodic_geometry``: when True, sphere bounding boxes wrap around the
  periodic lattice instead of being clamped (psc3d.py:73-78 clamps); lever arms
  then use the unwrapped cell position.  Default False = reference.

One extension has no reference counterpart: ``smooth_alpha`` (a smooth solid
fraction of sphere lists, the geometry of permeability gradients) and the
``solid_field`` option of ``OracleSolver`` that steps a given solid-fraction
field.  The step itself is unchanged; ``smooth_alpha`` is pinned by closed
forms (interior / exterior cells, the union rule, the sphere volume) in
``tests/test_adjoint_cpu.py``.

Parity status per function (how each is pinned) is listed in the test
modules ``tests/test_oracle_golden.py`` and ``tests/test_oracle_invariants.py``.
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

# --------------------------------------------------------------------------
# a1: D3Q19 lattice (psc3d.py:17-43).  Direction order: rest, +-x, +-y, +-z,
# then the 12 diagonals in (+,+),(-,-),(+,-),(-,+) pairs over xy, xz, yz.
# --------------------------------------------------------------------------
C = np.array(
    [[0, 0, 0],
     [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
     [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
     [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
     [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]],
    dtype=np.int64,
)
W = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12, dtype=np.float64)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17], dtype=np.int64)
CS2 = 1.0 / 3.0
Q = 19


@dataclass
class Sphere:
    """Sphere in lattice units (psc3d.py:46-57): centre, radius, U, omega."""

    x: float
    y: float
    z: float
    r: float
    vx: float = 0.0
    vy: float = 0.0
    vz: float = 0.0
    wx: float = 0.0
    wy: float = 0.0
    wz: float = 0.0


# --------------------------------------------------------------------------
# a3/a4: solid fraction by s^3 sub-sample counting
# --------------------------------------------------------------------------
def sphere_box(p, n: int, periodic: bool = False) -> tuple[int, int, int, int, int, int]:
    """Cell bounding box [lo, hi) per axis (psc3d.py:73-78).

    The reference clamps the box to [0, n); ``periodic=True`` keeps it
    unclamped (cells are wrapped when addressed).
    """
    lo = []
    hi = []
    for c in (p.x, p.y, p.z):
        a = int(np.floor(c - p.r - 1.0))
        b = int(np.ceil(c + p.r + 1.0))
        if not periodic:
            a = max(0, a)
            b = min(n, b)
        lo.append(a)
        hi.append(b)
    return lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]


def sphere_counts(p, n: int, subsamples: int, periodic: bool = False):
    """Integer count of sub-samples inside the sphere for every bbox cell.

    Follows psc3d.py:71-90: offsets (k+0.5)/s, sample coordinate
    ``(i + o) - centre``, distance ``(dx*dx + dy*dy) + dz*dz`` compared with
    ``r*r`` using ``<=``.  Returns (ix, iy, iz) index vectors of the box
    (unwrapped) and the count array of shape (bx, by, bz).
    """
    s = max(1, int(subsamples))
    offsets = (np.arange(s, dtype=np.float64) + 0.5) / s
    x0, x1, y0, y1, z0, z1 = sphere_box(p, n, periodic)
    ix = np.arange(x0, x1, dtype=np.int64)
    iy = np.arange(y0, y1, dtype=np.int64)
    iz = np.arange(z0, z1, dtype=np.int64)
    if ix.size == 0 or iy.size == 0 or iz.size == 0:
        return ix, iy, iz, np.zeros((ix.size, iy.size, iz.size), dtype=np.int64)
    dx = (ix[:, None] + offsets[None, :]) - p.x
    dy = (iy[:, None] + offsets[None, :]) - p.y
    dz = (iz[:, None] + offsets[None, :]) - p.z
    dx2 = dx * dx
    dy2 = dy * dy
    dz2 = dz * dz
    d2 = (dx2[:, None, None, :, None, None] + dy2[None, :, None, None, :, None]) \
        + dz2[None, None, :, None, None, :]
    inside = d2 <= p.r * p.r
    counts = inside.sum(axis=(3, 4, 5)).astype(np.int64)
    return ix, iy, iz, counts


def sparse_alpha(p, n: int, subsamples: int, periodic: bool = False):
    """(ids, alpha) of cells with alpha > 0, lexicographic (ix, iy, iz) order.

    Mirrors ``_sparse_particle_alpha_task`` (psc3d.py:60-94): alpha =
    count / float(s**3), flat id = (ix*n + iy)*n + iz.
    """
    s = max(1, int(subsamples))
    ix, iy, iz, counts = sphere_counts(p, n, s, periodic)
    alpha = counts.astype(np.float64) / float(s ** 3)
    sel = np.nonzero(alpha > 0.0)
    if periodic:
        gx, gy, gz = ix[sel[0]] % n, iy[sel[1]] % n, iz[sel[2]] % n
    else:
        gx, gy, gz = ix[sel[0]], iy[sel[1]], iz[sel[2]]
    ids = (gx * n + gy) * n + gz
    # unwrapped coordinates are kept for lever arms in periodic mode
    ux = ix[sel[0]]
    uy = iy[sel[1]]
    uz = iz[sel[2]]
    return ids.astype(np.int64), alpha[sel], (ux, uy, uz)


def dense_alpha(particles, n: int, subsamples: int, periodic: bool = False) -> np.ndarray:
    """alpha_p fields (P, n, n, n) of the dense solver (psc3d.py:151-173)."""
    out = np.zeros((len(particles), n, n, n), dtype=np.float64)
    s = max(1, int(subsamples))
    for ip, p in enumerate(particles):
        ix, iy, iz, counts = sphere_counts(p, n, s, periodic)
        if counts.size == 0:
            continue
        a = counts.astype(np.float64) / float(s ** 3)
        if periodic:
            out[ip][np.ix_(ix % n, iy % n, iz % n)] = a
        else:
            out[ip][np.ix_(ix, iy, iz)] = a
    return out


# --------------------------------------------------------------------------
# a6/a7: equilibrium and Guo forcing (psc3d.py:97-114), same operation order
# --------------------------------------------------------------------------
def smoothstep5(t):
    """Quintic smoothstep 6t^5 - 15t^4 + 10t^3 on [0, 1], 0 below, 1 above."""
    t = np.clip(t, 0.0, 1.0)
    return t * t * t * (10.0 + t * (-15.0 + 6.0 * t))


def smooth_alpha(table, shape):
    """Smooth solid fraction of spheres on a periodic lattice (no reference
    counterpart; the geometry of the permeability gradient):

        alpha(x) = 1 - prod_p (1 - S(r_p + 1/2 - |x - c_p|)),

    x the cell centre (i + 1/2) and |x - c| the minimum-image distance.  A cell
    fully inside a sphere (S = 1) gets alpha = 1.  ``table`` rows start with
    x, y, z, r (the particle-table layout)."""
    nx, ny, nz = shape
    X = [np.arange(m, dtype=np.float64) + 0.5 for m in (nx, ny, nz)]
    logs = np.zeros(shape)
    full = np.zeros(shape, dtype=bool)
    for row in np.asarray(table, dtype=np.float64):
        d2 = np.zeros(shape)
        for a, (m, xa) in enumerate(zip((nx, ny, nz), X)):
            dd = xa - row[a]
            dd = dd - m * np.round(dd / m)
            sh = [1, 1, 1]
            sh[a] = m
            d2 = d2 + (dd * dd).reshape(sh)
        ap = smoothstep5(row[3] + 0.5 - np.sqrt(d2))
        full |= ap >= 1.0
        part = (ap > 0.0) & (ap < 1.0)
        logs[part] += np.log1p(-ap[part])
    return np.where(full, 1.0, np.where(logs == 0.0, 0.0, -np.expm1(logs)))


def equilibrium(rho, ux, uy, uz):
    uu = ux * ux + uy * uy + uz * uz
    feq = np.empty((Q, *np.shape(rho)), dtype=np.float64)
    for i in range(Q):
        cx, cy, cz = C[i]
        cu = cx * ux + cy * uy + cz * uz
        feq[i] = W[i] * rho * (1.0 + cu / CS2 + 0.5 * cu * cu / (CS2 * CS2) - 0.5 * uu / CS2)
    return feq


def guo_force(fx, fy, fz, ux, uy, uz, tau):
    term = np.empty((Q, *np.shape(ux)), dtype=np.float64)
    pref = 1.0 - 0.5 / tau
    uf = ux * fx + uy * fy + uz * fz
    for i in range(Q):
        cx, cy, cz = C[i]
        cu = cx * ux + cy * uy + cz * uz
        cf = cx * fx + cy * fy + cz * fz
        term[i] = pref * W[i] * ((cf - uf) / CS2 + cu * cf / (CS2 * CS2))
    return term


def axis_moment(f, axis: int):
    """sum_i f_i c_{i,axis} as the reference forms it (psc3d.py:212-214)."""
    return np.sum(f * C[:, axis, None, None, None], axis=0)


def push_stream(f):
    """Periodic push streaming f_i(x + c_i) <- f_i(x) (psc3d.py:239-242)."""
    out = np.empty_like(f)
    for i in range(Q):
        out[i] = np.roll(f[i], shift=tuple(int(v) for v in C[i]), axis=(0, 1, 2))
    return out


def moments(f, body_force):
    """Post-stream rho, ux, uy, uz with the +F/2 shift (psc3d.py:243-246).

    The half force is added only on axes with a non-zero component, which for
    (0, 0, Fz) is exactly the reference's expression.
    """
    rho = np.sum(f, axis=0)
    u = []
    for a in range(3):
        m = axis_moment(f, a)
        if body_force[a] != 0.0:
            m = m + 0.5 * body_force[a]
        u.append(m / rho)
    return rho, u[0], u[1], u[2]


def b_weight(solid, tau):
    """B = alpha / (alpha + tau (1 - alpha) + 1e-14) (psc3d.py:228-229)."""
    alpha = np.clip(solid, 0.0, 1.0)
    return alpha / (alpha + tau * (1.0 - alpha) + 1.0e-14)


# --------------------------------------------------------------------------
# The solver (a8-a11)
# --------------------------------------------------------------------------
class OracleSolver:
    """FP64 restatement of ``PscD3Q19`` / ``SparsePscD3Q19`` (psc3d.py:117-407).

    ``layout='dense'`` follows the dense class (per-particle alpha fields,
    full-lattice load sums); ``layout='sparse'`` follows the sparse class
    (per-particle id lists, np.add.at alpha sum, gathered load sums).
    """

    def __init__(self, n, particles, tau=0.72, subsamples=4, body_force=(0.0, 0.0, 0.0),
                 static_geometry=False, layout="sparse", periodic_geometry=False, solid_field=None):
        if solid_field is not None and len(particles):
            raise ValueError("a solid_field replaces the particle list")
        self.solid_field = None if solid_field is None else np.asarray(solid_field, dtype=np.float64)
        self.n = int(n)
        self.particles = particles
        self.tau = float(tau)
        self.subsamples = int(subsamples)
        self.body_force = tuple(float(v) for v in body_force)
        self.static_geometry = bool(static_geometry)
        self.layout = layout
        self.periodic_geometry = bool(periodic_geometry)
        self._built = False
        shape = (self.n,) * 3
        self.rho = np.ones(shape)
        self.ux = np.zeros(shape)
        self.uy = np.zeros(shape)
        self.uz = np.zeros(shape)
        self.f = equilibrium(self.rho, self.ux, self.uy, self.uz)
        self.solid = np.zeros(shape)
        self.alpha_sum = np.zeros(shape)
        self.imb_mx = np.zeros(shape)
        self.imb_my = np.zeros(shape)
        self.imb_mz = np.zeros(shape)
        self.hydro_forces = np.zeros((len(particles), 3))
        self.hydro_torques = np.zeros((len(particles), 3))
        self.momentum_residual = 0.0
        self.node_ids = []
        self.alpha_vals = []
        self.node_pos = []
        self.alpha_p = None

    # -- geometry (psc3d.py:151-179 dense, 328-345 sparse) -----------------
    def rebuild_geometry(self):
        n = self.n
        if self.solid_field is not None:
            # a given solid fraction (no particles, no solid velocity, no loads)
            self.alpha_sum = self.solid_field.copy()
            self.alpha_p = np.zeros((0, n, n, n))
            self.node_ids, self.alpha_vals, self.node_pos = [], [], []
        elif self.layout == "dense":
            self.alpha_p = dense_alpha(self.particles, n, self.subsamples, self.periodic_geometry)
            self.alpha_sum = self.alpha_p.sum(axis=0)
        else:
            flat = np.zeros(n ** 3)
            self.node_ids, self.alpha_vals, self.node_pos = [], [], []
            for p in self.particles:
                ids, vals, pos = sparse_alpha(p, n, self.subsamples, self.periodic_geometry)
                if len(ids):
                    np.add.at(flat, ids, vals)
                self.node_ids.append(ids)
                self.alpha_vals.append(vals)
                self.node_pos.append(pos)
            self.alpha_sum = flat.reshape((n, n, n))
        self.solid = np.clip(self.alpha_sum, 0.0, 1.0)
        self._built = True

    # -- a5: solid velocity (psc3d.py:181-206 dense, 347-373 sparse) -----------
    def solid_velocity(self):
        n = self.n
        if self.layout == "dense":
            usx = np.zeros((n, n, n))
            usy = np.zeros_like(usx)
            usz = np.zeros_like(usx)
            xx, yy, zz = np.indices((n, n, n))
            for ip, p in enumerate(self.particles):
                a = self.alpha_p[ip]
                rx = (xx + 0.5) - p.x
                ry = (yy + 0.5) - p.y
                rz = (zz + 0.5) - p.z
                if self.periodic_geometry:
                    rx, ry, rz = self._unwrap_dense(rx, ry, rz)
                usx += a * (p.vx + p.wy * rz - p.wz * ry)
                usy += a * (p.vy + p.wz * rx - p.wx * rz)
                usz += a * (p.vz + p.wx * ry - p.wy * rx)
            asum = self.alpha_sum
        else:
            usx = np.zeros(n ** 3)
            usy = np.zeros_like(usx)
            usz = np.zeros_like(usx)
            for p, ids, a, pos in zip(self.particles, self.node_ids, self.alpha_vals, self.node_pos):
                if len(ids) == 0:
                    continue
                rx = pos[0] + 0.5 - p.x
                ry = pos[1] + 0.5 - p.y
                rz = pos[2] + 0.5 - p.z
                usx[ids] += a * (p.vx + p.wy * rz - p.wz * ry)
                usy[ids] += a * (p.vy + p.wz * rx - p.wx * rz)
                usz[ids] += a * (p.vz + p.wx * ry - p.wy * rx)
            asum = self.alpha_sum.ravel()
        mask = asum > 1.0e-14
        usx[mask] /= asum[mask]
        usy[mask] /= asum[mask]
        usz[mask] /= asum[mask]
        shape = (n, n, n)
        return usx.reshape(shape), usy.reshape(shape), usz.reshape(shape)

    def _unwrap_dense(self, rx, ry, rz):
        # periodic mode on the dense layout: nearest image of the cell centre
        n = self.n
        return (rx - n * np.round(rx / n), ry - n * np.round(ry / n), rz - n * np.round(rz / n))

    # -- a8: the step (psc3d.py:208-253) ---------------------------------------
    def collide_stream_psc(self):
        if (not self.static_geometry) or (not self._built):
            self.rebuild_geometry()
        fx, fy, fz = self.body_force
        f = self.f
        rho = np.sum(f, axis=0)
        jx = axis_moment(f, 0)
        jy = axis_moment(f, 1)
        jz = axis_moment(f, 2)
        ux = (jx + 0.5 * fx) / rho if fx != 0.0 else jx / rho
        uy = (jy + 0.5 * fy) / rho if fy != 0.0 else jy / rho
        uz = (jz + 0.5 * fz) / rho if fz != 0.0 else jz / rho
        usx, usy, usz = self.solid_velocity()
        # second half-force shift (quirk: psc3d.py:217-219)
        ux_e = ux + 0.5 * fx / rho if fx != 0.0 else ux
        uy_e = uy + 0.5 * fy / rho if fy != 0.0 else uy
        uz_e = uz + 0.5 * fz / rho if fz != 0.0 else uz
        feq_f = equilibrium(rho, ux_e, uy_e, uz_e)
        om_bgk = -(f - feq_f) / self.tau
        feq_s = equilibrium(rho, usx, usy, usz)
        om_s = f[OPP] - f + feq_s - feq_f[OPP]
        b = b_weight(self.solid, self.tau)
        coll = (1.0 - b) * om_bgk + b * om_s
        forcing = guo_force(fx, fy, fz, ux_e, uy_e, uz_e, self.tau)
        f = f + (coll + forcing)
        bos = b * om_s
        self.imb_mx = axis_moment(bos, 0)
        self.imb_my = axis_moment(bos, 1)
        self.imb_mz = axis_moment(bos, 2)
        self.accumulate_particle_loads()
        self.f = push_stream(f)
        self.rho, self.ux, self.uy, self.uz = moments(self.f, self.body_force)
        return self.diag()

    def diag(self):
        hf, ht = self.hydro_forces, self.hydro_torques
        return {
            "solid_volume_lu": float(np.sum(self.solid)),
            "mean_uz": float(np.mean(self.uz)),
            "momentum_residual": float(self.momentum_residual),
            "max_force": float(np.max(np.linalg.norm(hf, axis=1))) if len(hf) else 0.0,
            "max_torque": float(np.max(np.linalg.norm(ht, axis=1))) if len(ht) else 0.0,
        }

    # -- a10: node-summed loads (psc3d.py:255-281 dense, 375-407 sparse) -------
    def accumulate_particle_loads(self):
        n = self.n
        total = np.zeros(3)
        self.hydro_forces = np.zeros((len(self.particles), 3))
        self.hydro_torques = np.zeros((len(self.particles), 3))
        if self.layout == "dense":
            xx, yy, zz = np.indices((n, n, n))
            mask = self.alpha_sum > 1.0e-14
            for ip, p in enumerate(self.particles):
                w = np.zeros((n, n, n))
                w[mask] = self.alpha_p[ip][mask] / self.alpha_sum[mask]
                nfx = -w * self.imb_mx
                nfy = -w * self.imb_my
                nfz = -w * self.imb_mz
                rx = (xx + 0.5) - p.x
                ry = (yy + 0.5) - p.y
                rz = (zz + 0.5) - p.z
                if self.periodic_geometry:
                    rx, ry, rz = self._unwrap_dense(rx, ry, rz)
                self.hydro_forces[ip] = [float(nfx.sum()), float(nfy.sum()), float(nfz.sum())]
                self.hydro_torques[ip] = [float((ry * nfz - rz * nfy).sum()),
                                          float((rz * nfx - rx * nfz).sum()),
                                          float((rx * nfy - ry * nfx).sum())]
                total += self.hydro_forces[ip]
        else:
            asum = self.alpha_sum.ravel()
            ix_, iy_, iz_ = self.imb_mx.ravel(), self.imb_my.ravel(), self.imb_mz.ravel()
            for ip, (p, ids, a, pos) in enumerate(zip(self.particles, self.node_ids,
                                                      self.alpha_vals, self.node_pos)):
                if len(ids) == 0:
                    continue
                owner = a / np.maximum(asum[ids], 1.0e-14)
                nfx = -owner * ix_[ids]
                nfy = -owner * iy_[ids]
                nfz = -owner * iz_[ids]
                rx = pos[0] + 0.5 - p.x
                ry = pos[1] + 0.5 - p.y
                rz = pos[2] + 0.5 - p.z
                self.hydro_forces[ip] = [float(nfx.sum()), float(nfy.sum()), float(nfz.sum())]
                self.hydro_torques[ip] = [float((ry * nfz - rz * nfy).sum()),
                                          float((rz * nfx - rx * nfz).sum()),
                                          float((rx * nfy - ry * nfx).sum())]
                total += self.hydro_forces[ip]
        fluid = np.array([self.imb_mx.sum(), self.imb_my.sum(), self.imb_mz.sum()])
        self.momentum_residual = float(np.linalg.norm(total + fluid))


# --------------------------------------------------------------------------
# a9: readouts used by the validation drivers
# --------------------------------------------------------------------------
def fluid_statistics(solid, u, n):
    """(intrinsic, superficial, porosity) of one velocity component.

    run_precision_validation.py:142-147 (z component there).
    """
    w = 1.0 - np.clip(solid, 0.0, 1.0)
    vol = float(np.sum(w))
    intrinsic = float(np.sum(w * u) / max(vol, 1.0e-30))
    superficial = float(np.mean(w * u))
    return intrinsic, superficial, vol / float(n ** 3)


def permeability(solid, u, n, tau, force):
    """k = nu * sum((1 - solid) u) / n^3 / F, nu = (tau - 1/2)/3.

    run_psc_permeability_resolution.py:144-150.
    """
    w = 1.0 - solid
    superficial = float(np.sum(u * w) / n ** 3)
    nu = (tau - 0.5) / 3.0
    return float(nu * superficial / max(force, 1.0e-300))


def fluid_weighted_mean(solid, u):
    """run_d3q19_psc_moving_sphere_benchmark.py:57-59."""
    w = 1.0 - np.clip(solid, 0.0, 1.0)
    return float(np.sum(w * u) / max(float(np.sum(w)), 1.0e-30))


def hasimoto_factor(radius, box):
    """Periodic-array Stokes drag factor (run_d3q19_psc_moving_sphere_benchmark.py:50-54)."""
    lam = radius / box
    return float(1.0 / (1.0 - 2.837297 * lam + 4.19 * lam ** 3 - 27.4 * lam ** 6))


def sphere_volume(r):
    return 4.0 / 3.0 * math.pi * r ** 3
