"""Plain bounce-back D3Q19 window solver of the drivers, on the GPU.

The reference's drag-proxy coupling and fines permeability probes use a
compact single-phase solver, ``run_lbm`` of scripts/run_3d_window_lbm.py:95-146,
on binary voxel masks from ``voxelize`` (run_3d_window_lbm.py:73-92) or
``voxelize_variable`` (run_local_fines_lbm_permeability.py:39-59), and a Darcy
fit over three forces, ``run_force_set`` (run_local_fines_lbm_permeability.py:
62-79).  This module keeps those four functions (same arguments and result
keys) and runs them through libpsc_hip:

* the voxel test runs on the GPU (``psc_voxelize_binary``) with the host
  computing the same cell boxes and cell-centre coordinates as the reference;
* the step is ``psc_bb_step``: FP64 (default) reproduces the reference bit
  for bit; ``precision="fp32"`` stores g = f - w in FP32;
* the result dictionary is formed from the velocity field the GPU computes
  (``psc_moments``), with the reference's NumPy reductions, so every entry
  equals the reference's in FP64.
"""

from __future__ import annotations

import numpy as np

from . import _lib

CS2 = 1.0 / 3.0


def _torch():
    import torch
    return torch


def _device(device):
    torch = _torch()
    _lib.load()
    if not torch.cuda.is_available():
        raise _lib.PscError("the bounce-back solver needs a ROCm GPU (gfx950)")
    return torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())


def _stream(dev):
    return _torch().cuda.current_stream(dev).cuda_stream


def _mark_spheres(spheres, boxes, axes, n, device):
    """uint8 device mask of the spheres (rows x, y, z, r) over their boxes."""
    torch = _torch()
    dev = _device(device)
    solid = torch.zeros(n ** 3, dtype=torch.uint8, device=dev)
    if len(spheres):
        sp = torch.from_numpy(np.ascontiguousarray(spheres, dtype=np.float64)).to(dev)
        bx = torch.from_numpy(np.ascontiguousarray(boxes, dtype=np.int32)).to(dev)
        ax = torch.from_numpy(np.ascontiguousarray(np.concatenate(axes), dtype=np.float64)).to(dev)
        _lib.call("psc_voxelize_binary", sp.data_ptr(), bx.data_ptr(), len(spheres), ax.data_ptr(), n, n, n,
                  solid.data_ptr(), _stream(dev))
    return solid


def _inclusive_boxes(lo_edges, hi_edges, dx, n):
    """floor(lo/dx) .. ceil(hi/dx) clamped to [0, n-1], per sphere and axis."""
    i0 = np.maximum(0, np.floor(lo_edges / dx).astype(np.int64))
    i1 = np.minimum(n - 1, np.ceil(hi_edges / dx).astype(np.int64))
    return i0, i1


def voxelize(points, center, size, radius, n, device=None):
    """(solid (n, n, n) bool, dx (3,)) of spheres of one ``radius`` centred
    at ``points`` in the window [center - size/2, center + size/2]; a cell is
    solid when its centre lies in a sphere."""
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    center = np.asarray(center, dtype=np.float64)
    size = np.asarray(size, dtype=np.float64)
    lo = center - 0.5 * size
    hi = center + 0.5 * size
    dx = size / n
    axes = [lo[a] + (np.arange(n) + 0.5) * dx[a] for a in range(3)]
    margin = radius + np.linalg.norm(dx)
    pts = points[np.all((points >= lo - margin) & (points <= hi + margin), axis=1)]
    boxes = np.zeros((len(pts), 6), dtype=np.int64)
    for a in range(3):
        c = pts[:, a]
        boxes[:, 2 * a], boxes[:, 2 * a + 1] = _inclusive_boxes(c - radius - lo[a], c + radius - lo[a], dx[a], n)
    spheres = np.column_stack([pts, np.full(len(pts), float(radius))]) if len(pts) else np.zeros((0, 4))
    solid = _mark_spheres(spheres, boxes, axes, n, device)
    return solid.cpu().numpy().reshape(n, n, n).astype(bool), dx


def voxelize_variable(particles, box_size, n, fine_radius_inflate_dx=0.0, device=None):
    """solid (n, n, n) bool of rows (id, type, x, y, z, r) in [0, box_size]^3;
    fines (type 2) grow by ``fine_radius_inflate_dx`` cells."""
    rows = np.asarray(particles, dtype=np.float64).reshape(-1, 6)
    dx = box_size / n
    axes = [(np.arange(n) + 0.5) * dx for _ in range(3)]
    r = rows[:, 5].copy()
    fines = rows[:, 1].astype(np.int64) == 2
    r[fines] = r[fines] + fine_radius_inflate_dx * dx
    margin = r + np.sqrt(3.0) * dx
    x = rows[:, 2:5]
    keep = ~(np.any(x < -margin[:, None], axis=1) | np.any(x > box_size + margin[:, None], axis=1))
    x, r = x[keep], r[keep]
    boxes = np.zeros((len(r), 6), dtype=np.int64)
    for a in range(3):
        boxes[:, 2 * a], boxes[:, 2 * a + 1] = _inclusive_boxes(x[:, a] - r, x[:, a] + r, dx, n)
    spheres = np.column_stack([x, r]) if len(r) else np.zeros((0, 4))
    solid = _mark_spheres(spheres, boxes, axes, n, device)
    return solid.cpu().numpy().reshape(n, n, n).astype(bool)


class BounceBackD3Q19:
    """The run_lbm solver as an object: a periodic lattice with a binary
    solid mask, BGK + Guo + bounce-back, stepped on the GPU."""

    def __init__(self, solid, tau: float = 0.72, force_z: float = 0.0, body_force=None, precision: str = "fp64",
                 device=None, block="auto", nontemporal: bool = False):
        torch = _torch()
        self.device = _device(device)
        solid = np.asarray(solid, dtype=bool)
        if solid.ndim != 3:
            raise ValueError("solid must be a 3-D mask")
        self.shape = solid.shape
        self.solid = solid
        self.N = int(np.prod(self.shape))
        self.tau = float(tau)
        if not self.tau > 0.5:
            raise ValueError("tau must exceed 1/2")
        self.force = tuple(float(v) for v in body_force) if body_force is not None else (0.0, 0.0, float(force_z))
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.storage = _lib.PSC_F64_PLAIN if precision == "fp64" else _lib.PSC_F32_DEV
        dt = torch.float64 if precision == "fp64" else torch.float32
        dev = self.device
        self._solid = torch.from_numpy(solid.astype(np.uint8).ravel()).to(dev)
        self._alpha = torch.from_numpy(solid.astype(np.float64 if precision == "fp64" else np.float32).ravel()).to(dev)
        self._a = torch.empty(19 * self.N, dtype=dt, device=dev)
        self._b = torch.empty(19 * self.N, dtype=dt, device=dev)
        _lib.call("psc_init_state", self._a.data_ptr(), self.storage, self.N, _stream(dev))
        self._rho_u = torch.empty(4 * self.N, dtype=torch.float64, device=dev)
        self._partials = torch.empty(int(_lib.load().psc_moments_blocks(self.N)) * _lib.PSC_NSUM,
                                     dtype=torch.float64, device=dev)
        self._sums = torch.empty(_lib.PSC_NSUM, dtype=torch.float64, device=dev)
        # "auto": the launch (workgroup shape and non-temporal access; for FP32 also the
        # occupancy cap) is timed on the first step; explicit blocks are used as given
        self._autotune = isinstance(block, str) and block == "auto"
        self.block = (0, 0) if isinstance(block, str) else tuple(int(v) for v in block)
        self.flags = _lib.PSC_STEP_NONTEMPORAL if nontemporal else 0
        self._graph = None
        self.steps = 0

    def _launch(self, a, b, s):
        nx, ny, nz = self.shape
        fx, fy, fz = self.force
        _lib.call("psc_bb_step", a.data_ptr(), b.data_ptr(), self._solid.data_ptr(), self.storage, nx, ny, nz,
                  self.tau, fx, fy, fz, self.block[0], self.block[1], self.flags, s)

    def _tune(self, budget_ms: float = 300.0):
        """Time the launch candidates on A -> B (the state is not modified),
        then the four fastest again, longer; keep the fastest.  Every
        candidate gives the same bits."""
        self._autotune = False
        torch = _torch()
        from . import autotune
        st = torch.cuda.current_stream(self.device)

        def timed(bz, by, flags, reps):
            self.block, self.flags = (bz, by), flags
            self._launch(self._a, self._b, st.cuda_stream)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                self._launch(self._a, self._b, st.cuda_stream)
            e1.record(st)
            e1.synchronize()
            return e0.elapsed_time(e1) / reps

        if self.storage == _lib.PSC_F64_PLAIN:  # the FP64 step's only launch flag is non-temporal access
            nz = self.shape[2]
            cands = [(bz, by, nt, 0) for bz, by in autotune.SHAPES + autotune.FP64_SHAPES
                     if not (bz > 64 and nz <= bz // 2) for nt in autotune.NT]
        else:
            cands = [(bz, by, nt, bpc) for bz, by, nt, bpc, flat in autotune.candidates(self.shape[2]) if not flat]
        probe = timed(64, 4, 0, 1)
        reps = max(2, min(20, int(budget_ms / max(probe, 1e-3) / len(cands))))
        res = sorted((timed(bz, by, autotune.flags_of(nt, bpc, 0), reps), bz, by, autotune.flags_of(nt, bpc, 0))
                     for bz, by, nt, bpc in cands)
        best = min((timed(bz, by, f, 4 * reps), bz, by, f) for _, bz, by, f in res[:4])
        self.block, self.flags = (best[1], best[2]), best[3]

    def step(self):
        if self._autotune:
            self._tune()
        self._launch(self._a, self._b, _stream(self.device))
        self._a, self._b = self._b, self._a
        self.steps += 1

    def run(self, steps: int, graph: bool = True):
        """``steps`` steps; pairs of steps replay a captured hipGraph."""
        torch = _torch()
        steps = int(steps)
        if self._autotune:
            self._tune()
        if not graph or steps < 4:
            for _ in range(steps):
                self.step()
            return
        if self._graph is None:
            self.step()  # loads the kernel before the capture
            steps -= 1
            a, b = self._a, self._b
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                s = torch.cuda.current_stream(self.device).cuda_stream
                self._launch(a, b, s)
                self._launch(b, a, s)
            self._graph = (g, a)
        g, a = self._graph
        if self._a is not a:
            self.step()
            steps -= 1
        for _ in range(steps // 2):
            g.replay()
        self.steps += 2 * (steps // 2)
        if steps % 2:
            self.step()

    def fields(self):
        """rho, ux, uy, uz of the current (post-stream) state as (n, n, n)
        FP64 arrays, the reference's end-of-run moments (velocities zero in
        solid cells)."""
        nx, ny, nz = self.shape
        fx, fy, fz = self.force
        _lib.call("psc_moments", self._a.data_ptr(), self.storage, _lib.PSC_PULL, self._alpha.data_ptr(), None, nx, ny,
                  nz, 0, _lib.PSC_ALPHA_FIELD, fx, fy, fz, self._rho_u.data_ptr(), self._partials.data_ptr(),
                  self._sums.data_ptr(), _stream(self.device))
        ru = self._rho_u.cpu().numpy().reshape(4, *self.shape)
        rho, ux, uy, uz = (ru[i].copy() for i in range(4))
        for u in (ux, uy, uz):
            u[self.solid] = 0.0
        return rho, ux, uy, uz

    def result(self, return_fields: bool = False, axis: int = 2) -> dict:
        """The run_lbm result dictionary of the current state (flow along
        ``axis``; 2 = z, the reference)."""
        _, ux, uy, uz = self.fields()
        u = (ux, uy, uz)[axis]
        fluid = ~self.solid
        F = self.force[axis]
        out = {
            "mean_uz_fluid": float(np.mean(u[fluid])),
            "superficial_uz": float(np.mean(u)),
            "max_uz": float(np.max(u[fluid])) if np.any(fluid) else 0.0,
            "porosity_voxel": float(np.mean(fluid)),
            "speed_slice": u[:, :, self.shape[2] // 2].copy(),
            "solid_slice": self.solid[:, :, self.shape[2] // 2].copy(),
        }
        out["permeability_lu"] = CS2 * (self.tau - 0.5) * out["superficial_uz"] / max(F, 1e-30)
        if return_fields:
            out.update({"ux": ux, "uy": uy, "uz": uz})
        return out


def run_lbm(solid, tau, force_z, steps, return_fields=False, precision="fp64", device=None) -> dict:
    """The reference's run_lbm on the GPU: ``steps`` steps from rest, then
    the result dictionary (mean_uz_fluid, superficial_uz, max_uz,
    permeability_lu, porosity_voxel, speed_slice, solid_slice [, ux, uy, uz])."""
    s = BounceBackD3Q19(solid, tau=tau, force_z=force_z, precision=precision, device=device)
    s.run(steps)
    return s.result(return_fields)


def run_force_set(solid, tau, force_z, steps, force_sweep, precision="fp64", device=None):
    """(result at the middle force, Darcy slope, intercept, R^2) over the
    forces 0.5 F, F, 2 F (or F alone)."""
    forces = [0.5 * force_z, force_z, 2.0 * force_z] if force_sweep else [force_z]
    results = [run_lbm(solid, tau, fz, steps, precision=precision, device=device) for fz in forces]
    fa = np.array(forces, dtype=float)
    sup = np.array([r["superficial_uz"] for r in results], dtype=float)
    if len(fa) > 1:
        slope, intercept = np.polyfit(fa, sup, 1)
        pred = slope * fa + intercept
        r2 = 1.0 - float(np.sum((sup - pred) ** 2)) / max(float(np.sum((sup - sup.mean()) ** 2)), 1.0e-30)
    else:
        slope, intercept, r2 = float(sup[0] / max(fa[0], 1.0e-30)), 0.0, float("nan")
    return results[len(results) // 2], float(slope), float(intercept), float(r2)
