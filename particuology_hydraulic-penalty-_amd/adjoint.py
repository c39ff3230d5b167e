"""Permeability gradients with respect to particle centres and radii.

The reference computes the permeability of a sphere pack by running the PSC
step to (near) steady state and reading k = nu <(1 - solid) u>/F
(run_psc_permeability_resolution.py:137-150); its solid fraction is a
sub-sample count (psc3d.py:83-89), piecewise constant in the particle
positions, so it has no useful derivative.  This module makes k a smooth
function of the particles and differentiates it exactly:

* geometry: ``psc_smooth_alpha`` gives alpha(x) = 1 - prod_p (1 - S(r_p + 1/2
  - |x - c_p|)) with S the quintic smoothstep (a C2 cell-overlap fraction);
* forward: T steps of ``psc_step`` from rest with that field
  (PSC_ALPHA_FIELD, pull layout), FP32 deviation storage (or FP64 plain for
  reference runs); k from ``psc_moments`` in FP64;
* reverse: ``psc_adjoint_seed`` (dk/d state at T), then T reverse steps of
  ``psc_adjoint_step`` (the transposed collision Jacobian and reversed
  streaming, accumulating dk/dalpha per cell), then
  ``psc_smooth_alpha_grad`` (the chain rule to every centre and radius);
* memory: sqrt(T)-step checkpointing.  The forward keeps the state every K
  steps (K ~ sqrt(T), see ``checkpoint_interval``); the reverse sweep
  recomputes each segment's K states from its checkpoint into a segment
  buffer and runs the reverse steps over it, so the states held are about
  2 sqrt(T) and the forward work is 2T steps.

``PermeabilityFunction`` wraps forward and reverse as a torch.autograd.Function
of (centres, radii).  All device work runs through the C ABI of libpsc_hip;
there is no CPU fallback.
"""

from __future__ import annotations

import math

import numpy as np

from . import _lib
from .geometry import build_tables


def _torch():
    import torch
    return torch


class SmoothGeometry:
    """Device tables and the smooth solid fraction of one particle table on a
    periodic lattice."""

    def __init__(self, table, shape, device, chunk_cells: int = 2048):
        torch = _torch()
        self.shape = tuple(int(v) for v in shape)
        self.device = device
        table = np.ascontiguousarray(np.asarray(table, dtype=np.float64))
        if table.ndim != 2 or table.shape[1] != 10:
            raise ValueError(f"particle table must be (P, 10), got {table.shape}")
        if table.shape[0] and np.any(table[:, 4:10] != 0.0):
            raise ValueError("the permeability gradient assumes stationary particles")
        self.table = table
        self.P = table.shape[0]
        t = build_tables(table, self.shape, periodic=True, chunk_cells=chunk_cells)
        self.tabs = t
        self.chunk_cells = chunk_cells
        self.C = t.chunks.shape[0]
        self._parts = torch.from_numpy(t.particles.copy()).to(device)
        self._boxes = torch.from_numpy(t.boxes.copy()).to(device)
        self._chunks = torch.from_numpy(t.chunks.copy()).to(device)
        self._cfirst = torch.from_numpy(t.chunk_first.copy()).to(device)
        self._partials = torch.empty((max(1, self.C), 4), dtype=torch.float64, device=device)
        N = self.shape[0] * self.shape[1] * self.shape[2]
        self.N = N
        self.nfull = torch.empty(N, dtype=torch.int32, device=device)
        self.logsum = torch.empty(N, dtype=torch.int64, device=device)  # fixed point, 2^-50 units
        self.alpha64 = torch.empty(N, dtype=torch.float64, device=device)
        self.alpha32 = torch.empty(N, dtype=torch.float32, device=device)
        self.grad = torch.empty((max(1, self.P), 4), dtype=torch.float64, device=device)

    def build(self, stream):
        nx, ny, nz = self.shape
        _lib.call("psc_smooth_alpha", self._parts.data_ptr(), self._boxes.data_ptr(), self._chunks.data_ptr(), self.C,
                  self.chunk_cells, nx, ny, nz, self.nfull.data_ptr(), self.logsum.data_ptr(),
                  self.alpha64.data_ptr(), self.alpha32.data_ptr(), stream)

    def alpha(self, storage):
        return self.alpha32 if storage == _lib.PSC_F32_DEV else self.alpha64

    def backward(self, galpha, stream):
        """(P, 4) dL/d(x, y, z, r) from dL/dalpha per cell."""
        nx, ny, nz = self.shape
        _lib.call("psc_smooth_alpha_grad", self._parts.data_ptr(), self._boxes.data_ptr(), self._chunks.data_ptr(),
                  self._cfirst.data_ptr(), self.C, self.chunk_cells, self.P, nx, ny, nz, self.nfull.data_ptr(),
                  self.logsum.data_ptr(), galpha.data_ptr(), self._partials.data_ptr(), self.grad.data_ptr(), stream)
        return self.grad[:self.P]


def checkpoint_interval(steps: int) -> int:
    """Segment length K for sqrt(T) checkpointing: the smallest divisor of T
    in [sqrt(T), 1.5 sqrt(T)] (equal segments replay one captured graph),
    else ceil(sqrt(T))."""
    r = math.sqrt(steps)
    lo = int(math.ceil(r))
    for k in range(lo, int(1.5 * r) + 1):
        if steps % k == 0:
            return k
    return lo


class PermeabilityAdjoint:
    """k(particles) after ``steps`` steps from rest and its exact gradient.

    ``precision`` "fp32" runs the FP32 deviation step and an FP32 adjoint
    (FP64 reductions and dk/dalpha accumulation); "fp64" runs both in FP64
    (the reference gradient).  ``checkpoint_every`` = K (default
    ``checkpoint_interval(steps)``).  The body force sets the flow
    direction; k is the permeability along it.

    Launches: the forward runs each segment through psc_run_steps (hipGraph
    replay of the native step loop), with one device copy per checkpoint; the
    reverse sweep replays two captured graphs per full segment (the K-1
    recomputed steps, then the K reverse steps) after copying the segment's
    checkpoint into the segment buffer.  ``graph=False`` launches every
    kernel directly (same results, bit for bit)."""

    def __init__(self, shape, tau: float = 0.8, body_force=(0.0, 0.0, 1.0e-6), steps: int = 2000,
                 precision: str = "fp32", checkpoint_every: int | None = None, device=None, chunk_cells: int = 2048,
                 nontemporal: bool = True, graph: bool = True, tune: bool = True):
        torch = _torch()
        _lib.load()
        if not torch.cuda.is_available():
            raise _lib.PscError("PermeabilityAdjoint needs a ROCm GPU (gfx950)")
        self.shape = (int(shape),) * 3 if np.isscalar(shape) else tuple(int(v) for v in shape)
        self.tau = float(tau)
        if not self.tau > 0.5:
            raise ValueError("tau must exceed 1/2")
        self.body_force = tuple(float(v) for v in body_force)
        self.fmag = math.sqrt(sum(v * v for v in self.body_force))
        if self.fmag == 0.0:
            raise ValueError("the permeability needs a non-zero body force")
        self.steps = int(steps)
        if self.steps < 1:
            raise ValueError("steps must be >= 1")
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.precision = precision
        self.storage = _lib.PSC_F32_DEV if precision == "fp32" else _lib.PSC_F64_PLAIN
        self.dtype = torch.float32 if precision == "fp32" else torch.float64
        K = int(checkpoint_every) if checkpoint_every else checkpoint_interval(self.steps)
        self.K = max(1, min(K, self.steps))
        self.nseg = -(-self.steps // self.K)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.chunk_cells = chunk_cells
        self.flags = _lib.PSC_STEP_NONTEMPORAL if (nontemporal and precision == "fp32") else 0
        # (block_z, block_y, flags) of psc_step and psc_adjoint_step; (0, 0) = the library default
        # shape; tune=True times the candidates on the first forward()
        self.step_launch = (0, 0, self.flags)
        self.adjoint_launch = (0, 0, self.flags)
        self._tuned = not tune
        self.graph = bool(graph)
        nx, ny, nz = self.shape
        self.N = nx * ny * nz
        self._words = 19 * self.N * (1 if precision == "fp32" else 2)  # state size in 4-byte words
        dev = self.device
        # checkpoints: state at steps 0, K, 2K, ...; segment buffer: states 0..K-1 of one segment
        self._ckpt = torch.empty((self.nseg, 19 * self.N), dtype=self.dtype, device=dev)
        self._seg = torch.empty((self.K, 19 * self.N), dtype=self.dtype, device=dev)
        self._tmp = torch.empty((2, 19 * self.N), dtype=self.dtype, device=dev)
        self._lam = torch.empty((2, 19 * self.N), dtype=self.dtype, device=dev)
        self._galpha = torch.empty(self.N, dtype=torch.float64, device=dev)
        self._partials = torch.empty(int(_lib.load().psc_moments_blocks(self.N)) * _lib.PSC_NSUM,
                                     dtype=torch.float64, device=dev)
        self._sums = torch.empty(_lib.PSC_NSUM, dtype=torch.float64, device=dev)
        self.geometry = None
        self._final = None
        self._graphs = None
        self._version = 0
        self.k = None

    # ------------------------------------------------------------------ helpers
    def _stream(self):
        return _torch().cuda.current_stream(self.device).cuda_stream

    def _alpha_ptr(self):
        return self.geometry.alpha(self.storage).data_ptr()

    def _step(self, src, dst, s):
        nx, ny, nz = self.shape
        fx, fy, fz = self.body_force
        bz, by, fl = self.step_launch
        _lib.call("psc_step", src.data_ptr(), dst.data_ptr(), self._alpha_ptr(), None, None, self.storage,
                  _lib.PSC_PULL, nx, ny, nz, _lib.PSC_ALPHA_FIELD, self.tau, fx, fy, fz, bz, by, fl, s)

    def _reverse(self, state, lam_in, lam_out, s):
        nx, ny, nz = self.shape
        fx, fy, fz = self.body_force
        bz, by, fl = self.adjoint_launch
        _lib.call("psc_adjoint_step", state.data_ptr(), lam_in.data_ptr(), lam_out.data_ptr(), self._alpha_ptr(),
                  self._galpha.data_ptr(), self.storage, nx, ny, nz, self.tau, fx, fy, fz, bz, by, fl, s)

    def _copy(self, src, dst, s):
        _lib.call("psc_copy_f32", src.data_ptr(), dst.data_ptr(), self._words, s)

    def _plan(self):
        p = _lib.PscStepPlan()
        nx, ny, nz = self.shape
        p.state_a, p.state_b = self._tmp[0].data_ptr(), self._tmp[1].data_ptr()
        p.count_sum = self._alpha_ptr()
        p.storage, p.layout = self.storage, _lib.PSC_PULL
        p.nx, p.ny, p.nz, p.subsamples = nx, ny, nz, _lib.PSC_ALPHA_FIELD
        p.tau = self.tau
        p.fx, p.fy, p.fz = self.body_force
        p.block_z, p.block_y, p.flags = self.step_launch
        return p

    def _launch_candidates(self):
        """(block_z, block_y) shapes to try: row tiles of 64 and 128 and, when
        a row fits one workgroup, the row itself rounded to 32 cells."""
        nz = self.shape[2]
        shapes = [(64, 4), (64, 2), (128, 2), (256, 1)]
        r = -(-nz // 32) * 32
        if r <= 256:
            for by in (8, 4, 2, 1):
                if r * by <= 256 and (r * by) % 64 == 0:
                    shapes.append((r, by))
                    break
        out = []
        for bz, by in shapes:
            if bz > 64 and nz <= bz // 2:
                continue  # more than half of every row tile would idle
            if (bz, by) not in out:
                out.append((bz, by))
        return out

    def _tune(self, s):
        """Time every candidate launch of the forward step and of the reverse
        step (HIP events, 5 launches each on scratch buffers) and keep the
        fastest; results are bit-identical across launches."""
        torch = _torch()
        if self._tuned:
            return
        nx, ny, nz = self.shape
        fx, fy, fz = self.body_force
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream(self.device)

        def timed(fn):
            fn()
            ev0.record(stream)
            for _ in range(5):
                fn()
            ev1.record(stream)
            ev1.synchronize()
            return ev0.elapsed_time(ev1)

        caps = (0, 4, 6)
        nts = (0, 1) if self.precision == "fp32" else (0,)
        _lib.call("psc_init_state", self._tmp[0].data_ptr(), self.storage, self.N, s)
        _lib.call("psc_init_state", self._lam[0].data_ptr(), self.storage, self.N, s)
        best_f = best_a = None
        for bz, by in self._launch_candidates():
            for nt in nts:
                for cap in caps:
                    fl = nt | (cap << 8)
                    ms = timed(lambda: _lib.call("psc_step", self._tmp[0].data_ptr(), self._tmp[1].data_ptr(),
                                                 self._alpha_ptr(), None, None, self.storage, _lib.PSC_PULL, nx, ny,
                                                 nz, _lib.PSC_ALPHA_FIELD, self.tau, fx, fy, fz, bz, by,
                                                 fl | (1 << 16), s))
                    if best_f is None or ms < best_f[0]:
                        best_f = (ms, bz, by, fl | (1 << 16))
                    self.adjoint_launch = (bz, by, fl)
                    ms = timed(lambda: self._reverse(self._tmp[0], self._lam[0], self._lam[1], s))
                    if best_a is None or ms < best_a[0]:
                        best_a = (ms, bz, by, fl)
        self.step_launch = best_f[1:]
        self.adjoint_launch = best_a[1:]
        self._tuned = True

    def _segment_bounds(self, c):
        t0 = c * self.K
        return t0, min(t0 + self.K, self.steps)

    def _permeability(self, state, s):
        nx, ny, nz = self.shape
        fx, fy, fz = self.body_force
        _lib.call("psc_moments", state.data_ptr(), self.storage, _lib.PSC_PULL, self._alpha_ptr(), None, nx, ny, nz, 0,
                  _lib.PSC_ALPHA_FIELD, fx, fy, fz, None, self._partials.data_ptr(), self._sums.data_ptr(), s)
        sums = self._sums.cpu().numpy()
        e = np.array(self.body_force) / self.fmag
        superficial = float(e @ sums[5:8]) / self.N
        nu = (self.tau - 0.5) / 3.0
        return nu * superficial / self.fmag

    # ------------------------------------------------------------------ public
    def forward(self, table) -> float:
        """Build the smooth geometry of ``table`` ((P, 10) particle table or
        Particle3D list of stationary spheres), run the steps from rest keeping
        checkpoints, return k."""
        import ctypes as ct
        from .geometry import pack_particles
        s = self._stream()
        self.geometry = SmoothGeometry(pack_particles(table), self.shape, self.device, self.chunk_cells)
        self.geometry.build(s)
        self._graphs = None  # captured graphs hold the previous geometry's pointers
        self._tune(s)
        _lib.call("psc_init_state", self._tmp[0].data_ptr(), self.storage, self.N, s)
        plan = self._plan()
        par = ct.c_int32(0)
        for c in range(self.nseg):
            t0, t1 = self._segment_bounds(c)
            self._copy(self._tmp[par.value], self._ckpt[c], s)
            n = t1 - t0
            gs = n - (n & 1)
            _lib.call("psc_run_steps", ct.byref(plan), n, 1 if (self.graph and gs >= 2) else 0, max(gs, 2),
                      ct.byref(par), s)
        self._final = self._tmp[par.value]
        self.k = self._permeability(self._final, s)
        self._version += 1
        return self.k

    def _segment_reverse(self, n, s):
        """Recompute states 1..n-1 of a segment (state 0 in self._seg[0]) and
        run its n reverse steps; lam ends in self._lam[n % 2]."""
        for j in range(1, n):
            self._step(self._seg[j - 1], self._seg[j], s)
        for j in range(n - 1, -1, -1):
            k = n - 1 - j  # reverse steps done so far
            self._reverse(self._seg[j], self._lam[k % 2], self._lam[(k + 1) % 2], s)

    def _capture(self, s):
        torch = _torch()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._segment_reverse(self.K, torch.cuda.current_stream(self.device).cuda_stream)
        return g

    def backward(self, weight: float = 1.0):
        """weight * dk/d(x, y, z, r) per particle as a (P, 4) float64 array,
        for the geometry of the last forward()."""
        if self.geometry is None or self._final is None:
            raise RuntimeError("backward() needs a forward() first")
        torch = _torch()
        s = self._stream()
        nx, ny, nz = self.shape
        fx, fy, fz = self.body_force
        _lib.call("psc_adjoint_seed", self._final.data_ptr(), self._alpha_ptr(), self._lam[0].data_ptr(),
                  self._galpha.data_ptr(), self.storage, nx, ny, nz, self.tau, fx, fy, fz, float(weight), s)
        lam = 0  # buffer holding the adjoint of the current segment's end state
        first = True
        for c in range(self.nseg - 1, -1, -1):
            t0, t1 = self._segment_bounds(c)
            n = t1 - t0
            self._copy(self._ckpt[c], self._seg[0], s)
            if lam == 1:
                self._copy(self._lam[1], self._lam[0], s)
                lam = 0
            # the first segment runs directly (it also loads every kernel
            # before a capture); full segments after it replay one graph
            if self.graph and n == self.K and not first:
                if self._graphs is None:
                    self._graphs = self._capture(s)
                self._graphs.replay()
            else:
                self._segment_reverse(n, s)
            lam = n % 2
            first = False
        grad = self.geometry.backward(self._galpha, s)
        out = grad.cpu().numpy().copy()
        torch.cuda.current_stream(self.device).synchronize()
        return out

    def value_and_grad(self, table):
        k = self.forward(table)
        return k, self.backward()

    def memory_bytes(self) -> int:
        """Device bytes held for checkpoints, the segment buffer and the adjoint."""
        el = 4 if self.precision == "fp32" else 8
        return (self._ckpt.numel() + self._seg.numel() + self._tmp.numel() + self._lam.numel()) * el + 8 * self.N


def _table_of(centres, radii):
    c = centres.detach().to("cpu", dtype=_torch().float64).numpy()
    r = radii.detach().to("cpu", dtype=_torch().float64).numpy()
    t = np.zeros((c.shape[0], 10))
    t[:, 0:3] = c
    t[:, 3] = r
    return t


def _make_function():
    torch = _torch()

    class PermeabilityFunction(torch.autograd.Function):
        """k = PermeabilityFunction.apply(centres (P, 3), radii (P,), problem)
        with ``problem`` a PermeabilityAdjoint; backward gives dk/dcentres
        and dk/dradii through the adjoint sweep."""

        @staticmethod
        def forward(ctx, centres, radii, problem):
            k = problem.forward(_table_of(centres, radii))
            ctx.problem = problem
            ctx.version = problem._version
            ctx.meta = (centres.dtype, centres.device, radii.dtype, radii.device)
            return torch.tensor(k, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gk):
            problem = ctx.problem
            if problem._version != ctx.version:
                raise RuntimeError("the PermeabilityAdjoint ran another forward() before this backward()")
            g = problem.backward(float(gk))
            cd, cdev, rd, rdev = ctx.meta
            gc = torch.from_numpy(g[:, 0:3].copy()).to(device=cdev, dtype=cd)
            gr = torch.from_numpy(g[:, 3].copy()).to(device=rdev, dtype=rd)
            return gc, gr, None

    return PermeabilityFunction


_FN = None


def permeability(centres, radii, problem):
    """Differentiable permeability of the spheres (centres (P, 3), radii (P,))
    under ``problem`` (a PermeabilityAdjoint)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(centres, radii, problem)


def PermeabilityFunction():
    """The torch.autograd.Function class behind ``permeability``."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN
