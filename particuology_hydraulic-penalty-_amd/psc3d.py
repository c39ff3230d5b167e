"""Drop-in MI355X implementation of ``src/lbm_dem/psc3d.py``.

Same public names as the reference module (``C, W, OPP, CS2, Particle3D,
equilibrium, guo_force, PscD3Q19, SparsePscD3Q19``), same constructor
signatures, methods, attributes and ``diag`` keys (psc3d.py:17-407).  The lattice lives in HBM as torch-ROCm
tensors; every step runs the hand-written gfx950 kernels of libpsc_hip.so
(voxelize, fused collide+stream, particle loads, moments).  NumPy arrays are
produced only when an attribute is read.

Differences a caller can opt into (keywords after ``*``):

* ``precision="fp64"`` (default) runs the FP64 plain step in the reference's
  operation order: f, u, imb and the solid field equal the reference bit for
  bit on cells owned by one particle.  ``precision="fp32"`` stores
  g = f - w in FP32 (half the bytes, the fast path).
* ``body_force=(fx, fy, fz)`` generalises ``body_force_z``.
* ``periodic_geometry=True`` wraps sphere boxes around the periodic lattice
  instead of clamping them (psc3d.py:73-78 clamps).
* ``stream_layout="pull"|"push"`` selects the streaming form of the kernel;
  both give identical results.
* ``diag="lazy"|"eager"``: by default the returned diagnostics are computed
  when first read, which must happen before the next step.
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

from . import _lib
from .geometry import build_tables, pack_particles

PIPELINE_MIN_CELLS = 1 << 19  # unfused FP32 lattices from this size overlap the loads with the next step
FUSE_MAX_CELLS = 1 << 23      # FP32 lattices below this size fuse the loads into the step kernels
FUSE_PERIOD = 1               # fused loads (untuned): every load workgroup ahead of the cell workgroups
MAX_F32_CELLS = 1 << 30        # FP32 kernels address a direction plane with 32-bit byte offsets
GRAPH_STEPS = 256              # steps per captured graph of run() (graph launches leave ~8 us gaps between them)
MAX_CELLS = (1 << 31) - 1      # cell ids are 32-bit

# Lattice constants (psc3d.py:17-43): identical values and order.
C = np.array(
    [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
     [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
     [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]],
    dtype=int,
)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=float)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
CS2 = 1.0 / 3.0

DIAG_KEYS = ("solid_volume_lu", "mean_uz", "momentum_residual", "max_force", "max_torque")


@dataclass
class Particle3D:
    """Sphere in lattice units (psc3d.py:46-57).  Mutable: with dynamic
    geometry, edits take effect at the next step."""

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


def _torch():
    import torch
    return torch


def _on_device(shape, *arrays):
    """FP64 device copies of ``arrays`` broadcast to ``shape`` (flattened),
    on the current GPU; raises PscError without one (no host path)."""
    torch = _torch()
    _lib.load()
    if not torch.cuda.is_available():
        raise _lib.PscError("equilibrium/guo_force run on a ROCm GPU (gfx950); torch.cuda.is_available() is False")
    dev = torch.device("cuda", torch.cuda.current_device())
    out = []
    for a in arrays:
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape)).reshape(-1)
        out.append(torch.from_numpy(b).to(dev))
    return dev, out


def equilibrium(rho, ux, uy, uz) -> np.ndarray:
    """feq_i = w_i rho (1 + c_i.u/cs2 + (c_i.u)^2/(2 cs2^2) - u.u/(2 cs2)) as
    the (19, *rho.shape) float64 array of psc3d.py:97-103, computed on the GPU
    (psc_equilibrium, FP64, the reference's operation order)."""
    torch = _torch()
    shape = np.shape(rho)
    dev, (r, x, y, z) = _on_device(shape, rho, ux, uy, uz)
    n = r.numel()
    out = torch.empty(19 * n, dtype=torch.float64, device=dev)
    _lib.call("psc_equilibrium", r.data_ptr(), x.data_ptr(), y.data_ptr(), z.data_ptr(), n, out.data_ptr(),
              torch.cuda.current_stream(dev).cuda_stream)
    return out.cpu().numpy().reshape((19,) + tuple(shape))


def guo_force(fx, fy, fz, ux, uy, uz, tau) -> np.ndarray:
    """Guo forcing term (1 - 1/(2 tau)) w_i ((c_i.F - u.F)/cs2 + (c_i.u)(c_i.F)/cs2^2)
    as the (19, *ux.shape) float64 array of psc3d.py:106-114, computed on the
    GPU (psc_guo_force, FP64, the reference's operation order)."""
    torch = _torch()
    shape = np.shape(ux)
    dev, (x, y, z) = _on_device(shape, ux, uy, uz)
    n = x.numel()
    out = torch.empty(19 * n, dtype=torch.float64, device=dev)
    _lib.call("psc_guo_force", float(fx), float(fy), float(fz), x.data_ptr(), y.data_ptr(), z.data_ptr(), n,
              float(tau), out.data_ptr(), torch.cuda.current_stream(dev).cuda_stream)
    return out.cpu().numpy().reshape((19,) + tuple(shape))


class StaleDiagnostics(RuntimeError):
    """A lazily evaluated diag dict was read after the solver advanced."""


class LazyDiag(dict):
    """The dict ``collide_stream_psc`` returns (keys of psc3d.py:247-253).

    Values are computed on the GPU at the first read, which must come before
    the solver's next step (otherwise StaleDiagnostics is raised)."""

    def __init__(self, solver, step):
        super().__init__()
        self._solver = solver
        self._step = step
        self._ready = False

    def _fill(self):
        if self._ready:
            return
        if self._solver._step_count != self._step:
            raise StaleDiagnostics(
                "diagnostics of step %d read after step %d; read them before the next step or construct the "
                "solver with diag='eager'" % (self._step, self._solver._step_count))
        super().update(self._solver._diagnostics())
        self._ready = True
        self._solver = None

    def __getitem__(self, k):
        self._fill()
        return super().__getitem__(k)

    def get(self, k, default=None):
        self._fill()
        return super().get(k, default)

    def __iter__(self):
        self._fill()
        return super().__iter__()

    def __len__(self):
        self._fill()
        return super().__len__()

    def __contains__(self, k):
        self._fill()
        return super().__contains__(k)

    def keys(self):
        self._fill()
        return super().keys()

    def values(self):
        self._fill()
        return super().values()

    def items(self):
        self._fill()
        return super().items()

    def copy(self):
        self._fill()
        return dict(super().items())

    def __repr__(self):
        self._fill()
        return super().__repr__()

    def __eq__(self, other):
        self._fill()
        return dict(super().items()) == other

    __hash__ = None


class PscD3Q19:
    """D3Q19 PSC-IMB solver with node-summed particle loads (psc3d.py:117-281)."""

    _LAYOUT_NAME = "dense"

    def __init__(self, n, particles, tau: float = 0.72, subsamples: int = 4, body_force_z: float = 0.0,
                 static_geometry: bool = False, *, body_force=None, precision: str = "fp64",
                 stream_layout: str = "pull", periodic_geometry: bool = False, device=None,
                 chunk_cells: int = 2048, block="auto", nontemporal: bool = False, diag: str = "lazy",
                 owner_table=None):
        if isinstance(n, (tuple, list)):
            shape = tuple(int(v) for v in n)
            if len(shape) != 3:
                raise ValueError("n must be an int or a 3-tuple")
        else:
            shape = (int(n),) * 3
        if min(shape) < 1:
            raise ValueError(f"lattice {shape} has an empty axis")
        self._shape = shape
        self.n = shape[0]
        self._N = shape[0] * shape[1] * shape[2]
        if precision == "fp32" and self._N > MAX_F32_CELLS:
            raise ValueError(f"an FP32 lattice holds at most {MAX_F32_CELLS} cells (32-bit byte offsets in a plane); "
                             f"{shape} has {self._N}")
        if self._N > MAX_CELLS:
            raise ValueError(f"a lattice holds at most {MAX_CELLS} cells; {shape} has {self._N}")
        torch = _torch()
        self._lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.PscError("PscD3Q19 needs a ROCm GPU (gfx950); torch.cuda.is_available() is False")
        self._particles = particles
        self.tau = float(tau)
        if not self.tau > 0.5:
            raise ValueError("tau must exceed 1/2")
        self.subsamples = int(subsamples)
        self._s = max(1, self.subsamples)
        if body_force is not None:
            self._force = tuple(float(v) for v in body_force)
            if len(self._force) != 3:
                raise ValueError("body_force must have 3 components")
        else:
            self._force = (0.0, 0.0, float(body_force_z))
        self.static_geometry = bool(static_geometry)
        # run(): reduce the loads of step t inside the kernels of steps t+1 and t+2 (None =
        # where the kernel supports it -- FP32 pull storage, stationary particles -- on
        # lattices below FUSE_MAX_CELLS); fuse_period: the load workgroups lead those
        # grids, one in every fuse_period (None = as the launch tuner measured it, else
        # FUSE_PERIOD; odd: even periods put them on half the XCDs)
        self.fuse_loads = None
        self.fuse_period = None
        self._period_tuned = FUSE_PERIOD
        # fused run(): step pairs in place on one state buffer (AA pattern: half the
        # footprint, the same bits); None = as the launch tuner measured it, else forced
        self.inplace = None
        self._inplace_tuned = False
        self.loads_mode = None  # how the last native run() reduced the loads
        # run() without fused loads: overlap the loads of step t with step t+1 on a side
        # stream (None = by lattice size and precision)
        self.pipeline_loads = None
        # collide_stream_psc with static geometry: leave each step's loads pending and
        # reduce them inside the next step's launch, or at the first read of the loads
        # (None = wherever run() fuses the loads; True = wherever the kernel can)
        self.defer_loads = None
        self._pending = None      # (plan, parity, PSC_PENDING_* bits) of deferred steps
        self._held = None         # plan of a deferred step not launched yet (first of an in-place pair)
        self._step_plan = None    # (key, plan) of the deferred steps
        self.periodic_geometry = bool(periodic_geometry)
        # moving spheres re-voxelized every step, FP32 pull storage: record one owner
        # particle per covered cell and form its U + omega x r in the step kernel
        # instead of accumulating solid-velocity planes (None = wherever it applies)
        self.owner_table = owner_table
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.precision = precision
        self._storage = _lib.PSC_F32_DEV if precision == "fp32" else _lib.PSC_F64_PLAIN
        if stream_layout not in ("pull", "push"):
            raise ValueError("stream_layout must be 'pull' or 'push'")
        self.stream_layout = stream_layout
        self._layout = _lib.PSC_PULL if stream_layout == "pull" else _lib.PSC_PUSH
        if diag not in ("lazy", "eager"):
            raise ValueError("diag must be 'lazy' or 'eager'")
        self._diag_mode = diag
        self._chunk_cells = int(chunk_cells)
        if self._chunk_cells < 64 or self._chunk_cells & (self._chunk_cells - 1):
            raise ValueError("chunk_cells must be a power of two >= 64")
        # launch configuration of psc_step: "auto" = measured by the autotuner
        # at the first step (FP32: shape, non-temporal access, occupancy cap, cells per thread,
        # flat mapping; FP64: shape and non-temporal access)
        self._autotune = isinstance(block, str) and block == "auto"
        self._block = (0, 0) if self._autotune else (int(block[0]), int(block[1]))
        self._flags = _lib.PSC_STEP_NONTEMPORAL if nontemporal else 0
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self._fdt = torch.float32 if self._storage == _lib.PSC_F32_DEV else torch.float64
        N = self._N
        # FP32 direction planes start on 256-byte boundaries: planes of N cells with 4N
        # not a multiple of 64 bytes would leave every plane after the first misaligned,
        # and the step's stores would then cover partial 64-byte granules, which HBM
        # completes by read-modify-write (PSC_STEP_PLANES64: planes of N rounded up to 64)
        planes64 = self._storage == _lib.PSC_F32_DEV and N % 64 != 0
        self._state_flags = _lib.PSC_STEP_PLANES64 if planes64 else 0
        self._ps = -(-N // 64) * 64 if planes64 else N  # cells per direction plane of the state
        self._a = torch.empty(19 * self._ps, dtype=self._fdt, device=self.device)
        self._b = torch.empty(19 * self._ps, dtype=self._fdt, device=self.device)
        self._csum = torch.zeros(N, dtype=torch.int32, device=self.device)
        # imb of the last step; a second buffer is added for the pipelined loop
        self._imb_bufs = [torch.zeros(3 * N, dtype=self._fdt, device=self.device), None]
        self._imb = self._imb_bufs[0]
        self._partials_alt = None
        self._us = None
        self._own = None          # owner table (one particle id per covered cell)
        self._motion = None       # per-particle motion records of the owner table
        self._geom_mode = None    # how the current geometry was voxelized: count / planes / owners
        self._ent = None          # (ent_cell u32, ent_loc u32, ent_cnt u16) as int32/int32/int16 tensors
        self._old_ent = None
        self._us_clean = False    # count_sum/us_num are zero outside the current entries
        self._chunk_nnz = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._partials = torch.zeros(6, dtype=torch.float64, device=self.device)
        self._tables = None
        self._dev_tables = None
        self._loads = torch.zeros((0, 6), dtype=torch.float64, device=self.device)
        nblk = int(self._lib.psc_moments_blocks(N))
        self._mom_partials = torch.zeros(nblk * _lib.PSC_NSUM, dtype=torch.float64, device=self.device)
        self._sums = torch.zeros(_lib.PSC_NSUM, dtype=torch.float64, device=self.device)
        self._rho_u = None
        self._geometry_built = False
        self._step_count = 0
        self._cache = {}
        self._plan_epoch = 0      # bumped whenever a buffer that run()'s plan holds is replaced
        self._plan_cache = None
        _lib.call("psc_init_state", self._ptr(self._a), self._storage, self._ps, self._stream())
        self.geometry_workers = 1

    # ------------------------------------------------------------------ utils
    @property
    def _step_flags(self):
        """Flag word of the step calls: the launch flags and the state layout."""
        return self._flags | self._state_flags

    def _stream(self):
        return _torch().cuda.current_stream(self.device).cuda_stream

    @staticmethod
    def _ptr(t):
        return None if t is None else t.data_ptr()

    def _np(self, t):
        return t.detach().cpu().numpy()

    # ------------------------------------------------------------- particles
    @property
    def particles(self):
        return self._particles

    @particles.setter
    def particles(self, value):
        self._particles = value
        self._cache.pop("hydro_forces", None)
        self._cache.pop("hydro_torques", None)

    @property
    def body_force_z(self):
        return self._force[2]

    @property
    def body_force(self):
        return self._force

    # --------------------------------------------------------------- geometry
    def _upload_tables(self, arrays):
        """Device copies of host tables without stalling the host: the arrays
        are packed into one of two pinned staging buffers and copied
        asynchronously on the solver's stream into the matching device buffer
        (8-byte aligned views are returned).  A slot is reused two rebuilds
        later, after the copy out of it has completed (event wait), and the
        kernels that read the previous contents of its device buffer were
        queued before the new copy on the same stream."""
        torch = _torch()
        offs, total = [], 0
        for a in arrays:
            offs.append(total)
            total += (a.nbytes + 7) // 8 * 8
        total = max(total, 8)
        slot = self._stage_slot = 1 - getattr(self, "_stage_slot", 1)
        if not hasattr(self, "_stage"):
            self._stage = [None, None]
        entry = self._stage[slot]
        if entry is not None and entry[2] is not None:
            entry[2].synchronize()
        if entry is None or entry[0].numel() < total:
            cap = int(total * 1.25) + 64
            entry = [torch.empty(cap, dtype=torch.uint8, pin_memory=True),
                     torch.empty(cap, dtype=torch.uint8, device=self.device), None]
            self._stage[slot] = entry
        host, dev = entry[0], entry[1]
        hv = host.numpy()
        for a, o in zip(arrays, offs):
            hv[o:o + a.nbytes] = np.ascontiguousarray(a).view(np.uint8).ravel()
        stream = torch.cuda.current_stream(self.device)
        dev[:total].copy_(host[:total], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
        entry[2] = ev
        out = []
        for a, o in zip(arrays, offs):
            dt = {np.dtype(np.float64): torch.float64, np.dtype(np.int32): torch.int32}[a.dtype]
            out.append(dev[o:o + max(a.nbytes, 8)].view(dt)[: a.size].view(*a.shape) if a.size else
                       torch.empty(a.shape, dtype=dt, device=self.device))
        return out

    def _voxel_mode(self, moving: bool) -> str:
        """How a voxelization stores the solid velocity: "count" (static spheres:
        none), "owners" (owner table; FP32 pull storage with the geometry
        rebuilt every step, unless owner_table is False) or "planes"."""
        if not moving:
            return "count"
        fits = self._storage == _lib.PSC_F32_DEV and self._layout == _lib.PSC_PULL and not self.static_geometry
        if self.owner_table is None:
            return "owners" if fits else "planes"
        if self.owner_table and not fits:
            raise ValueError("owner_table needs FP32 pull storage and static_geometry=False")
        return "owners" if self.owner_table else "planes"

    def rebuild_geometry(self, table=None) -> None:
        """Voxelize every sphere (psc3d.py:151-179 / 328-345) on the GPU.
        ``table``: the packed (P, 10) particle table, if the caller has it."""
        # loads still pending are those of the geometry being replaced
        self._flush_loads()
        torch = _torch()
        table = pack_particles(self._particles) if table is None else table
        tabs = build_tables(table, self._shape, self.periodic_geometry, self._chunk_cells)
        dev = self.device
        parts, boxes, chunks, cfirst = self._upload_tables(
            [tabs.particles, tabs.boxes, tabs.chunks, tabs.chunk_first])
        C = int(tabs.chunks.shape[0])
        cap = max(1, C * tabs.chunk_cells)
        self._old_ent = self._ent
        old_nnz = self._chunk_nnz
        if self._ent is None or self._ent[0].numel() < cap:
            cap = int(cap * 1.25) + 1
            self._ent = (torch.empty(cap, dtype=torch.int32, device=dev),
                         torch.empty(cap, dtype=torch.int32, device=dev),
                         torch.empty(cap, dtype=torch.int16, device=dev))
        if self._chunk_nnz.numel() < max(1, C):
            self._chunk_nnz = torch.empty(max(1, int(C * 1.25)), dtype=torch.int32, device=dev)
            # the clear below still needs the previous counts
            self._chunk_nnz[: old_nnz.numel()].copy_(old_nnz)
        if self._partials.numel() < 6 * max(1, C):
            self._partials = torch.empty(6 * max(1, int(C * 1.25)), dtype=torch.float64, device=dev)
        P = table.shape[0]
        if self._loads.shape[0] != P:
            self._loads = torch.zeros((P, 6), dtype=torch.float64, device=dev)
        stream = self._stream()
        N = self._N
        mode = self._voxel_mode(tabs.moving)
        if mode != "count" and self._us is None:
            self._us = torch.empty(3 * N, dtype=self._fdt, device=dev)
            _lib.call("psc_clear", self._ptr(self._us), 3 * N * self._us.element_size(), stream)
        if mode == "owners" and self._own is None:
            self._own = torch.empty(N, dtype=torch.int32, device=dev)
        if mode == "owners" and (self._motion is None or self._motion.numel() < 12 * max(1, P)):
            self._motion = torch.empty(12 * max(1, P), dtype=torch.float64, device=dev)  # 96 B records
        old = self._tables
        if old is not None and self._old_ent is not None and self._us_clean:
            # zero only the cells the previous geometry touched (count_sum and us_num
            # are zero everywhere else)
            oc, _, on = self._old_ent
            nold = int(old.chunks.shape[0])
            if self._geom_mode == "owners":
                _lib.call("psc_clear_entries_owners", self._ptr(oc), self._ptr(on), self._ptr(self._chunk_nnz), nold,
                          old.chunk_cells, *self._shape, self._ptr(self._csum), self._ptr(self._us), stream)
            elif self._geom_mode == "planes":
                _lib.call("psc_clear_entries", self._ptr(oc), self._ptr(self._chunk_nnz), nold, old.chunk_cells,
                          *self._shape, self._ptr(self._csum), self._ptr(self._us),
                          1 if self._storage == _lib.PSC_F32_DEV else 2, stream)
            else:
                _lib.call("psc_clear_entries", self._ptr(oc), self._ptr(self._chunk_nnz), nold, old.chunk_cells,
                          *self._shape, self._ptr(self._csum), None, 0, stream)
        else:
            _lib.call("psc_clear", self._ptr(self._csum), 4 * N, stream)
            if self._us is not None:
                _lib.call("psc_clear", self._ptr(self._us), 3 * N * self._us.element_size(), stream)
        nx, ny, nz = self._shape
        ec, el, en = self._ent
        if mode == "owners":
            _lib.call("psc_voxelize_owners", self._ptr(parts), self._ptr(boxes), P, self._ptr(chunks), C,
                      tabs.chunk_cells, nx, ny, nz, self._s, self._ptr(ec), self._ptr(el), self._ptr(en),
                      self._ptr(self._chunk_nnz), self._ptr(self._csum), self._ptr(self._own), self._ptr(self._us),
                      self._ptr(self._motion), stream)
        else:
            us_ptr = self._ptr(self._us) if mode == "planes" else None
            usp = (1 if self._storage == _lib.PSC_F32_DEV else 2) if mode == "planes" else 0
            _lib.call("psc_voxelize", self._ptr(parts), self._ptr(boxes), self._ptr(chunks), C, tabs.chunk_cells,
                      nx, ny, nz, 0, self._s, self._ptr(ec), self._ptr(el), self._ptr(en),
                      self._ptr(self._chunk_nnz), self._ptr(self._csum), us_ptr, usp, 0, stream)
        self._geom_mode = mode
        self._tables = tabs
        # keep the device tables alive until the kernels that read them have run
        self._dev_tables = (parts, boxes, chunks, cfirst)
        self._moving = tabs.moving
        self._geometry_built = True
        self._us_clean = True
        self._old_ent = None
        # a copy: an array passed as the particles may be edited in place by the caller
        self._voxelized_table = np.array(table, dtype=np.float64, copy=True)
        self._plan_epoch += 1
        self._cache.clear()

    def _refresh_geometry(self) -> None:
        """The voxelization a step needs: built once with static geometry; with
        dynamic geometry whenever the particle table (centres, radii, U,
        omega) differs from the one last voxelized.  The reference re-voxelizes
        on every call (psc3d.py:209-210); for an unchanged table that gives
        the same solid fraction, solid velocity and entries as the last
        voxelization, which are still in place, so it is skipped (callers that
        hold spheres still with static_geometry=False, as the moving-sphere
        benchmark does, then pay for the step only)."""
        if self._geometry_built and self.static_geometry:
            return
        table = pack_particles(self._particles)
        last = getattr(self, "_voxelized_table", None)
        if self._geometry_built and last is not None and last.shape == table.shape and np.array_equal(last, table):
            return
        self.rebuild_geometry(table)

    # ------------------------------------------------------------------- step
    def collide_stream_psc(self):
        """One PSC collide + stream step (psc3d.py:208-253) on the GPU.

        With static geometry where run() fuses the loads, the step's loads are
        reduced inside the next step's launch or at their first read, and
        where run() steps in place in pairs the step may be held for the next
        call (defer_loads); every read through the solver sees every step
        taken, with the bits of a step followed by its load reduction."""
        self._advance()
        d = LazyDiag(self, self._step_count)
        if self._diag_mode == "eager":
            d._fill()
        return d

    def run(self, steps: int, graph: bool = True, graph_steps: int = GRAPH_STEPS) -> None:
        """``steps`` calls of collide_stream_psc without building diag dicts.

        With static geometry the loop runs natively (psc_run_steps): direct
        launches, or captured hipGraphs of at most ``graph_steps`` steps
        (a run that short is one graph).  With dynamic geometry every step
        re-voxelizes from the Python objects when they changed."""
        steps = int(steps)
        if steps <= 0:
            return
        if (not self.static_geometry) or (not self._geometry_built):
            for _ in range(steps):
                self._advance()
            return
        import ctypes as ct
        self._sync_steps()
        plan, fused, pipelined = self._run_plan()
        # the run reduces its last step's loads, which supersede any left pending
        self._pending = None
        self.loads_mode = "fused" if fused else ("pipelined" if pipelined else "in line")
        parity = ct.c_int32(0)
        _lib.call("psc_run_steps", ct.byref(plan), steps, 1 if graph else 0, int(graph_steps), ct.byref(parity),
                  self._stream())
        if parity.value:
            self._a, self._b = self._b, self._a
        # the last step read parity 1 - p and wrote that imb buffer
        self._imb = self._imb_bufs[1 - parity.value] if (pipelined or fused) else self._imb_bufs[0]
        self._step_count += steps
        self._cache.clear()

    def _run_plan(self):
        """(plan, fused, pipelined) of the native loop, tuned and with its
        buffers allocated.  The ctypes plan is kept while nothing it holds
        changed (buffers, launch, load settings), so a run starts with no
        host work beyond the call."""
        if self._autotune:
            self._tune()
        fused, pipelined = self._load_modes()
        if pipelined or fused:
            self._ensure_pipeline_buffers()
        key = (self._plan_epoch, id(self._a), fused, pipelined, self._block, self._flags, self.inplace,
               self._inplace_tuned, self.effective_fuse_period(), self.tau, self._force)
        if self._plan_cache is None or self._plan_cache[0] != key:
            self._plan_cache = (key, self._plan(pipelined, fused))
        return self._plan_cache[1], fused, pipelined

    def prepare_run(self, steps: int, graph_steps: int = GRAPH_STEPS) -> None:
        """Build the geometry, tune the launch and capture the hipGraphs that
        ``run(steps, graph_steps=graph_steps)`` replays, without stepping
        (psc_prepare_steps), so that run() itself starts at once.  The graphs
        hold the state buffers the next run starts from: steps taken in
        between (an odd number of them moves the state to the other buffer)
        can leave the run needing graphs of its own, so call this right
        before the run it prepares."""
        import ctypes as ct
        if not self._geometry_built:
            self.rebuild_geometry()
        if not self.static_geometry or int(steps) <= 0:
            return
        self._sync_steps()  # the run starts from the buffer a held step moves the state to
        plan, _, _ = self._run_plan()
        parity = ct.c_int32(0)
        _lib.call("psc_prepare_steps", ct.byref(plan), int(steps), int(graph_steps), ct.byref(parity))

    def effective_fuse_period(self) -> int:
        """Placement period of the load workgroups in fused steps."""
        return max(1, int(self._period_tuned if self.fuse_period is None else self.fuse_period))

    def _load_modes(self):
        """(fused, pipelined): how run() reduces the per-particle loads.

        FP32 pull lattices with stationary particles below FUSE_MAX_CELLS reduce the
        loads inside the next steps' kernels (measured, us/step, fused vs side stream:
        64^3 11.7 vs 23.3, 96^3 25.5 vs 35.9, 128^3 63.7 vs 70.7, 256^3 517 vs 510).
        Otherwise the loads of step t overlap step t+1 on a side stream when the step
        is long enough to hide them (FP64, or FP32 from ~80^3 up); on small lattices
        the fork/join costs more than it hides and the loads run in line."""
        has_loads = self._tables is not None and self._tables.particles.shape[0] > 0
        fusable = self._fusable()
        fused = fusable and (self._N < FUSE_MAX_CELLS if self.fuse_loads is None else bool(self.fuse_loads))
        auto = self._storage == _lib.PSC_F64_PLAIN or self._N >= PIPELINE_MIN_CELLS
        pipelined = (has_loads and not fused
                     and (auto if self.pipeline_loads is None else bool(self.pipeline_loads)))
        return fused, pipelined

    def _fusable(self) -> bool:
        """The step kernel can carry the loads: particles, FP32 pull storage,
        no solid velocity."""
        return (self._tables is not None and self._tables.particles.shape[0] > 0
                and self._storage == _lib.PSC_F32_DEV and self._layout == _lib.PSC_PULL and not self._moving)

    def _defer(self) -> bool:
        """Whether collide_stream_psc leaves its loads pending (defer_loads)."""
        if not (self.static_geometry and self._geometry_built):
            return False
        if self.defer_loads is None:
            return self._load_modes()[0]
        return bool(self.defer_loads) and self._fusable()

    def _ensure_pipeline_buffers(self):
        torch = _torch()
        if self._imb_bufs[1] is None:
            self._imb_bufs[1] = torch.zeros(3 * self._N, dtype=self._fdt, device=self.device)
            self._plan_epoch += 1
        need = self._partials.numel()
        if self._partials_alt is None or self._partials_alt.numel() < need:
            self._partials_alt = torch.empty(need, dtype=torch.float64, device=self.device)
            self._plan_epoch += 1

    def _plan(self, pipelined: bool = False, fused: bool = False, inplace=None):
        """PscStepPlan of the current buffers (state in self._a); ``inplace``
        None = the in-place pairs as tuned or set."""
        p = _lib.PscStepPlan()
        nx, ny, nz = self._shape
        p.state_a, p.state_b = self._ptr(self._a), self._ptr(self._b)
        p.count_sum = self._ptr(self._csum)
        p.us_num = self._ptr(self._us) if self._moving else None
        p.imb = self._ptr(self._imb_bufs[0])
        p.imb_alt = self._ptr(self._imb_bufs[1]) if (pipelined or fused) else None
        p.storage, p.layout = self._storage, self._layout
        p.nx, p.ny, p.nz, p.subsamples = nx, ny, nz, self._s
        p.tau = self.tau
        p.fx, p.fy, p.fz = self._force
        p.block_z, p.block_y = self._block
        if inplace is None:
            inplace = self._inplace_tuned if self.inplace is None else bool(self.inplace)
        p.flags = self._step_flags | (_lib.PSC_RUN_INPLACE if fused and inplace else 0)
        tabs = self._tables
        if tabs is not None and tabs.particles.shape[0] > 0:
            parts, boxes, chunks, cfirst = self._dev_tables
            p.particles, p.boxes, p.chunks, p.chunk_first = (self._ptr(parts), self._ptr(boxes), self._ptr(chunks),
                                                             self._ptr(cfirst))
            p.nparticles, p.nchunks, p.chunk_cells = (int(tabs.particles.shape[0]), int(tabs.chunks.shape[0]),
                                                      tabs.chunk_cells)
            ec, el, en = self._ent
            p.ent_cell, p.ent_loc, p.ent_cnt = self._ptr(ec), self._ptr(el), self._ptr(en)
            p.chunk_nnz, p.partials = self._ptr(self._chunk_nnz), self._ptr(self._partials)
            p.partials_alt = self._ptr(self._partials_alt) if (pipelined or fused) else None
            p.loads = self._ptr(self._loads)
            p.fuse_loads = self.effective_fuse_period() if fused else 0
        return p

    def _tune(self):
        """Pick the psc_step launch configuration once (autotune.py).  When run()
        will fuse the loads into the step kernels (static geometry built), the
        candidates are timed in that fused loop, whose best launch differs from
        the plain step's."""
        if self._autotune:
            from . import autotune
            self._flush_loads()  # the unfused timing steps write the solver's own imb
            fused = self.static_geometry and self._geometry_built and self._load_modes()[0]
            if fused:
                self._ensure_pipeline_buffers()
            bz, by, flags = autotune.tune(self, fused=fused)
            self._block = (bz, by)
            self._flags = flags & ~_lib.PSC_RUN_INPLACE
            self._inplace_tuned = bool(flags & _lib.PSC_RUN_INPLACE)
            if fused:
                self._period_tuned = autotune.tune_period(self, bz, by, flags)
        self._autotune = False

    def _advance(self):
        self._refresh_geometry()
        if self._autotune:
            self._tune()
        if self._defer():
            self._step_deferred()
            return
        self._sync_steps()
        nx, ny, nz = self._shape
        stream = self._stream()
        fx, fy, fz = self._force
        us_ptr = self._ptr(self._us) if self._moving else None
        self._imb = self._imb_bufs[0]
        if self._geom_mode == "owners":
            _lib.call("psc_step_owners", self._ptr(self._a), self._ptr(self._b), self._ptr(self._csum),
                      self._ptr(self._own), us_ptr, self._ptr(self._motion), int(self._tables.particles.shape[0]),
                      self._ptr(self._imb), nx, ny, nz, self._s, self.tau, fx, fy, fz, self._block[0], self._block[1],
                      self._step_flags, stream)
        else:
            _lib.call("psc_step", self._ptr(self._a), self._ptr(self._b), self._ptr(self._csum), us_ptr,
                      self._ptr(self._imb), self._storage, self._layout, nx, ny, nz, self._s, self.tau, fx, fy, fz,
                      self._block[0], self._block[1], self._step_flags, stream)
        self._a, self._b = self._b, self._a
        self._step_count += 1
        self._cache.clear()
        self.accumulate_particle_loads()

    def _step_deferred(self):
        """A step whose loads stay pending (psc_run_steps_deferred): its launch
        also reduces the loads earlier steps left pending, so a step is one
        launch, and the last step's loads are reduced when they are read
        (_flush_loads).  Where run() steps in place in pairs, a step is held
        until the next call and the two launch as a pair on one state buffer;
        anything that reads the state, imb or loads first launches a held step
        alone (_sync_steps).  Same bits as a step followed by
        accumulate_particle_loads."""
        plan = self._deferred_plan()
        if self._held is not None and self._held is not plan:
            self._sync_steps()
        if (plan.flags & _lib.PSC_RUN_INPLACE) and self._held is None:
            self._held = plan
        else:
            pair = self._held is not None
            self._held = None
            self._launch_deferred(plan, 2 if pair else 1)
        self._step_count += 1
        self._cache.clear()

    def _deferred_plan(self):
        """The fused plan of the deferred steps, kept while nothing it holds
        changed."""
        self._ensure_pipeline_buffers()
        key = (self._plan_epoch, self._block, self._flags, self.inplace, self._inplace_tuned,
               self.effective_fuse_period(), self.tau, self._force)
        cached = self._step_plan
        if (cached is None or cached[0] != key
                or {cached[1].state_a, cached[1].state_b} != {self._ptr(self._a), self._ptr(self._b)}):
            cached = self._step_plan = (key, self._plan(fused=True))
        return cached[1]

    def _launch_deferred(self, plan, nsteps: int) -> None:
        import ctypes as ct
        parity = ct.c_int32(0 if self._ptr(self._a) == plan.state_a else 1)
        pend = self._pending
        if pend is not None and (pend[0] is not plan or pend[1] != parity.value):
            self._reduce_pending()
            pend = None
        bits = ct.c_int32(0 if pend is None else pend[2])
        _lib.call("psc_run_steps_deferred", ct.byref(plan), nsteps, ct.byref(parity), ct.byref(bits), self._stream())
        if nsteps & 1:  # an in-place pair leaves the state in its buffer
            self._a, self._b = self._b, self._a
        # the last step read parity 1 - p and wrote that imb buffer
        self._imb = self._imb_bufs[1 - parity.value]
        self._pending = (plan, parity.value, bits.value)

    def _sync_steps(self) -> None:
        """Launch a step held back for an in-place pair on its own (an
        ordinary step to the other buffer), so that the state, imb and
        pending loads are those of every step taken."""
        held = self._held
        if held is not None:
            self._held = None
            self._launch_deferred(held, 1)

    def _reduce_pending(self) -> None:
        pend = self._pending
        if pend is None:
            return
        if pend[2]:
            import ctypes as ct
            bits = ct.c_int32(pend[2])
            _lib.call("psc_flush_loads", ct.byref(pend[0]), pend[1], ct.byref(bits), self._stream())
        self._pending = None

    def _flush_loads(self) -> None:
        """Launch a held step and reduce the loads deferred steps left
        pending (psc_flush_loads), so that the loads buffer holds the last
        step's."""
        self._sync_steps()
        self._reduce_pending()

    def accumulate_particle_loads(self) -> None:
        """Node-summed force and torque per particle (psc3d.py:255-281 / 375-407)."""
        self._sync_steps()
        # the last step's loads computed afresh: whatever deferred steps left pending is superseded
        self._pending = None
        tabs = self._tables
        if tabs is None:
            return
        parts, boxes, chunks, cfirst = self._dev_tables
        nx, ny, nz = self._shape
        imbp = 1 if self._storage == _lib.PSC_F32_DEV else 2
        ec, el, en = self._ent
        _lib.call("psc_particle_loads", self._ptr(parts), self._ptr(boxes), self._ptr(chunks), self._ptr(cfirst),
                  int(tabs.particles.shape[0]), int(tabs.chunks.shape[0]), tabs.chunk_cells, self._ptr(ec),
                  self._ptr(el), self._ptr(en), self._ptr(self._chunk_nnz), self._ptr(self._csum),
                  self._ptr(self._imb), imbp, nx, ny, nz, self._ptr(self._partials), self._ptr(self._loads),
                  self._stream())
        self._cache.pop("hydro_forces", None)
        self._cache.pop("hydro_torques", None)
        self._cache.pop("loads", None)

    # --------------------------------------------------------------- readouts
    def _moments(self, fields: bool, with_state: bool = True):
        self._sync_steps()
        torch = _torch()
        nx, ny, nz = self._shape
        if fields and self._rho_u is None:
            self._rho_u = torch.empty(4 * self._N, dtype=torch.float64, device=self.device)
        fx, fy, fz = self._force
        _lib.call("psc_moments", self._ptr(self._a) if with_state else None, self._storage, self._layout,
                  self._ptr(self._csum), self._ptr(self._imb), nx, ny, nz, nz if self._state_flags else 0, self._s,
                  fx, fy, fz,
                  self._ptr(self._rho_u) if fields else None, self._ptr(self._mom_partials), self._ptr(self._sums),
                  self._stream())
        return self._np(self._sums)

    def _fields(self):
        if "fields" not in self._cache:
            sums = self._moments(True)
            self._cache["sums"] = sums
            self._cache["fields"] = self._np(self._rho_u).reshape((4,) + self._shape)
        return self._cache["fields"]

    def readout_sums(self) -> dict:
        """FP64 sums of the post-stream state (see psc_moments in psc_hip.h)."""
        if "sums" not in self._cache:
            self._cache["sums"] = self._moments(False)
        s = self._cache["sums"]
        return {
            "solid": s[0], "fluid": s[1], "u": s[2:5].copy(), "fluid_u": s[5:8].copy(), "imb": s[8:11].copy(),
            "rho": s[11], "nonfinite": int(s[12]),
        }

    def _geometry_sums(self):
        if "gsums" not in self._cache:
            self._cache["gsums"] = self._moments(False, with_state=False)
        return self._cache["gsums"]

    @property
    def rho(self):
        return self._fields()[0]

    @property
    def ux(self):
        return self._fields()[1]

    @property
    def uy(self):
        return self._fields()[2]

    @property
    def uz(self):
        return self._fields()[3]

    @property
    def f(self):
        """Post-stream distributions (19, n, n, n) float64 (psc3d.py:242)."""
        if "f" not in self._cache:
            torch = _torch()
            out = torch.empty(19 * self._N, dtype=torch.float64, device=self.device)
            nx, ny, nz = self._shape
            _lib.call("psc_export_f64", self._ptr(self._compact_state()), self._storage, self._layout, nx, ny, nz,
                      self._ptr(out), self._stream())
            self._cache["f"] = self._np(out).reshape((19,) + self._shape)
        return self._cache["f"]

    @f.setter
    def f(self, value):
        self._sync_steps()  # a held step came before this write
        torch = _torch()
        arr = np.ascontiguousarray(np.asarray(value, dtype=np.float64).reshape(19 * self._N))
        src = torch.from_numpy(arr).to(self.device)
        nx, ny, nz = self._shape
        dst = self._compact_state() if self._ps != self._N else self._a
        _lib.call("psc_import_f64", self._ptr(src), self._storage, self._layout, nx, ny, nz, self._ptr(dst),
                  self._stream())
        if dst is not self._a:
            self._a.view(19, self._ps)[:, :self._N].copy_(dst.view(19, self._N))
        _torch().cuda.current_stream(self.device).synchronize()
        self._cache.clear()

    def _compact_state(self):
        """The state with planes of exactly N cells (self._a itself unless its
        planes are padded to 64 cells)."""
        self._sync_steps()
        if self._ps == self._N:
            return self._a
        return self._a.view(19, self._ps)[:, :self._N].contiguous().view(-1)

    def _alpha_sum_np(self):
        if "asum" not in self._cache:
            cnt = self._np(self._csum).astype(np.float64)
            self._cache["asum"] = (cnt / float(self._s ** 3)).reshape(self._shape)
        return self._cache["asum"]

    @property
    def solid(self):
        """clip(sum_p alpha_p, 0, 1) (psc3d.py:175 / 344)."""
        if "solid" not in self._cache:
            self._cache["solid"] = np.clip(self._alpha_sum_np(), 0.0, 1.0)
        return self._cache["solid"]

    def _imb_np(self):
        if "imb" not in self._cache:
            self._sync_steps()
            mask = self._np(self._csum).reshape(self._shape) != 0
            # interleaved on the device: imb[3 cell + axis]
            imb = np.moveaxis(self._np(self._imb).astype(np.float64).reshape(self._shape + (3,)), -1, 0)
            self._cache["imb"] = np.where(mask[None], imb, 0.0)
        return self._cache["imb"]

    @property
    def imb_mx(self):
        return self._imb_np()[0]

    @property
    def imb_my(self):
        return self._imb_np()[1]

    @property
    def imb_mz(self):
        return self._imb_np()[2]

    def _loads_np(self):
        if "loads" not in self._cache:
            self._flush_loads()
            self._cache["loads"] = self._np(self._loads).reshape(-1, 6)
        return self._cache["loads"]

    @property
    def hydro_forces(self):
        if "hydro_forces" not in self._cache:
            self._cache["hydro_forces"] = self._loads_np()[:, 0:3].copy()
        return self._cache["hydro_forces"]

    @hydro_forces.setter
    def hydro_forces(self, value):
        self._cache["hydro_forces"] = np.asarray(value, dtype=float)

    @property
    def hydro_torques(self):
        if "hydro_torques" not in self._cache:
            self._cache["hydro_torques"] = self._loads_np()[:, 3:6].copy()
        return self._cache["hydro_torques"]

    @hydro_torques.setter
    def hydro_torques(self, value):
        self._cache["hydro_torques"] = np.asarray(value, dtype=float)

    @property
    def momentum_residual(self) -> float:
        """|sum_p F_p + sum_x imb| (psc3d.py:280-281)."""
        if "residual" not in self._cache:
            loads = self._loads_np()
            total = loads[:, 0:3].sum(axis=0) if len(loads) else np.zeros(3)
            fluid = self._geometry_sums()[8:11]
            self._cache["residual"] = float(np.linalg.norm(total + fluid))
        return self._cache["residual"]

    def solid_velocity(self):
        """Per-cell solid velocity sum_p alpha_p (U + omega x r) / sum alpha
        (psc3d.py:181-206 / 347-373), as three (n, n, n) arrays."""
        self._refresh_geometry()
        if not self._moving or self._us is None:
            z = np.zeros(self._shape)
            return z, z.copy(), z.copy()
        asum = self._alpha_sum_np()
        num = self._np(self._us_num_planes()).astype(np.float64).reshape((3,) + self._shape)
        mask = asum > 1.0e-14
        out = []
        for a in range(3):
            u = num[a].copy()
            u[mask] /= asum[mask]
            out.append(u)
        return tuple(out)

    def _us_num_planes(self):
        """The solid-velocity numerator planes sum_p alpha_p (U + omega x r) at
        every covered cell.  The owner table keeps them at shared cells only,
        so they are voxelized afresh into scratch buffers for this readout."""
        if self._geom_mode != "owners":
            return self._us
        torch = _torch()
        N, dev = self._N, self.device
        tabs = self._tables
        parts, boxes, chunks, _ = self._dev_tables
        C = int(tabs.chunks.shape[0])
        cap = max(1, C * tabs.chunk_cells)
        csum = torch.zeros(N, dtype=torch.int32, device=dev)
        us = torch.zeros(3 * N, dtype=self._fdt, device=dev)
        ent = (torch.empty(cap, dtype=torch.int32, device=dev), torch.empty(cap, dtype=torch.int32, device=dev),
               torch.empty(cap, dtype=torch.int16, device=dev))
        nnz = torch.empty(max(1, C), dtype=torch.int32, device=dev)
        _lib.call("psc_voxelize", self._ptr(parts), self._ptr(boxes), self._ptr(chunks), C, tabs.chunk_cells,
                  *self._shape, 0, self._s, *(self._ptr(t) for t in ent), self._ptr(nnz), self._ptr(csum),
                  self._ptr(us), 1, 0, self._stream())
        return us

    def _diagnostics(self) -> dict:
        sums = self.readout_sums()
        loads = self._loads_np()
        if len(loads):
            mf = float(np.max(np.linalg.norm(loads[:, 0:3], axis=1)))
            mt = float(np.max(np.linalg.norm(loads[:, 3:6], axis=1)))
        else:
            mf = mt = 0.0
        total = loads[:, 0:3].sum(axis=0) if len(loads) else np.zeros(3)
        residual = float(np.linalg.norm(total + sums["imb"]))
        self._cache["residual"] = residual
        return {
            "solid_volume_lu": float(sums["solid"]),
            "mean_uz": float(sums["u"][2] / self._N),
            "momentum_residual": residual,
            "max_force": mf,
            "max_torque": mt,
        }

    # --------------------------------------------------------------- physics
    def mean_uz(self) -> float:
        """Mean of the post-stream u_z over all cells, np.mean(solver.uz) of
        the drivers (couple_psc_liggghts.py:303), from the FP64 GPU sums
        without copying the field."""
        return float(self.readout_sums()["u"][2] / self._N)

    def permeability(self, axis=None) -> float:
        """k = nu sum((1 - solid) u_axis) / N / F_axis, nu = (tau - 1/2)/3
        (run_psc_permeability_resolution.py:144-150; axis defaults to the body
        force direction)."""
        if axis is None:
            axis = int(np.argmax(np.abs(self._force)))
        F = self._force[axis]
        sums = self.readout_sums()
        nu = (self.tau - 0.5) / 3.0
        return float(nu * (sums["fluid_u"][axis] / self._N) / max(F, 1.0e-300))

    def fluid_statistics(self, axis: int = 2):
        """(intrinsic, superficial, porosity) of u_axis
        (run_precision_validation.py:142-147)."""
        sums = self.readout_sums()
        vol = float(sums["fluid"])
        intrinsic = float(sums["fluid_u"][axis] / max(vol, 1.0e-30))
        superficial = float(sums["fluid_u"][axis] / self._N)
        return intrinsic, superficial, vol / float(self._N)

    def state_dict(self) -> dict:
        """Distributions and step count for checkpoint/resume."""
        return {"f": self.f.copy(), "step": self._step_count, "precision": self.precision}

    def load_state_dict(self, state: dict) -> None:
        self.f = state["f"]
        self._step_count = int(state.get("step", 0))

    # dense-layout views (psc3d.py:141-142, 176-178); P * n^3 arrays
    @property
    def alpha_p(self):
        if (not self._geometry_built):
            self.rebuild_geometry()
        ids, vals = self._sparse_lists()
        out = np.zeros((len(ids),) + self._shape)
        for p, (i, v) in enumerate(zip(ids, vals)):
            out[p].ravel()[i] = v
        return out

    @property
    def owner_weights(self):
        ap = self.alpha_p
        asum = self._alpha_sum_np()
        out = np.zeros_like(ap)
        mask = asum > 1.0e-14
        out[:, mask] = ap[:, mask] / asum[mask]
        return out

    def _sparse_lists(self):
        """Per-particle (ids, alpha) of cells with alpha > 0, box order
        (the reference's node_ids / alpha_vals, psc3d.py:339-343)."""
        if "lists" not in self._cache:
            tabs = self._tables
            C = int(tabs.chunks.shape[0])
            CK = tabs.chunk_cells
            nnz = self._np(self._chunk_nnz)[:C].astype(np.int64)
            cells = self._np(self._ent[0][: max(1, C * CK)]).view(np.uint32)
            cnts = self._np(self._ent[2][: max(1, C * CK)]).view(np.uint16)
            ids, vals = [], []
            s3 = float(self._s ** 3)
            for p in range(tabs.particles.shape[0]):
                c0, c1 = int(tabs.chunk_first[p]), int(tabs.chunk_first[p + 1])
                sel = [np.arange(c * CK, c * CK + nnz[c]) for c in range(c0, c1)]
                idx = np.concatenate(sel) if sel else np.zeros(0, dtype=np.int64)
                ids.append(cells[idx].astype(np.int64))
                vals.append(cnts[idx].astype(np.float64) / s3)
            self._cache["lists"] = (ids, vals)
        return self._cache["lists"]


class SparsePscD3Q19(PscD3Q19):
    """Sparse-geometry variant (psc3d.py:284-407): same step, per-particle
    cell lists.  On the GPU both classes share the bounding-box count
    layout; ``geometry_workers`` is accepted for compatibility and unused."""

    _LAYOUT_NAME = "sparse"

    def __init__(self, n, particles, tau: float = 0.72, subsamples: int = 3, body_force_z: float = 0.0,
                 static_geometry: bool = False, geometry_workers: int = 1, **kw):
        super().__init__(n, particles, tau=tau, subsamples=subsamples, body_force_z=body_force_z,
                         static_geometry=static_geometry, **kw)
        self.geometry_workers = max(1, int(geometry_workers))

    @property
    def node_ids(self):
        if not self._geometry_built:
            return []
        return self._sparse_lists()[0]

    @property
    def alpha_vals(self):
        if not self._geometry_built:
            return []
        return self._sparse_lists()[1]

    @property
    def alpha_sum_flat(self):
        return self._alpha_sum_np().ravel()


def sphere_volume(r: float) -> float:
    return 4.0 / 3.0 * math.pi * r ** 3
