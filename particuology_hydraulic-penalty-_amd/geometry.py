"""Host-side geometry tables for the voxelizer and the particle-load kernels.

The kernels take three small tables (layouts in include/psc_hip.h):

* particles (P, 10) float64: x, y, z, r, vx, vy, vz, wx, wy, wz;
* boxes (P, 8) int32: cell bounding box origin and extent, entry offset;
* chunks (C, 2) int32: (particle, first box entry) work items, plus
  chunk_first (P+1,) int32, the first chunk of every particle.

Bounding boxes follow src/lbm_dem/psc3d.py:73-78: [floor(c - r - 1),
ceil(c + r + 1)) per axis, clamped to [0, n) (the reference) or left unclamped
and wrapped by the kernels (``periodic=True``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FIELDS = ("x", "y", "z", "r", "vx", "vy", "vz", "wx", "wy", "wz")
MAX_ENTRIES = 2 ** 31 - 1
MAX_BOX_EXTENT = 1024  # box-local coordinates are packed in 10 bits per axis


def pack_particles(particles) -> np.ndarray:
    """(P, 10) float64 table from Particle3D-like objects or an array."""
    if isinstance(particles, np.ndarray):
        t = np.asarray(particles, dtype=np.float64)
        if t.ndim != 2 or t.shape[1] != 10:
            raise ValueError(f"particle table must be (P, 10), got {t.shape}")
        return np.ascontiguousarray(t)
    if len(particles) == 0:
        return np.zeros((0, 10), dtype=np.float64)
    return np.array([[float(getattr(p, k)) for k in FIELDS] for p in particles], dtype=np.float64)


@dataclass
class GeometryTables:
    particles: np.ndarray    # (P, 10) float64
    boxes: np.ndarray        # (P, 8) int32
    chunks: np.ndarray       # (C, 2) int32
    chunk_first: np.ndarray  # (P + 1,) int32
    n_entries: int           # total box entries (size of the count array)
    moving: bool             # any non-zero U or omega
    chunk_cells: int


def build_tables(table: np.ndarray, shape, periodic: bool = False, chunk_cells: int = 2048) -> GeometryTables:
    """Bounding boxes, entry offsets and chunk work items for a particle table."""
    nx, ny, nz = (int(v) for v in shape)
    P = table.shape[0]
    lo = np.empty((P, 3), dtype=np.int64)
    hi = np.empty((P, 3), dtype=np.int64)
    for a, n in enumerate((nx, ny, nz)):
        c = table[:, a]
        r = table[:, 3]
        l0 = np.floor(c - r - 1.0).astype(np.int64)
        h0 = np.ceil(c + r + 1.0).astype(np.int64)
        if periodic:
            if np.any(h0 - l0 > n):
                raise ValueError("periodic geometry needs every sphere box to fit in the lattice "
                                 f"(extent {int(np.max(h0 - l0))} > {n} on axis {a})")
        else:
            l0 = np.maximum(l0, 0)
            h0 = np.minimum(h0, n)
        lo[:, a] = l0
        hi[:, a] = np.maximum(h0, l0)
    ext = hi - lo
    if P and int(ext.max()) > MAX_BOX_EXTENT:
        raise ValueError(f"sphere box extent {int(ext.max())} exceeds {MAX_BOX_EXTENT} cells")
    vols = ext[:, 0] * ext[:, 1] * ext[:, 2]
    offs = np.zeros(P + 1, dtype=np.int64)
    np.cumsum(vols, out=offs[1:])
    total = int(offs[-1])
    nch_total = int(((vols + chunk_cells - 1) // chunk_cells).sum()) * chunk_cells
    if total > MAX_ENTRIES or nch_total > MAX_ENTRIES:
        raise ValueError(f"{total} box entries exceed the int32 entry space")
    boxes = np.zeros((P, 8), dtype=np.int32)
    boxes[:, 0:3] = lo
    boxes[:, 3:6] = ext
    boxes[:, 6] = offs[:-1]
    nch = (vols + chunk_cells - 1) // chunk_cells
    chunk_first = np.zeros(P + 1, dtype=np.int64)
    np.cumsum(nch, out=chunk_first[1:])
    C = int(chunk_first[-1])
    pid = np.repeat(np.arange(P, dtype=np.int64), nch)
    k = np.arange(C, dtype=np.int64) - np.repeat(chunk_first[:-1], nch)
    chunks = np.stack([pid, k * chunk_cells], axis=1).astype(np.int32) if C else np.zeros((0, 2), np.int32)
    moving = bool(P and np.any(table[:, 4:10] != 0.0))
    return GeometryTables(table, boxes, chunks, chunk_first.astype(np.int32), total, moving, chunk_cells)


def box_cells(boxes: np.ndarray, p: int, shape):
    """Unwrapped and wrapped cell coordinates of particle p's box entries, in
    entry order (ix, iy, iz lexicographic)."""
    x0, y0, z0, ex, ey, ez = (int(v) for v in boxes[p, :6])
    ix, iy, iz = np.meshgrid(np.arange(x0, x0 + ex), np.arange(y0, y0 + ey), np.arange(z0, z0 + ez), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    nx, ny, nz = shape
    return (ix, iy, iz), (ix % nx, iy % ny, iz % nz)
