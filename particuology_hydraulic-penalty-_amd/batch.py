"""Many independent pore windows stepped together on one GPU.

The reference runs one window per process (``SparsePscD3Q19`` per case, sharded
by manifest row, run_synced_campaign.py:59-73).  Here a batch of periodic
windows of uneven sizes shares one set of device buffers:

* state: the 19 FP32 deviation planes of each window, window after window;
* count_sum / imb: one array over the concatenated cells of all windows;
* particles, boxes, chunks and voxel entries: concatenated, with global
  particle ids, so one launch of the load kernels covers every window.

Each step is ONE launch of ``psc_step_batched`` (work items are spans of
cells or groups of whole rows of a window), followed by the particle-load
reduction.  ``run`` captures a few steps into a hipGraph (torch.cuda.CUDAGraph
around the C-ABI launches) and replays it; there each launch
(``psc_step_batched_loads``) also carries the load reduction of the two
previous steps, so a step stays one launch.  The launch configuration
(workgroup size, item mode, non-temporal access, occupancy cap) is autotuned
on first use.
Every window is the reference's ``collide_stream_psc`` (psc3d.py:208-253) with
periodic geometry; per-window permeability follows
run_psc_permeability_resolution.py:144-150.
"""

from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .geometry import build_tables, pack_particles

MAX_WINDOW_CELLS = 1 << 22
FUSE_PERIOD = 5  # fused loads: one load item in every 5 at the head of a launch's item sequence
# Default z-row padding: none.  A padding cell is never written, so a row pitch
# above the row length leaves every 64-byte memory granule that holds padding
# partly written by the step's stores, and a partial granule costs HBM a
# read-modify-write: empty windows of 137, 141, 101 and 75 cells per row
# stream at 4.1-4.5 TB/s with rows padded to a multiple of 4 cells and at
# 5.3-5.6 TB/s unpadded (spans of 256 cells, whose wavefront stores then
# cover whole aligned 256-byte runs).
ROW_ALIGN = 1
PLANE_ALIGN = 64  # direction planes padded to a multiple of 64 cells (256-byte aligned planes)


class BatchedPscD3Q19:
    def __init__(self, windows, tau: float = 0.8, subsamples: int = 3, body_force=(0.0, 0.0, 1.0e-6),
                 device=None, chunk_cells: int = 2048, launch="auto", row_align: int = ROW_ALIGN):
        import torch
        self._lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.PscError("BatchedPscD3Q19 needs a ROCm GPU (gfx950)")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.tau = float(tau)
        self.subsamples = int(subsamples)
        self.body_force = tuple(float(v) for v in body_force)
        self._s = max(1, self.subsamples)
        self.shapes = []
        tables = []
        for shape, parts in windows:
            shp = (int(shape),) * 3 if np.isscalar(shape) else tuple(int(v) for v in shape)
            if shp[0] * shp[1] * shp[2] > MAX_WINDOW_CELLS:
                raise ValueError(f"window {shp} exceeds {MAX_WINDOW_CELLS} cells")
            self.shapes.append(shp)
            tables.append(pack_particles(parts))
        self.nwin = len(self.shapes)
        # rows padded to row_align cells (32: every row a whole number of 128-byte cache lines)
        ra = max(1, int(row_align))
        self.row_align = ra
        self.pitch = [-(-c // ra) * ra for _, _, c in self.shapes]
        # cells per window incl. row padding, and per direction plane (64-cell multiple: 256-byte aligned planes,
        # the layout psc_moments assumes for pitch_z > nz)
        ncell = np.array([a * b * p for (a, b, _), p in zip(self.shapes, self.pitch)], dtype=np.int64)
        self.plane = np.array([-(-int(v) // PLANE_ALIGN) * PLANE_ALIGN for v in ncell], dtype=np.int64)
        self.cell_off = np.concatenate([[0], np.cumsum(self.plane)[:-1]]).astype(np.int64)
        self.state_off = 19 * self.cell_off
        self.ncell = ncell                       # padded cells per window
        self.nlive = np.array([a * b * c for a, b, c in self.shapes], dtype=np.int64)
        self.Ntot = int(self.plane.sum())
        dev = self.device
        self._a = torch.zeros(19 * self.Ntot, dtype=torch.float32, device=dev)
        self._b = torch.zeros(19 * self.Ntot, dtype=torch.float32, device=dev)
        self._csum = torch.zeros(self.Ntot, dtype=torch.int32, device=dev)
        self._imb = torch.zeros(3 * self.Ntot, dtype=torch.float32, device=dev)
        # window tables and work items per (workgroup size, item mode), built on demand
        self._item_tables = {}
        # geometry: per-window boxes and chunks, concatenated with global ids
        self.tables = tables
        self.pbase = np.concatenate([[0], np.cumsum([t.shape[0] for t in tables])[:-1]]).astype(np.int64)
        tabs = [build_tables(t, shp, periodic=True, chunk_cells=chunk_cells) for t, shp in zip(tables, self.shapes)]
        self.chunk_cells = chunk_cells
        self.cbase = np.concatenate([[0], np.cumsum([t.chunks.shape[0] for t in tabs])[:-1]]).astype(np.int64)
        parts = np.concatenate([t.particles for t in tabs]) if tabs else np.zeros((0, 10))
        boxes = np.concatenate([t.boxes for t in tabs])
        chunks = np.concatenate([t.chunks + np.array([self.pbase[w], 0], dtype=np.int32)
                                 for w, t in enumerate(tabs)])
        cfirst = [0]
        for w, t in enumerate(tabs):
            cfirst.extend((t.chunk_first[1:] + self.cbase[w]).tolist())
        self.P = parts.shape[0]
        self.C = chunks.shape[0]
        self._parts = torch.from_numpy(np.ascontiguousarray(parts)).to(dev)
        self._boxes = torch.from_numpy(np.ascontiguousarray(boxes)).to(dev)
        self._chunks = torch.from_numpy(np.ascontiguousarray(chunks.astype(np.int32))).to(dev)
        self._cfirst = torch.from_numpy(np.asarray(cfirst, dtype=np.int32)).to(dev)
        cap = max(1, self.C * chunk_cells)
        self._ent = (torch.empty(cap, dtype=torch.int32, device=dev), torch.empty(cap, dtype=torch.int32, device=dev),
                     torch.empty(cap, dtype=torch.int16, device=dev))
        self._nnz = torch.zeros(max(1, self.C), dtype=torch.int32, device=dev)
        self._partials = torch.zeros(6 * max(1, self.C), dtype=torch.float64, device=dev)
        self._loads = torch.zeros((max(1, self.P), 6), dtype=torch.float64, device=dev)
        self._sums = torch.zeros(self.nwin * _lib.PSC_NSUM, dtype=torch.float64, device=dev)
        self._mom_partials = torch.zeros(int(self._lib.psc_moments_blocks(int(ncell.max()))) * _lib.PSC_NSUM,
                                         dtype=torch.float64, device=dev)
        self._flags = 0
        self._grid = 0
        self._mode = "spans"
        self._tuned = launch != "auto"
        if launch != "auto":
            self._flags, self._grid = int(launch[0]), int(launch[1])
            if len(launch) > 2:
                self._mode = str(launch[2])
        self._graph = None
        self._imb_alt = None
        self._partials_alt = None
        # run(): reduce each step's loads inside the next two step launches (False: on a
        # side stream, concurrently with the next step)
        self.fuse_loads = True
        # placement period of the load items (None = as the launch tuner measured it)
        self.fuse_period = None
        self._period_tuned = FUSE_PERIOD
        self._steps = 0
        self._voxelize(tabs)

    # ----------------------------------------------------------------- setup
    def _stream(self):
        import torch
        return torch.cuda.current_stream(self.device).cuda_stream

    def _voxelize(self, tabs):
        s = self._stream()
        ec, el, en = self._ent
        CK = self.chunk_cells
        for w, t in enumerate(tabs):
            nch = t.chunks.shape[0]
            if nch == 0:
                continue
            cb = int(self.cbase[w])
            nx, ny, nz = self.shapes[w]
            _lib.call("psc_voxelize", self._parts.data_ptr(), self._boxes.data_ptr(),
                      self._chunks.data_ptr() + 8 * cb, nch, CK, nx, ny, nz, self.pitch[w], self._s,
                      ec.data_ptr() + 4 * cb * CK, el.data_ptr() + 4 * cb * CK, en.data_ptr() + 2 * cb * CK,
                      self._nnz.data_ptr() + 4 * cb, self._csum.data_ptr() + 4 * int(self.cell_off[w]), None, 0,
                      int(self.cell_off[w]), s)

    def _items_for(self, bs: int, mode: str):
        """(window table, item list, item count) for workgroups of ``bs``
        threads.  mode "spans": items of bs consecutive cells (rows may be
        split between workgroups); mode "rows": items of floor(bs/pz) whole
        rows, so a row's cache lines are only ever touched by one workgroup
        (windows whose rows exceed bs fall back to spans)."""
        import torch
        key = (int(bs), mode)
        if key in self._item_tables:
            return self._item_tables[key]
        if mode not in ("spans", "rows"):
            raise ValueError(f"unknown item mode {mode!r}")
        wt = (_lib.PscWindow * self.nwin)()
        items = []
        for w, (nx, ny, nz) in enumerate(self.shapes):
            pz = self.pitch[w]
            wt[w].state_off = int(self.state_off[w])
            wt[w].cell_off = int(self.cell_off[w])
            wt[w].nx, wt[w].ny, wt[w].nz, wt[w].pz = nx, ny, nz, pz
            wt[w].inv_pz = 1.0 / pz
            wt[w].inv_nypz = 1.0 / (ny * pz)
            wt[w].ps = int(self.plane[w])
            wt[w].fx, wt[w].fy, wt[w].fz = self.body_force
            rows = bs // pz if mode == "rows" else 0
            span = rows * pz if rows >= 1 else bs
            wt[w].span = span
            starts = np.arange(0, int(self.ncell[w]), span, dtype=np.int64)
            items.append(np.stack([np.full(len(starts), w, dtype=np.int64), starts], axis=1))
        raw = np.frombuffer(bytes(wt), dtype=np.uint8)
        windows = torch.from_numpy(raw.copy()).to(self.device)
        it = np.concatenate(items).astype(np.int32).reshape(-1)
        entry = (windows, torch.from_numpy(it).to(self.device), len(it) // 2)
        self._item_tables[key] = entry
        return entry

    # ------------------------------------------------------------------ step
    def _launch_step(self, a, b, flags, grid, stream, mode=None, imb=None):
        bs = 512 if (flags & _lib.PSC_BATCH_WG512) else 256
        windows, items, nitems = self._items_for(bs, mode or self._mode)
        imb = self._imb if imb is None else imb
        _lib.call("psc_step_batched", a.data_ptr(), b.data_ptr(), self._csum.data_ptr(), imb.data_ptr(),
                  self.Ntot, windows.data_ptr(), items.data_ptr(), nitems, self._s, self.tau,
                  flags, grid, stream)

    def _launch_loads(self, stream, imb=None):
        if self.P == 0:
            return
        ec, el, en = self._ent
        imb = self._imb if imb is None else imb
        _lib.call("psc_particle_loads", self._parts.data_ptr(), self._boxes.data_ptr(), self._chunks.data_ptr(),
                  self._cfirst.data_ptr(), self.P, self.C, self.chunk_cells, ec.data_ptr(), el.data_ptr(),
                  en.data_ptr(), self._nnz.data_ptr(), self._csum.data_ptr(), imb.data_ptr(), 1, 1, 1,
                  self.Ntot, self._partials.data_ptr(), self._loads.data_ptr(), stream)

    def _loads_tables(self):
        t = _lib.PscLoadsTables()
        ec, el, en = self._ent
        t.particles, t.boxes, t.chunks = self._parts.data_ptr(), self._boxes.data_ptr(), self._chunks.data_ptr()
        t.chunk_first = self._cfirst.data_ptr()
        t.nparticles, t.nchunks, t.chunk_cells = self.P, self.C, self.chunk_cells
        t.ent_cell, t.ent_loc, t.ent_cnt = ec.data_ptr(), el.data_ptr(), en.data_ptr()
        t.chunk_nnz, t.count_sum = self._nnz.data_ptr(), self._csum.data_ptr()
        return t

    def _ensure_fused_buffers(self):
        import torch
        if self._imb_alt is None:
            self._imb_alt = torch.zeros_like(self._imb)
        if self._partials_alt is None:
            self._partials_alt = torch.zeros_like(self._partials)

    def effective_fuse_period(self) -> int:
        """Placement period of the load items in fused step launches."""
        return max(1, int(self._period_tuned if self.fuse_period is None else self.fuse_period))

    def _launch_fused(self, a, b, q, pc, pf, stream, flags=None, mode=None, period=None):
        """One step from a into b writing imb[q], plus the chunk sums of
        imb[1-q] (pc) and the per-particle sums of partials[q] (pf)."""
        flags = self._flags if flags is None else flags
        period = self.effective_fuse_period() if period is None else period
        bs = 512 if (flags & _lib.PSC_BATCH_WG512) else 256
        windows, items, nitems = self._items_for(bs, mode or self._mode)
        imbs = (self._imb, self._imb_alt)
        prt = (self._partials, self._partials_alt)
        t = self._loads_tables()
        _lib.call("psc_step_batched_loads", a.data_ptr(), b.data_ptr(), self._csum.data_ptr(), imbs[q].data_ptr(),
                  self.Ntot, windows.data_ptr(), items.data_ptr(), nitems, self._s, self.tau, flags,
                  self._grid, ct.byref(t), imbs[1 - q].data_ptr() if pc else None, prt[1 - q].data_ptr(),
                  prt[q].data_ptr() if pf else None, self._loads.data_ptr(), max(1, int(period)), stream)

    def step(self):
        """One step of every window (collide_stream_psc) plus the loads."""
        self._tune()
        s = self._stream()
        self._launch_step(self._a, self._b, self._flags, self._grid, s)
        self._a, self._b = self._b, self._a
        self._launch_loads(s)
        self._steps += 1

    def _tune(self):
        """Pick the workgroup size (256 or 512), item mode (cell spans or
        whole rows), non-temporal access (state, and separately imb) and
        occupancy cap, timing each candidate with HIP events on the solver's
        own buffers (the state is not modified: A -> B).  The launch timed is
        the one run() uses: with fused loads, the step that also carries the
        load reduction of two earlier steps."""
        if self._tuned:
            return
        import torch
        from . import autotune
        s = self._stream()
        fused = self.fuse_loads and self.P > 0
        if fused:
            self._ensure_fused_buffers()
        def timed(flags, mode, reps):
            def launch():
                if fused:
                    self._launch_fused(self._a, self._b, 0, True, True, s, flags, mode)
                else:
                    self._launch_step(self._a, self._b, flags, 0, s, mode)
            launch()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                launch()
            e1.record()
            e1.synchronize()
            return e0.elapsed_time(e1) / reps

        results = []
        for bs in (256, 512):
            for mode in ("rows", "spans"):
                for nt in (0, 1):
                    for bpc in (0, 2, 4, 8):
                        for inb in (0, 1):
                            flags = (autotune.flags_of(nt, bpc, 0) | (_lib.PSC_BATCH_WG512 if bs == 512 else 0)
                                     | (_lib.PSC_BATCH_IMB_NT if inb else 0))
                            results.append((timed(flags, mode, 3), flags, mode))
        # the leaders are often within 1% of each other: time them again, longer
        best = min((timed(flags, mode, 12), flags, mode) for _, flags, mode in sorted(results)[:4])
        self._flags, self._grid, self._mode = best[1], 0, best[2]
        if fused:  # then the placement period of the load items, for that launch
            bestp = None
            for period in (1, 3, 5, 7, 9):
                self._launch_fused(self._a, self._b, 0, True, True, s, period=period)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(6):
                    self._launch_fused(self._a, self._b, 0, True, True, s, period=period)
                e1.record()
                e1.synchronize()
                ms = e0.elapsed_time(e1) / 6
                if bestp is None or ms < bestp[0]:
                    bestp = (ms, period)
            self._period_tuned = bestp[1]
        self._tuned = True

    def run(self, steps: int, graph: bool = True, graph_steps: int = 8) -> None:
        """``steps`` steps.  With graph=True, ``graph_steps`` (even) steps are
        captured once into a hipGraph and replayed.  The steps alternate
        between two imb buffers.  With fuse_loads (the default), the launch of
        step t also sums the chunks of step t-1's imb and adds up, per
        particle, the chunk partials of step t-2 (two partial-sum buffers
        alternate too); the loads still pending after the last step are
        reduced at the end.  Otherwise the load reduction of step t runs on a
        side stream while step t+1 runs on the main one, step t+2 waits for
        it (it overwrites the imb those loads read), and the graph ends
        joined.  Results are those of stepping one by one."""
        import torch
        steps = int(steps)
        self._tune()
        # the item tables are uploaded here, never inside the graph capture
        self._items_for(512 if (self._flags & _lib.PSC_BATCH_WG512) else 256, self._mode)
        G = max(2, int(graph_steps) // 2 * 2)
        if self.fuse_loads and self.P > 0 and steps > 0:
            self._run_fused(steps, graph, G)
            return
        if not graph or steps < G:
            for _ in range(steps):
                self.step()
            return
        if self._graph is None:
            a, b = self._a, self._b
            if self._imb_alt is None:
                self._imb_alt = torch.zeros_like(self._imb)
            imbs = (self._imb, self._imb_alt)
            side = torch.cuda.Stream(device=self.device)
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                main = torch.cuda.current_stream(self.device)
                done = []
                for j in range(G):
                    src, dst = (a, b) if j % 2 == 0 else (b, a)
                    if j >= 2:
                        main.wait_event(done[j - 2])  # imbs[j % 2] is free again
                    self._launch_step(src, dst, self._flags, self._grid, main.cuda_stream, imb=imbs[j % 2])
                    side.wait_stream(main)
                    self._launch_loads(side.cuda_stream, imb=imbs[j % 2])
                    ev = torch.cuda.Event()
                    ev.record(side)
                    done.append(ev)
                main.wait_stream(side)
            self._graph = (g, a, b, side, done)
        g, a, b = self._graph[:3]
        if self._a is not a:  # odd number of eager steps since capture
            self.step()
            steps -= 1
        for _ in range(steps // G):
            g.replay()
        self._steps += G * (steps // G)
        for _ in range(steps % G):
            self.step()

    def _run_fused(self, steps, graph, G):
        import torch
        self._ensure_fused_buffers()
        s = self._stream()
        if not hasattr(self, "_buf0"):
            self._buf0 = self._a.data_ptr()
        q = 0 if self._a.data_ptr() == self._buf0 else 1
        pc = pf = False
        done = 0

        def direct(n):
            nonlocal q, pc, pf
            for _ in range(n):
                self._launch_fused(self._a, self._b, q, pc, pf, s)
                self._a, self._b = self._b, self._a
                pf, pc, q = pc, True, 1 - q

        head = min(steps, 2)
        direct(head)
        done += head
        if graph and steps - done >= G:
            key = ("fused", q, G, self._flags, self._mode, self._grid, self.effective_fuse_period())
            if getattr(self, "_fgraph", None) is None or self._fgraph[0] != key:
                a, b = self._a, self._b
                torch.cuda.synchronize(self.device)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    cs = torch.cuda.current_stream(self.device).cuda_stream
                    for j in range(G):
                        src, dst = (a, b) if j % 2 == 0 else (b, a)
                        self._launch_fused(src, dst, (q + j) % 2, True, True, cs)
                self._fgraph = (key, g, a, b)
            g = self._fgraph[1]
            while steps - done >= G:
                g.replay()
                done += G
        direct(steps - done)
        # the loads still pending: step t-1 (its imb is intact) and step t
        if pf:
            self._launch_loads(s, imb=(self._imb, self._imb_alt)[q])
        self._launch_loads(s, imb=(self._imb, self._imb_alt)[1 - q])
        self._steps += steps

    # -------------------------------------------------------------- readouts
    def readout_sums(self):
        """(nwin, PSC_NSUM) FP64 sums of each window's post-stream state."""
        s = self._stream()
        for w, (nx, ny, nz) in enumerate(self.shapes):
            _lib.call("psc_moments", self._a.data_ptr() + 4 * int(self.state_off[w]), _lib.PSC_F32_DEV,
                      _lib.PSC_PULL, self._csum.data_ptr() + 4 * int(self.cell_off[w]), None, nx, ny, nz,
                      self.pitch[w], self._s,
                      *self.body_force, None, self._mom_partials.data_ptr(),
                      self._sums.data_ptr() + 8 * _lib.PSC_NSUM * w, s)
        return self._sums.cpu().numpy().reshape(self.nwin, _lib.PSC_NSUM)

    def permeability(self, axis=None):
        """Per-window k = nu sum((1-solid) u_axis) / N_w / F_axis."""
        if axis is None:
            axis = int(np.argmax(np.abs(self.body_force)))
        sums = self.readout_sums()
        nu = (self.tau - 0.5) / 3.0
        return nu * (sums[:, 5 + axis] / self.nlive) / self.body_force[axis]

    def porosity(self):
        sums = self.readout_sums()
        return sums[:, 1] / self.nlive

    def loads(self):
        """Per-window (P_w, 6) FP64 force/torque arrays of the last step."""
        L = self._loads.cpu().numpy()[: self.P]
        out = []
        for w, t in enumerate(self.tables):
            p0 = int(self.pbase[w])
            out.append(L[p0: p0 + t.shape[0]])
        return out

    def window_state(self, w):
        """FP32 deviation state of window w (19, nx, ny, nz), post-collision."""
        n = int(self.ncell[w])
        ps = int(self.plane[w])
        o = int(self.state_off[w])
        nx, ny, nz = self.shapes[w]
        st = self._a[o: o + 19 * ps].cpu().numpy().reshape(19, ps)[:, :n].reshape(19, nx, ny, self.pitch[w])
        return st[..., :nz]
