"""Launch-configuration autotuner for the collide+stream kernel (psc_step).

The FP32 step is HBM-bound; how close it gets to the roofline depends on the
workgroup shape, on non-temporal (streaming) loads/stores and on how many
workgroups are resident per CU (fewer resident waves keep fewer DRAM pages
open at once).  The best point moves with the lattice shape, so it is
measured, not guessed: every candidate is timed with HIP events on the
solver's own buffers (state A -> scratch B; the state is not modified); the
six fastest are timed again, four times longer (the leaders are often within
1% of each other), and the fastest is kept.  Results are cached per (device,
lattice, storage, stream layout, moving, fused loads) in memory and in a JSON
file ($PSC_HIP_TUNE_CACHE, default ~/.cache/psc_hip/step_tuning.json).  When the
native loop fuses the particle loads into the step kernels, candidates are
timed in that loop (on scratch copies of the buffers): the extra load
workgroups shift the best occupancy cap.  There each candidate is also timed
with the steps in place in pairs (PSC_RUN_INPLACE, one state buffer): that
pays where one state buffer fits the Infinity Cache and two do not.  The
leading launches of that quick pass are then timed together with the
placement period of the load workgroups, each pair through graph replays of
LONG_STEPS steps as run() drives them, twice in interleaved rounds (the
fastest time counts): runs that long repeat to about 0.05 %, while the
leaders differ by 0.5-2 %.

All candidates compute bit-identical results; only the launch differs.
"""

from __future__ import annotations

import json
import os
import threading

from . import _lib

_LOCK = threading.Lock()
_MEM: dict = {}

# (block_z, block_y, nontemporal, workgroups-per-CU cap; 0 = none, flat mapping)
SHAPES = [(64, 4), (64, 2), (64, 1), (96, 2), (128, 2), (128, 1), (192, 1), (256, 1)]
# Extra workgroup shapes of the FP64 step (narrow z tiles keep more rows in flight)
FP64_SHAPES = [(32, 4), (32, 8), (16, 16)]
# Workgroup sizes of the flat mapping (PSC_STEP_LINEAR: consecutive cells of
# the lattice's flat index), tried when z rows are not a whole number of
# wavefronts: a (z, y) tile then leaves the lanes past each row end idle and
# splits the 64-byte granules at row ends between workgroups.
FLAT = (64, 128, 256)
NT = (0, 1)
BPC = (0, 2, 3, 4, 6, 8)


def candidates(nz: int, moving: bool = False):
    """(block_z, block_y, non-temporal, workgroups per CU cap, flat) of the
    launch search.  The flat mapping is searched where rows are not a whole
    number of wavefronts, and for moving spheres, whose flat kernel runs at
    six wavefronts per SIMD where the tiled one holds five."""
    out = []
    for bz, by in SHAPES:
        if bz > 64 and nz < bz // 2:
            continue  # more than half the threads of a row would idle
        for nt in NT:
            for bpc in BPC:
                out.append((bz, by, nt, bpc, False))
    if nz % 64 or moving:
        for bz in FLAT:
            for nt in NT:
                for bpc in BPC:
                    out.append((bz, 1, nt, bpc, True))
    return out


def _cache_path():
    return os.environ.get("PSC_HIP_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "psc_hip", "step_tuning.json"))


# Part of every cache key: entries chosen from an earlier candidate set or by an
# earlier selection rule are not reused.
CANDIDATE_SET = 8


def _key(solver):
    import torch
    name = torch.cuda.get_device_properties(solver.device).name
    arch = getattr(torch.cuda.get_device_properties(solver.device), "gcnArchName", "")
    return "|".join(str(v) for v in (name, arch, solver._shape, solver._storage, solver._layout,
                                      bool(getattr(solver, "_moving", False)), getattr(solver, "_geom_mode", None),
                                      "c%d" % CANDIDATE_SET))


def _load_file():
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_file(key, cfg):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load_file()
        data[key] = cfg
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # the cache is an optimisation only


def flags_of(nt: int, bpc: int, cpt: int = 1, flat: bool = False) -> int:
    return ((_lib.PSC_STEP_NONTEMPORAL if nt else 0) | ((int(bpc) & 0xFF) << 8) | ((int(cpt) & 3) << 16)
            | (_lib.PSC_STEP_LINEAR if flat else 0))


def time_config(solver, bz, by, flags, repeats):
    """Average duration (ms) of one launch configuration of the step kernel
    the solver runs (psc_step, or psc_step_owners for moving spheres with an
    owner table).  flags: launch flags; the state's layout flag is added."""
    import torch
    flags = int(flags) | getattr(solver, "_state_flags", 0)
    nx, ny, nz = solver._shape
    fx, fy, fz = solver.body_force
    stream = torch.cuda.current_stream(solver.device)
    h = stream.cuda_stream
    us = solver._ptr(solver._us) if getattr(solver, "_moving", False) else None
    lib = _lib.load()
    if getattr(solver, "_geom_mode", None) == "owners":
        fn, name = lib.psc_step_owners, "psc_step_owners"
        args = (solver._ptr(solver._a), solver._ptr(solver._b), solver._ptr(solver._csum), solver._ptr(solver._own),
                us, solver._ptr(solver._motion), int(solver._tables.particles.shape[0]), solver._ptr(solver._imb), nx,
                ny, nz, solver._s, solver.tau, fx, fy, fz, bz, by, flags, h)
    else:
        fn, name = lib.psc_step, "psc_step"
        args = (solver._ptr(solver._a), solver._ptr(solver._b), solver._ptr(solver._csum), us,
                solver._ptr(solver._imb), solver._storage, solver._layout, nx, ny, nz, solver._s, solver.tau, fx, fy,
                fz, bz, by, flags, h)
    rc = fn(*args)
    _lib.check(rc, name)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(repeats):
        fn(*args)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / repeats


def _fused_scratch(solver):
    """Buffers of the fused timing loop, apart from the solver's own: it steps
    copies, so the solver's state, imb and loads are not modified."""
    import torch
    a = solver._a.clone()
    imb = solver._imb_bufs[0]
    return (a, torch.empty_like(a), torch.zeros_like(imb), torch.zeros_like(imb),
            torch.zeros_like(solver._partials), torch.zeros_like(solver._partials), torch.zeros_like(solver._loads))


def time_fused_config(solver, scratch, bz, by, flags, repeats, period=None, graph=False):
    """Average step duration (ms) of `repeats` steps of the native loop
    (psc_run_steps: direct launches, or graph replays as run() does) with the
    loads fused into the step kernels, on the scratch buffers of
    _fused_scratch."""
    import ctypes as ct
    import torch
    a, b, i0, i1, q0, q1, ld = scratch
    p = solver._plan(False, True)
    p.state_a, p.state_b = a.data_ptr(), b.data_ptr()
    p.imb, p.imb_alt = i0.data_ptr(), i1.data_ptr()
    p.partials, p.partials_alt, p.loads = q0.data_ptr(), q1.data_ptr(), ld.data_ptr()
    # flags may hold PSC_RUN_INPLACE; the state's layout flag is added
    p.block_z, p.block_y, p.flags = bz, by, int(flags) | getattr(solver, "_state_flags", 0)
    if period is not None:
        p.fuse_loads = int(period)
    stream = torch.cuda.current_stream(solver.device)
    h = stream.cuda_stream
    lib = _lib.load()
    parity = ct.c_int32(0)
    _lib.check(lib.psc_run_steps(ct.byref(p), 4, 0, 16, ct.byref(parity), h), "psc_run_steps")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    if graph:  # capture the graphs of the timed run outside the timed region
        _lib.check(lib.psc_prepare_steps(ct.byref(p), repeats, GRAPH_STEPS, ct.byref(parity)), "psc_prepare_steps")
    e0.record(stream)
    rc = lib.psc_run_steps(ct.byref(p), repeats, 1 if graph else 0, GRAPH_STEPS, ct.byref(parity), h)
    e1.record(stream)
    _lib.check(rc, "psc_run_steps")
    e1.synchronize()
    return e0.elapsed_time(e1) / repeats


PERIODS = (1, 3, 5, 7, 9)
GRAPH_STEPS = 32  # steps per captured graph, as run() replays them
LONG_STEPS = 8 * GRAPH_STEPS + 2  # steps of a leader's timed run (two direct steps, then graphs of 7 x 32 and 32)
LEADERS = 6  # launches of the quick pass timed again with every placement period


def tune_period(solver, bz, by, flags, repeats: int = 120, use_cache: bool = True):
    """Placement period of the load workgroups in the fused loop (odd: the
    workgroups go to the 8 XCDs round-robin), timed with the chosen launch."""
    key = _key(solver) + "|period|%d,%d,%d" % (bz, by, flags)
    with _LOCK:
        if use_cache:
            if key in _MEM:
                return _MEM[key]
            disk = _load_file().get(key)
            if disk:
                _MEM[key] = int(disk[0])
                return _MEM[key]
        scratch = _fused_scratch(solver)
        best = None
        for period in PERIODS:
            ms = time_fused_config(solver, scratch, bz, by, flags, repeats, period, graph=True)
            if best is None or ms < best[0]:
                best = (ms, period)
        _MEM[key] = best[1]
        _store_file(key, [best[1], best[0]])
        return best[1]


def tune(solver, budget_ms: float = 400.0, use_cache: bool = True, fused: bool = False):
    """Return (block_z, block_y, flags) of the fastest candidate for solver.

    fused: time each candidate in the native step loop with the loads fused
    into the step kernels (the solver's plan, on scratch copies), instead of
    the plain step kernel."""
    key = _key(solver) + ("|fused" if fused else "")
    with _LOCK:
        if use_cache:
            if key in _MEM:
                return _MEM[key]
            disk = _load_file().get(key)
            if disk:
                cfg = (int(disk[0]), int(disk[1]), int(disk[2]))
                _MEM[key] = cfg
                return cfg
        nz = solver._shape[2]
        if solver._storage == _lib.PSC_F64_PLAIN:
            # the FP64 step's only launch flag is non-temporal state access: shapes x that
            cands = [(bz, by, nt, 0, False) for bz, by in SHAPES + FP64_SHAPES for nt in NT
                     if not (bz > 64 and nz < bz // 2)]
        else:
            cands = candidates(nz, bool(getattr(solver, "_moving", False)))
        # one probe launch sizes the repeat count to the time budget
        probe = time_config(solver, 64, 4, 0, 1)
        reps = max(2, min(50, int(budget_ms / max(probe, 1e-3) / len(cands))))
        if fused:
            scratch = _fused_scratch(solver)
            reps = max(12, reps)  # the loop's two head steps and final flush are amortised

            def timer(bz, by, flags):
                return time_fused_config(solver, scratch, bz, by, flags, reps, period=1)
        else:
            def timer(bz, by, flags):
                return time_config(solver, bz, by, flags, reps)
        results = []
        for bz, by, nt, bpc, flat in cands:
            flags = flags_of(nt, bpc, 1, flat)
            for f in ((flags, flags | _lib.PSC_RUN_INPLACE) if fused else (flags,)):
                results.append((timer(bz, by, f), bz, by, f))
        leaders = [(bz, by, f) for _, bz, by, f in sorted(results)[:LEADERS]]
        if fused:
            # each leader with every placement period, through long graph-replayed runs
            best_t = {}
            for _ in range(2):
                for bz, by, f in leaders:
                    for period in PERIODS:
                        ms = time_fused_config(solver, scratch, bz, by, f, LONG_STEPS, period, graph=True)
                        k = (bz, by, f, period)
                        best_t[k] = min(best_t.get(k, ms), ms)
            (bz, by, f, period), ms = min(best_t.items(), key=lambda kv: kv[1])
            cfg = (bz, by, f)
            pkey = _key(solver) + "|period|%d,%d,%d" % cfg
            _MEM[pkey] = period
            _store_file(pkey, [period, ms])
        else:
            # the leaders are often within 1% of each other: time them again, longer
            reps = max(4 * reps, 2 * GRAPH_STEPS + 2)
            ms, bz, by, f = min((time_config(solver, bz, by, f, reps), bz, by, f) for bz, by, f in leaders)
            cfg = (bz, by, f)
        _MEM[key] = cfg
        _store_file(key, list(cfg) + [ms])
        return cfg
