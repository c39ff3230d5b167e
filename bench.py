#!/usr/bin/env python3
"""Benchmark of the D3Q19 PSC-IMB step on one MI355X.

Workload: the fixed-array permeability case -- periodic 128^3 lattice, 60
non-overlapping spheres of radius 12 placed by seeded random sequential
addition, tau 0.8, body force 1e-6 along x, 4x4x4 sub-samples, static
geometry.  A "step" is one collide_stream_psc: the fused collide+stream
kernel plus the per-particle force/torque reduction, exactly what the
reference performs per call.

The timed path is the FP32 deviation storage (g = f - w).  The same
trajectory is run with the FP64 plain kernel (reference operation order) to
state the permeability error of FP32 against FP64 after the same number of
steps, and to time the FP64 step.  The CPU baseline is the NumPy FP64
oracle (a restatement of the reference) on the same lattice, a bounded number
of steps on one host core.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG = "particuology_hydraulic-penalty-_amd"
METRIC = "D3Q19 PSC MLUPS + % HBM roofline at 128³; FP32 vs FP64 at matched k error"
HBM_PEAK_GBPS = 8000.0
BYTES_F32_STATIC = 19 * 4 * 2 + 4   # g read + g written + solid count
BYTES_F64_STATIC = 19 * 8 * 2 + 4


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4900)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--kernel-repeats", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--cpu-n", type=int, default=64)
    ap.add_argument("--cpu-big-steps", type=int, default=8,
                    help="timed oracle steps of the benchmarked lattice itself (the reported cpu_baseline)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="processes of the aggregate CPU baseline (capped at os.cpu_count(); 1: none)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--layout", default="pull", choices=["pull", "push"])
    ap.add_argument("--launch", default=None,
                    help="fixed psc_step launch 'block_z,block_y,flags' instead of autotuning (for profiling); "
                         "flags may add PSC_RUN_INPLACE (0x200000) for in-place step pairs")
    ap.add_argument("--batch-launch", default=None,
                    help="windows workload: fixed launch 'flags,grid,mode' (mode spans|rows) instead of autotuning")
    ap.add_argument("--fuse-period", type=int, default=None,
                    help="placement period of the fused load workgroups (default: tuned; with --launch: 5)")
    ap.add_argument("--no-copy", action="store_true")
    ap.add_argument("--per-call-steps", type=int, default=240,
                    help="fixed_array: also time this many collide_stream_psc() calls one at a time, reading the "
                         "diag dict every 60 (0: skip)")
    ap.add_argument("--no-large-check", action="store_true",
                    help="skip the 256^3 HBM-only rate of the plain step kernel")
    ap.add_argument("--workload", default="fixed_array",
                    choices=["fixed_array", "moving", "windows", "adjoint", "bounce_back", "coupling", "moving_sphere"],
                    help="fixed_array: 128^3 permeability (default); moving: 512^3 window, 2000 moving spheres; "
                         "windows: 64 batched pore windows; adjoint: 96^3 permeability gradient; "
                         "bounce_back: 256^3 plain bounce-back window; coupling: one exchange's flow solve; "
                         "moving_sphere: the 20^3 FP64 moving-sphere case")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes that measure roofline.traffic")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(msg, quiet=False):
    if not quiet:
        print(msg, file=sys.stderr, flush=True)


def lib_flag(name):
    return getattr(importlib.import_module(PKG + "._lib"), name)


def timed_run(psc3d, table, n, precision, layout, steps, warmup, torch, launch=None, period=None):
    kw = {}
    if launch is not None:
        bz, by, flags = launch
        kw = dict(block=(bz, by))
    sol = psc3d.PscD3Q19(n, table, tau=0.8, subsamples=4, body_force=(1.0e-6, 0.0, 0.0), static_geometry=True,
                         precision=precision, stream_layout=layout, device="cuda:0", **kw)
    if launch is not None:
        # a fixed launch: flags as psc_step takes them, plus PSC_RUN_INPLACE for in-place step pairs
        sol._flags = launch[2] & ~lib_flag("PSC_RUN_INPLACE")
        sol.inplace = bool(launch[2] & lib_flag("PSC_RUN_INPLACE"))
    if period is not None:
        sol.fuse_period = period
    sol.rebuild_geometry()
    sol._tune()   # launch-shape autotuning happens before the warm-up
    sol.run(warmup)
    # capture (and upload) the graphs the timed run replays, without stepping, after the
    # warm-up: an odd warm-up leaves the state in the other buffer, and graphs are captured
    # for the buffers the run starts from, so graphs prepared before it would not be the
    # timed run's, which would then capture its own inside the timed region
    sol.prepare_run(steps)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    sol.run(steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    # the permeability readout (FP64 flux reduction of the final state), timed on its own
    r0 = time.perf_counter()
    k = sol.permeability(axis=0)
    readout_ms = (time.perf_counter() - r0) * 1e3
    return sol, k, wall, gpu_ms, readout_ms


def per_call_rate(sol, calls, read_every, torch):
    """collide_stream_psc() one call at a time, as the reference's drivers step
    (run_psc_permeability_resolution.py:137-143 reads the diag dict every 60
    calls): HIP-event time per step on the solver's stream over `calls` calls
    (a multiple of read_every, so the last call's read leaves nothing held or
    owed), and the wall time per call."""
    stream = torch.cuda.current_stream()
    for _ in range(4):
        sol.collide_stream_psc()
    sol.hydro_forces
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(calls):
        d = sol.collide_stream_psc()
        if (k + 1) % read_every == 0:
            d["mean_uz"]
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) * 1e3 / calls, wall * 1e6 / calls


def step_kernel_ms(sol, repeats, torch, lib, moving=False, launch_cfg=None):
    """Average duration of the collide+stream kernel alone, HIP events on the
    stream it is launched on; launch_cfg = (block_z, block_y, flags), default
    the solver's own launch."""
    nx, ny, nz = sol._shape
    fx, fy, fz = sol.body_force
    stream = torch.cuda.current_stream()
    h = stream.cuda_stream
    bz, by, flags = launch_cfg if launch_cfg is not None else (sol._block[0], sol._block[1], sol._flags)
    flags |= getattr(sol, "_state_flags", 0)  # the state's plane layout

    owners = moving and getattr(sol, "_geom_mode", None) == "owners"

    def launch(a, b):
        if owners:  # the kernel the moving solver runs: solid velocity from the owner table
            lib.call("psc_step_owners", a.data_ptr(), b.data_ptr(), sol._csum.data_ptr(), sol._own.data_ptr(),
                     sol._us.data_ptr(), sol._motion.data_ptr(), int(sol._tables.particles.shape[0]),
                     sol._imb.data_ptr(), nx, ny, nz, sol._s, sol.tau, fx, fy, fz, bz, by, flags, h)
            return
        lib.call("psc_step", a.data_ptr(), b.data_ptr(), sol._csum.data_ptr(),
                 sol._us.data_ptr() if moving else None, sol._imb.data_ptr(),
                 sol._storage, sol._layout, nx, ny, nz, sol._s, sol.tau, fx, fy, fz, bz, by, flags, h)

    a, b = sol._a, sol._b
    for _ in range(20):
        launch(a, b)
        a, b = b, a
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(repeats):
        launch(a, b)
        a, b = b, a
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / repeats


def hbm_only_check(psc3d, workloads, torch, lib, repeats, quiet, n=256):
    """The plain FP32 step kernel (tuned launch) on an n^3 lattice of the same
    workload (60 spheres per 128^3, the same volume fraction), whose 2.5 GB
    two-buffer working set is ten times the 256 MB Infinity Cache: its rate is
    an HBM rate."""
    count = int(round(60 * (n / 128) ** 3))
    sol = psc3d.PscD3Q19(n, workloads.fixed_array(seed=0, n=n, count=count), tau=0.8, subsamples=4,
                         body_force=(1.0e-6, 0.0, 0.0), static_geometry=True, precision="fp32", device="cuda:0")
    sol.rebuild_geometry()
    cfg = importlib.import_module(PKG + ".autotune").tune(sol)
    ms = step_kernel_ms(sol, max(20, repeats), torch, lib, launch_cfg=cfg)
    ibytes, _ = imb_bytes(sol._csum)
    nbytes = BYTES_F32_STATIC * n ** 3 + ibytes
    gbps = nbytes / (ms * 1e-3) / 1e9
    log(f"hbm-only check: {n}^3 step kernel {ms*1e3:.1f} us = {gbps:.0f} GB/s", quiet)
    out = {"lattice": [n, n, n], "particles": count, "step_kernel_us": ms * 1e3, "bytes_per_launch": nbytes,
           "achieved_GBps": gbps, "frac": gbps / HBM_PEAK_GBPS, "launch": list(cfg),
           "note": "plain two-buffer step kernel on a 2.5 GB working set: an HBM-only rate"}
    del sol
    torch.cuda.empty_cache()
    return out


def copy_bandwidth(torch, lib, nbytes=2 << 30, repeats=20):
    n = nbytes // 4
    src = torch.empty(n, dtype=torch.float32, device="cuda:0").fill_(1.0)
    dst = torch.empty_like(src)
    h = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        lib.call("psc_copy_f32", src.data_ptr(), dst.data_ptr(), n, h)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(repeats):
        lib.call("psc_copy_f32", src.data_ptr(), dst.data_ptr(), n, h)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / repeats
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def _oracle_solver(table, n):
    from oracle import psc_oracle as O

    parts = [O.Sphere(*row) for row in table]
    return O.OracleSolver(n, parts, tau=0.8, subsamples=4, body_force=(1.0e-6, 0.0, 0.0), static_geometry=True,
                          layout="sparse")


def cpu_baseline(table, n, steps, quiet):
    """NumPy FP64 oracle (restatement of the reference) timed on one core
    for ``steps`` static-geometry steps of an n^3 lattice.  Returns (MLUPS,
    seconds, voxelization seconds)."""
    s = _oracle_solver(table, n)
    t0 = time.perf_counter()
    s.rebuild_geometry()
    tv = time.perf_counter() - t0
    s.collide_stream_psc()   # first step outside the timing (allocations settle)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.collide_stream_psc()
    dt = time.perf_counter() - t0
    log(f"cpu oracle: {steps} steps of {n}^3 in {dt:.2f}s (voxelization {tv:.2f}s)", quiet)
    return n ** 3 * steps / dt / 1e6, dt, tv


def _cpu_worker(table, n, steps, barrier, results):
    """One process of the aggregate CPU baseline: an independent n^3 box."""
    s = _oracle_solver(table, n)
    s.rebuild_geometry()
    s.collide_stream_psc()
    barrier.wait()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.collide_stream_psc()
    results.put(time.perf_counter() - t0)


def cpu_baseline_aggregate(table, n, steps, procs, quiet):
    """The reference's case-level parallelism: ``procs`` single-threaded
    processes, each stepping its own n^3 box (fresh interpreters), started
    together.  Aggregate MLUPS = procs * n^3 * steps / slowest process."""
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    barrier, results = ctx.Barrier(procs), ctx.Queue()
    ps = [ctx.Process(target=_cpu_worker, args=(table, n, steps, barrier, results)) for _ in range(procs)]
    for p in ps:
        p.start()
    dts = [results.get(timeout=600) for _ in ps]
    for p in ps:
        p.join()
    dt = max(dts)
    log(f"cpu oracle aggregate: {procs} processes x {steps} steps of {n}^3, slowest {dt:.2f}s", quiet)
    return procs * n ** 3 * steps / dt / 1e6, dt


def imb_bytes(csum):
    """Algorithmic bytes of the imb output: 12 per cell with solid (count_sum > 0)."""
    n = int((csum != 0).sum().item())
    return 12 * n, n


def loads_bytes(sol):
    """Algorithmic bytes of one step's load reduction: per compacted entry
    4 (cell) + 4 (box coordinates) + 2 (count) + 4 (count_sum) + 12 (imb),
    48 per chunk written and read back, 48 per particle written."""
    C = int(sol._tables.chunks.shape[0])
    nnz = int(sol._chunk_nnz[:C].sum().item())
    P = int(sol._tables.particles.shape[0])
    return 26 * nnz + 96 * C + 48 * P, nnz


def read_traffic(name, launch):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary ``profiles/<name>`` (FETCH_SIZE + WRITE_SIZE, calibrated on the
    copy kernel), or None when the summary is missing or was measured on
    another launch configuration than ``launch`` (a tuple of ints that the
    summary's ``launch:`` line must repeat)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    traffic, measured = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            traffic = float(line.split(":", 1)[1])
        elif line.startswith("launch:"):
            measured = tuple(int(v, 0) for v in line.split(":", 1)[1].split(","))
    return traffic if measured == tuple(int(v) for v in launch) else None


COPY_PROBE_BYTES = 256 << 20  # each buffer of the calibration copy in the traffic probe


def _copy_calibration(lib, torch):
    """The streaming copy kernel on 256 MiB buffers (read 256 MiB, write
    256 MiB per dispatch), three times: the counter calibration of a traffic
    probe."""
    m = COPY_PROBE_BYTES // 4
    src = torch.ones(m, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        lib.call("psc_copy_f32", src.data_ptr(), dst.data_ptr(), m, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()


def traffic_probe(args, psc3d, workloads, lib, torch):
    """Child of measure_traffic, run under rocprofv3 --pmc: the timed
    configuration with the launch fixed (--launch, --fuse-period) for
    12 direct steps, then the copy calibration.  With --workload windows: the
    batched windows with --batch-launch and --fuse-period for 12 steps of
    fused launches (no graph), then the copy calibration."""
    if args.workload == "windows":
        batch = importlib.import_module(PKG + ".batch")
        f_, g_, m_ = args.batch_launch.split(",")
        wins = workloads.pebble_windows(seed=0)
        B = batch.BatchedPscD3Q19([(L, t) for L, t, _ in wins], tau=0.8, subsamples=3, body_force=(0.0, 0.0, 1.0e-6),
                                  device="cuda:0", launch=(int(f_, 0), int(g_), m_))
        if args.fuse_period is not None:
            B.fuse_period = args.fuse_period
        B.run(12, graph=False)
        torch.cuda.synchronize()
        _copy_calibration(lib, torch)
        return
    n = args.n
    bz, by, flags = (int(v) for v in args.launch.split(","))
    sol = psc3d.PscD3Q19(n, workloads.fixed_array(seed=0, n=n), tau=0.8, subsamples=4, body_force=(1.0e-6, 0.0, 0.0),
                         static_geometry=True, precision="fp32", stream_layout=args.layout, device="cuda:0",
                         block=(bz, by))
    sol._flags = flags & ~lib.PSC_RUN_INPLACE
    sol.inplace = bool(flags & lib.PSC_RUN_INPLACE)
    sol.fuse_period = args.fuse_period or 1
    sol.rebuild_geometry()
    sol.run(12, graph=False)
    torch.cuda.synchronize()
    _copy_calibration(lib, torch)


def measure_traffic(args, launch, period, kernel_prefix, quiet, probe_args=None, timeout=180):
    """HBM traffic per launch of the timed step kernel: two rocprofv3 --pmc
    passes over ``bench.py --traffic-probe`` with the same launch (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter
    reads for the copy kernel of known size in that run.  probe_args: the
    probe's arguments when they are not the default workload's.  Returns a
    dict, or None when rocprofv3 is missing or a pass fails."""
    import csv
    import shutil
    import statistics
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="psc_traffic_", dir="/tmp")
        if probe_args is None:
            probe_args = ["--n", str(args.n), "--layout", args.layout, "--launch",
                          ",".join(str(int(v)) for v in launch), "--fuse-period", str(int(period))]
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "p", "--", sys.executable, "-u",
               os.path.join(ROOT, "bench.py"), "--traffic-probe"] + list(probe_args)
        log(f"traffic: rocprofv3 --pmc {counter} over the probe", quiet)
        try:
            subprocess.run(cmd, cwd=ROOT, env=env, timeout=timeout, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
        except (subprocess.TimeoutExpired, OSError):
            shutil.rmtree(d, ignore_errors=True)
            return None
        files = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        step, copy = [], []
        for f in files:
            for row in csv.DictReader(open(f)):
                if row.get("Counter_Name") != counter:
                    continue
                name = row["Kernel_Name"]
                if kernel_prefix in name:
                    step.append(float(row["Counter_Value"]))
                elif "k_copy4" in name:
                    copy.append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if len(step) < 8 or not copy:
            return None
        scale = COPY_PROBE_BYTES / (statistics.median(copy) * 1024.0)
        kib = statistics.median(step[4:])  # steady state: after the two head steps of each form
        out[counter] = {"kib_per_dispatch": kib, "copy_scale": scale, "bytes": kib * 1024.0 * scale}
    out["bytes_per_launch"] = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    return out


def bench_moving(args, psc3d, workloads, lib, torch, np):
    """Large moving-particle window: periodic n^3 (512 by default), 2000
    spheres r ~ U[6,18], |U|, |omega| r <= 1e-3, centres advanced by U every
    step and re-voxelized every step (s=3), tau 0.8, no body force.  FP32
    deviation storage; the same trajectory in FP64 gives the per-particle
    force/torque error."""
    n = args.n if args.n != 128 else 512
    steps = args.steps if args.steps != 4900 else 1000
    warm = min(args.warmup, 20)
    N = n ** 3
    base = workloads.moving_window(seed=0, n=n)

    def run(precision, nsteps, timed):
        table = base.copy()
        sol = psc3d.SparsePscD3Q19(n, table, tau=0.8, subsamples=3, static_geometry=False, precision=precision,
                                   periodic_geometry=True, device="cuda:0")
        for _ in range(warm):
            workloads.advance_centres(table, n)
            sol.particles = table
            sol.run(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            workloads.advance_centres(table, n)
            sol.particles = table
            sol.run(1)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return sol, wall

    log(f"moving window {n}^3: {warm} warm-up + {steps} steps, FP32", args.quiet)
    s32, w32 = run("fp32", steps, True)
    kern = step_kernel_ms(s32, 20, torch, lib, moving=True)
    solid = s32._csum != 0
    nsolid = int(solid.sum().item())   # cells that read their solid velocity and write imb
    owners = s32._geom_mode == "owners"
    # owner table: 4 B of owner id at every cell with solid, 12 B of numerators at the shared ones
    nshared = int(solid.logical_and(s32._own == -1).sum().item()) if owners else nsolid
    launch_cfg = [int(s32._block[0]), int(s32._block[1]), int(s32._flags)]
    loads32 = s32._loads_np().copy()
    mlups = N * steps / w32 / 1e6
    log(f"fp32: {mlups:.0f} MLUPS, step kernel {kern*1e3:.0f} us", args.quiet)
    del s32
    torch.cuda.empty_cache()
    s64, w64 = run("fp64", steps, True)
    loads64 = s64._loads_np().copy()
    del s64
    torch.cuda.empty_cache()
    Fs = np.linalg.norm(loads64[:, 0:3], axis=1)
    Ts = np.linalg.norm(loads64[:, 3:6], axis=1)
    eF = np.linalg.norm(loads32[:, 0:3] - loads64[:, 0:3], axis=1) / np.maximum(Fs, 1e-3 * Fs.max())
    eT = np.linalg.norm(loads32[:, 3:6] - loads64[:, 3:6], axis=1) / np.maximum(Ts, 1e-3 * Ts.max())
    # algorithmic bytes: 156 per cell plus, at the cells with solid, 12 of imb written and the
    # solid velocity: 12 of numerators (plane form) or 4 of owner id + 12 of numerators at the
    # shared cells (owner table; the motion records, 96 B per particle, are cache-resident)
    launch_bytes = BYTES_F32_STATIC * N + (16 * nsolid + 12 * nshared if owners else 24 * nsolid)
    achieved = launch_bytes / (kern * 1e-3) / 1e9
    out = {
        "metric": METRIC + " (moving-window configuration)",
        "value": mlups, "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": w32 * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"moving window: periodic {n}^3, 2000 spheres r~U[6,18] (seed 0), |U|,|w|r<=1e-3, "
                               "centres advanced and re-voxelized every step (s=3), tau 0.8, FP32 deviation "
                               "storage, FP64 loads; step = voxelize + collide+stream + loads",
                   "lattice": [n, n, n], "particles": int(base.shape[0])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "k_step_f32 (pull, moving, %s)" % ("owner table" if owners else "solid-velocity planes"),
                     "bytes_per_cell": BYTES_F32_STATIC, "solid_cells": nsolid, "shared_cells": nshared,
                     "bytes_per_launch": launch_bytes, "kernel_us": kern * 1e3, "launch": launch_cfg},
        "fp64_MLUPS": N * steps / w64 / 1e6,
        "force_rel_err_fp32_vs_fp64": {"max": float(eF.max()), "median": float(np.median(eF))},
        "torque_rel_err_fp32_vs_fp64": {"max": float(eT.max()), "median": float(np.median(eT))},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        from oracle import psc_oracle as O
        cn = 64
        cnt = max(1, int(round(base.shape[0] * (cn / n) ** 3)))
        ct_ = workloads.moving_window(seed=0, n=cn, count=cnt)
        ora = O.OracleSolver(cn, [O.Sphere(*row) for row in ct_], tau=0.8, subsamples=3, static_geometry=False,
                             layout="sparse", periodic_geometry=True)
        ora.collide_stream_psc()
        cs = 4
        tc = time.perf_counter()
        for _ in range(cs):
            workloads.advance_centres(ct_, cn)
            ora.particles = [O.Sphere(*row) for row in ct_]
            ora.collide_stream_psc()
        dt = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": cn ** 3 * cs / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
                               "sample": f"NumPy FP64 oracle, {cs} moving steps (re-voxelized each step) of a {cn}^3 "
                                         f"periodic window with {cnt} spheres (same density) in {dt:.1f} s, one thread"}
    print(json.dumps(out))


def bench_windows(args, workloads, torch, np):
    """Batched pore windows: 64 periodic windows, edges ~ U{48..160}, pebbles
    r=16 near solid fraction 0.6 plus 0-400 throat fines (skeleton / passive /
    preloaded), tau 0.8, F = 1e-6 along z, s=3, FP32, 3000 steps; one
    persistent launch per step (+ loads), replayed through a hipGraph."""
    batch = importlib.import_module(PKG + ".batch")
    steps = args.steps if args.steps != 4900 else 3000
    warm = min(args.warmup, 20)
    log("generating 64 windows", args.quiet)
    wins = workloads.pebble_windows(seed=0)
    launch = "auto"
    if args.batch_launch:
        f_, g_, m_ = args.batch_launch.split(",")
        launch = (int(f_, 0), int(g_), m_)
    B = batch.BatchedPscD3Q19([(L, t) for L, t, _ in wins], tau=0.8, subsamples=3, body_force=(0.0, 0.0, 1.0e-6),
                              device="cuda:0", launch=launch)
    if args.fuse_period is not None:
        B.fuse_period = args.fuse_period
    B.run(warm)
    stream = torch.cuda.current_stream()
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g0.record(stream)
    B.run(steps)
    g1.record(stream)
    k = B.permeability()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = g0.elapsed_time(g1)
    fused = bool(B.fuse_loads and B.P > 0)
    # batched step kernel alone, HIP events on its stream
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        B._launch_step(B._a, B._b, B._flags, B._grid, stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    kern = e0.elapsed_time(e1) / reps
    cells = int(B.nlive.sum())  # live cells only: row and plane padding is neither read nor written
    nnz = int(B._nnz[:B.C].sum().item())
    lbytes = 26 * nnz + 96 * B.C + 48 * B.P
    ibytes, nsolid = imb_bytes(B._csum)
    step_bytes = BYTES_F32_STATIC * cells + ibytes
    if fused:
        # every timed step is one k_step_f32_batched<..., loads> launch: the HIP-event time of
        # the timed region over the step count is its average duration (plus the pending
        # loads once at the end of the run)
        kern_l = gpu_ms / steps
        launch_bytes = step_bytes + lbytes
        achieved = launch_bytes / (kern_l * 1e-3) / 1e9
        kname, kus = "k_step_f32_batched (with the loads of the two previous steps)", kern_l * 1e3
    else:
        launch_bytes = step_bytes
        achieved = launch_bytes / (kern * 1e-3) / 1e9
        kname, kus = "k_step_f32_batched", kern * 1e3
    por = B.porosity()
    variants = [v for _, _, v in wins]
    out = {
        "metric": METRIC + " (batched-window configuration)",
        "value": cells * steps / wall / 1e6, "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 periodic pore windows, L~U{48..160} (seed 0), pebbles r=16 near solid fraction "
                               "0.6 + 0-400 throat fines (skeleton/passive/preloaded), tau 0.8, F=1e-6 z, s=3, "
                               "FP32; one persistent launch per step (loads "
                               f"{'fused into the launches' if fused else 'on a side stream'}), hipGraph replay",
                   "windows": len(wins), "total_cells": cells, "particles": int(B.P)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": None,
                     "kernel": kname,
                     "bytes_per_cell": BYTES_F32_STATIC, "bytes_per_launch": launch_bytes,
                     "solid_cells": nsolid, "imb_bytes": ibytes,
                     "loads_bytes_per_step": lbytes, "loads_entries": nnz, "kernel_us": kus,
                     "step_kernel_alone_us": kern * 1e3,
                     "step_kernel_alone_GBps": step_bytes / (kern * 1e-3) / 1e9,
                     "launch": {"flags": B._flags, "grid": B._grid, "items": B._mode,
                                "workgroup": 512 if B._flags & 0x40000 else 256,
                                "fuse_period": B.effective_fuse_period() if fused else None}},
        "permeability": {v: [float(x) for x, vv in zip(k, variants) if vv == v][:4] for v in ("skeleton", "passive",
                                                                                            "preloaded")},
        "porosity_range": [float(por.min()), float(por.max())],
        "cpu_baseline": None,
    }
    # FP32 batch vs FP64 single-window solvers on the three smallest windows
    psc3d = importlib.import_module(PKG + ".psc3d")
    small = list(np.argsort([L for L, _, _ in wins])[:3])
    errs = []
    for w in small:
        L, t, _ = wins[w]
        s64 = psc3d.SparsePscD3Q19(L, t, tau=0.8, subsamples=3, body_force=(0.0, 0.0, 1.0e-6), static_geometry=True,
                                   periodic_geometry=True, precision="fp64", device="cuda:0")
        s64.run(warm + steps)
        k64 = s64.permeability(axis=2)
        errs.append({"window": int(w), "L": int(L), "k_fp32": float(k[w]), "k_fp64": k64,
                     "rel_err": abs(float(k[w]) - k64) / abs(k64)})
        del s64
    out["k_fp32_vs_fp64_smallest_windows"] = errs
    if fused and not args.no_traffic:
        # HBM/fabric bytes per fused launch, measured on this launch by two counter passes
        probe = ["--workload", "windows", "--batch-launch", f"{B._flags:#x},{B._grid},{B._mode}",
                 "--fuse-period", str(B.effective_fuse_period())]
        tr = measure_traffic(args, None, None, "k_step_f32_batched<", args.quiet, probe_args=probe, timeout=400)
        if tr is not None:
            tr["source"] = "measured in this run: rocprofv3 --pmc over bench.py --traffic-probe, same launch"
            out["roofline"]["traffic"] = tr["bytes_per_launch"]
            out["roofline"]["traffic_over_algorithmic"] = tr["bytes_per_launch"] / launch_bytes
            out["roofline"]["traffic_detail"] = tr
    if not args.no_cpu_baseline:
        from oracle import psc_oracle as O
        w0 = int(np.argmin([L for L, _, _ in wins]))
        L0, t0_, _ = wins[w0]
        ora = O.OracleSolver(L0, [O.Sphere(*row) for row in t0_], tau=0.8, subsamples=3, body_force=(0.0, 0.0, 1.0e-6),
                             static_geometry=True, layout="sparse", periodic_geometry=True)
        ora.collide_stream_psc()
        cs = 8
        tc = time.perf_counter()
        for _ in range(cs):
            ora.collide_stream_psc()
        dt = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": L0 ** 3 * cs / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
                               "sample": f"NumPy FP64 oracle, {cs} steps of window {w0} ({L0}^3, {len(t0_)} particles, "
                                         f"periodic) in {dt:.1f} s, one thread"}
    print(json.dumps(out))


def bench_adjoint(args, workloads, torch, np):
    """Permeability gradient: periodic 96^3 window, 40 pebbles r=14 + 300
    fines r in [2, 4] (lognormal), smooth solid fraction, tau 0.8, F = 1e-6
    along z, T = 2000 steps from rest, sqrt(T) checkpointing.  Times one FP32
    gradient evaluation (forward + reverse sweep) after a warm-up one, the
    reverse-step kernel alone, and checks the FP32 gradient against the FP64
    adjoint and FP64 central differences (h = 1e-3 cells) on the four largest
    components."""
    adj = importlib.import_module(PKG + ".adjoint")
    lib = importlib.import_module(PKG + "._lib")
    T = args.steps if args.steps != 4900 else 2000
    n = 96
    F = (0.0, 0.0, 1.0e-6)
    tab = workloads.adjoint_window(seed=0)
    N = n ** 3
    p32 = adj.PermeabilityAdjoint(n, tau=0.8, body_force=F, steps=T, precision="fp32", device="cuda:0")
    log(f"adjoint: warm-up gradient ({T} steps, K={p32.K})", args.quiet)
    p32.value_and_grad(tab)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k32 = p32.forward(tab)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g32 = p32.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = t1 - t0, t2 - t1
    log(f"fp32: k={k32:.9e} forward {fwd*1e3:.1f} ms, reverse sweep {bwd*1e3:.1f} ms", args.quiet)
    alpha = p32.geometry.alpha32
    trans = float(((alpha > 0) & (alpha < 1)).float().mean())
    # kernels alone, HIP events on the stream they run on
    stream = torch.cuda.current_stream()
    h = stream.cuda_stream
    reps = 100

    def timed(fn):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    adj_ms = timed(lambda: p32._reverse(p32._seg[0], p32._lam[0], p32._lam[1], h))
    fwd_ms = timed(lambda: p32._step(p32._seg[0], p32._seg[1], h))
    bytes_adj = 3 * 76 + 4 + 16 * trans   # state gather, adjoint in/out, alpha, dk/dalpha update
    ach = bytes_adj * N / (adj_ms * 1e-3) / 1e9
    log(f"k_adjoint_step {adj_ms*1e3:.1f} us ({ach:.0f} GB/s), field step {fwd_ms*1e3:.1f} us", args.quiet)

    log("fp64 reference gradient", args.quiet)
    p64 = adj.PermeabilityAdjoint(n, tau=0.8, body_force=F, steps=T, precision="fp64", device="cuda:0")
    t3 = time.perf_counter()
    k64, g64 = p64.value_and_grad(tab)
    t4 = time.perf_counter()
    scale = float(np.max(np.abs(g64)))
    fd_rows = []
    for flat in np.argsort(-np.abs(g64).ravel())[:4]:
        pi, j = divmod(int(flat), 4)
        hh = 1.0e-3
        tp, tm = tab.copy(), tab.copy()
        tp[pi, j] += hh
        tm[pi, j] -= hh
        fd = (p64.forward(tp) - p64.forward(tm)) / (2 * hh)
        fd_rows.append({"particle": pi, "param": "xyzr"[j], "fd_fp64": fd, "adjoint_fp64": float(g64[pi, j]),
                        "adjoint_fp32": float(g32[pi, j]),
                        "rel_err_fp32_vs_fd": abs(float(g32[pi, j]) - fd) / abs(fd),
                        "rel_err_fp64_vs_fd": abs(float(g64[pi, j]) - fd) / abs(fd)})
    out = {
        "metric": METRIC + " (adjoint-gradient configuration)",
        "value": N * T / (fwd + bwd) / 1e6, "unit": "MLUPS", "n_gpus": 1, "steps": T, "warmup": 1,
        "ms_per_step": (fwd + bwd) * 1e3 / T, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "permeability gradient: periodic 96^3 window, 40 pebbles r=14 + 300 passive throat "
                               "fines r~lognormal(3, 0.25) clipped to [2,4] (seed 0), smooth solid fraction, tau 0.8, "
                               f"F=1e-6 z, {T} steps from rest, sqrt(T) checkpoints (K={p32.K}), FP32 forward and "
                               "adjoint, FP64 reductions; value = N*T / (forward + reverse sweep) seconds",
                   "particles": int(tab.shape[0]), "lattice": [n, n, n],
                   "porosity": 1.0 - float(alpha.double().mean()), "transition_cell_fraction": trans},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None, "kernel": "k_adjoint_step (FP32)",
                     "bytes_per_cell": bytes_adj, "kernel_us": adj_ms * 1e3},
        "forward_ms": fwd * 1e3, "reverse_sweep_ms": bwd * 1e3,
        "forward_MLUPS": N * T / fwd / 1e6,
        "field_step_kernel_us": fwd_ms * 1e3,
        "k_fp32": k32, "k_fp64": k64, "k_rel_err": abs(k32 - k64) / abs(k64),
        "grad_rel_err_fp32_vs_fp64": float(np.max(np.abs(g32 - g64)) / scale),
        "fp64_gradient_s": t4 - t3,
        "finite_differences": fd_rows,
        "checkpoint_memory_GB": p32.memory_bytes() / 1e9,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def bench_bounce_back(args, workloads, torch, np):
    """Plain bounce-back window solver (run_lbm of the drivers): periodic
    256^3 voxel window of 400 RSA spheres r = 10 (centre-in-sphere voxels),
    tau 0.72, force 2e-7 along z (the drivers' defaults), FP32 timed over the
    steps after a warm-up, FP64 run for the k error and its own MLUPS."""
    lbm = importlib.import_module(PKG + ".lbm")
    n = 256
    steps = args.steps if args.steps != 4900 else 1000
    warm = min(args.warmup, 50)
    pts = workloads.rsa_spheres(n, 400, 10.0, seed=0)
    solid, _ = lbm.voxelize(pts, np.full(3, n / 2.0), np.full(3, float(n)), 10.0, n, device="cuda:0")
    N = n ** 3
    res = {}
    for prec in ("fp32", "fp64"):
        s = lbm.BounceBackD3Q19(solid, tau=0.72, force_z=2.0e-7, precision=prec, device="cuda:0")
        s.run(warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        stream = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        reps = 50
        e0.record(stream)
        for _ in range(reps):
            s._launch(s._a, s._b, stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        kern = e0.elapsed_time(e1) / reps
        out = s.result()
        res[prec] = {"MLUPS": N * steps / wall / 1e6, "kernel_us": kern * 1e3, "k": out["permeability_lu"],
                     "bytes_per_cell": (19 * 4 * 2 + 1) if prec == "fp32" else (19 * 8 * 2 + 1),
                     "launch": [s.block[0], s.block[1], s.flags]}
        log(f"{prec}: {res[prec]['MLUPS']:.0f} MLUPS, kernel {kern*1e3:.1f} us, k={out['permeability_lu']:.9e}",
            args.quiet)
        del s
    r32 = res["fp32"]
    ach = r32["bytes_per_cell"] * N / (r32["kernel_us"] * 1e-6) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import bounce_back as BB
        # the drivers' default resolution, spheres scaled to keep the solid fraction
        cn, cs = 56, 20
        cr = 10.0 * cn / n * 2.0
        ccount = max(1, int(round(float(solid.mean()) * cn ** 3 / (4.0 / 3.0 * np.pi * cr ** 3))))
        csolid = BB.voxelize(workloads.rsa_spheres(cn, ccount, cr, seed=0), np.full(3, cn / 2.0),
                             np.full(3, float(cn)), cr, cn)[0]
        t0 = time.perf_counter()
        BB.run_lbm(csolid, 0.72, 2.0e-7, cs)
        dt = time.perf_counter() - t0
        cpu = {"value": cn ** 3 * cs / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
               "sample": f"NumPy FP64 restatement of run_lbm (oracle/bounce_back.py), {cs} steps of a {cn}^3 window "
                         f"(the drivers' default resolution) of {ccount} spheres r={cr:.2f}, solid fraction "
                         f"{float(csolid.mean()):.3f}, in {dt:.1f} s, one thread"}
    out = {
        "metric": METRIC + " (plain bounce-back window configuration)",
        "value": r32["MLUPS"], "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": N / (r32["MLUPS"] * 1e6) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "run_lbm bounce-back window: periodic 256^3, 400 RSA spheres r=10 (seed 0), "
                               "centre-in-sphere voxels, tau 0.72, force 2e-7 z, FP32 deviation storage; "
                               "two steps per hipGraph replay",
                   "lattice": [n, n, n], "porosity": float(1.0 - solid.mean())},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None, "kernel": "k_bb_step_f32",
                     "bytes_per_cell": r32["bytes_per_cell"], "kernel_us": r32["kernel_us"],
                     "launch": r32["launch"]},
        "fp64": res["fp64"], "k_rel_err_fp32_vs_fp64": abs(r32["k"] - res["fp64"]["k"]) / abs(res["fp64"]["k"]),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bench_coupling(args, workloads, torch, np):
    """One coupling exchange's flow solve (the drivers' lbm_converge): a
    40^3 window (the coupling driver's default resolution) of 12 pebbles r=6
    and 60 fines r=1.2, tau 0.72, force 2e-7 along z, s=2, residual 1e-8 on
    mean u_z, at most 2000 steps.  Times converge with the residual read after
    every step (the reference's loop) and read every 32 steps (native graph
    loop between reads), FP64 and FP32."""
    psc3d = importlib.import_module(PKG + ".psc3d")
    cpl = importlib.import_module(PKG + ".coupling")
    n = 40
    rng = np.random.default_rng(0)
    peb = workloads.rsa_spheres(n, 12, 6.0, seed=0, margin=-6.0)
    fines = rng.uniform(0, n, (60, 3))
    table = np.zeros((72, 10))
    table[:12, 0:3], table[:12, 3] = peb, 6.0
    table[12:, 0:3], table[12:, 3] = fines, 1.2
    rows = {}
    for prec in ("fp64", "fp32"):
        for K in (1, 32):
            sol = psc3d.SparsePscD3Q19(n, [psc3d.Particle3D(*r) for r in table], tau=0.72, subsamples=2,
                                       body_force_z=2.0e-7, precision=prec, device="cuda:0")
            cpl.converge(sol, 20, 40, 1e-8, check_every=K)  # warm-up (kernel loading, graph capture)
            sol = psc3d.SparsePscD3Q19(n, [psc3d.Particle3D(*r) for r in table], tau=0.72, subsamples=2,
                                       body_force_z=2.0e-7, precision=prec, device="cuda:0")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, res, ok, diag = cpl.converge(sol, 20, 2000, 1e-8, check_every=K)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            rows[f"{prec}_K{K}"] = {"steps": it, "residual": res, "converged": ok, "seconds": wall,
                                    "MLUPS": n ** 3 * it / wall / 1e6, "us_per_step": wall / it * 1e6}
            log(f"{prec} K={K}: {it} steps in {wall*1e3:.1f} ms ({wall/it*1e6:.1f} us/step)", args.quiet)
    best = rows["fp32_K32"]
    out = {
        "metric": METRIC + " (coupling-exchange configuration)",
        "value": best["MLUPS"], "unit": "MLUPS", "n_gpus": 1, "steps": best["steps"], "warmup": 40,
        "ms_per_step": best["us_per_step"] / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "coupling exchange flow solve: 40^3 window, 12 pebbles r=6 + 60 fines r=1.2 (seed 0), "
                               "tau 0.72, Fz 2e-7, s=2, converge to 1e-8 relative change of mean u_z (<= 2000 "
                               "steps); value = FP32 with the residual read every 32 steps"},
        "roofline": None,
        "runs": rows,
        "speedup_K32_over_K1": {p: rows[f"{p}_K1"]["us_per_step"] / rows[f"{p}_K32"]["us_per_step"]
                                for p in ("fp64", "fp32")},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def bench_moving_sphere(args, psc3d, workloads, torch, np):
    """The plain-method FP64 case: periodic 20^3, tau 0.8, rho 1 at rest, one
    sphere of radius 2.5 at a seeded uniform centre moving at u = (1e-4, 0, 0),
    2x2x2 sub-samples, re-voxelized every step (dynamic geometry, as the
    moving-sphere harness runs it, run_d3q19_psc_moving_sphere_benchmark.py:
    72-80), 100 steps of collide_stream_psc called from Python with the drag
    read after each.  Checked against the oracle on the same steps: f and the
    force/torque history."""
    n = 20
    steps = args.steps if args.steps != 4900 else 100
    table = workloads.moving_sphere(seed=0)

    def make():
        return psc3d.PscD3Q19(n, [psc3d.Particle3D(*row) for row in table], tau=0.8, subsamples=2, precision="fp64",
                              device="cuda:0")

    warm = make()
    for _ in range(min(args.warmup, 20)):
        warm.collide_stream_psc()
        _ = warm.hydro_forces[0, 0]
    sol = make()
    torch.cuda.synchronize()
    forces = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sol.collide_stream_psc()
        forces.append(sol.hydro_forces[0].copy())  # the harness reads the drag every step
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    F, T = sol.hydro_forces[0].copy(), sol.hydro_torques[0].copy()
    from oracle import psc_oracle as O
    ora = O.OracleSolver(n, [O.Sphere(*row) for row in table], tau=0.8, subsamples=2, layout="dense")
    tc = time.perf_counter()
    ref_forces = []
    for _ in range(steps):
        ora.collide_stream_psc()
        ref_forces.append(ora.hydro_forces[0].copy())
    dt = time.perf_counter() - tc
    ferr = float(np.max(np.abs(np.array(forces) - np.array(ref_forces))) / np.max(np.abs(ref_forces)))
    out = {
        "metric": METRIC + " (moving-sphere configuration)",
        "value": n ** 3 * steps / wall / 1e6, "unit": "MLUPS", "n_gpus": 1, "steps": steps,
        "warmup": min(args.warmup, 20), "ms_per_step": wall * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "moving sphere: periodic 20^3, one r=2.5 sphere at a seeded uniform centre (seed 0), "
                               "u=(1e-4,0,0), tau 0.8, s=2, re-voxelized every step, FP64 reference operation order; "
                               "step = voxelize + collide+stream + loads, called from Python, drag read every step",
                   "lattice": [n, n, n], "particles": 1},
        "roofline": None,
        "force": F.tolist(), "torque": T.tolist(),
        "f_bit_identical_to_oracle": bool(np.array_equal(sol.f, ora.f)),
        "force_history_rel_err_vs_oracle": ferr,
        "cpu_baseline": {"value": n ** 3 * steps / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
                         "sample": f"NumPy FP64 oracle, the same {steps} steps in {dt:.2f} s, one thread"},
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(--gpus must be 1)", "n_gpus": args.gpus}))
        sys.exit(2)
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import numpy as np
    import torch

    psc3d = importlib.import_module(PKG + ".psc3d")
    workloads = importlib.import_module(PKG + ".workloads")
    lib = importlib.import_module(PKG + "._lib")
    lib.load()
    if args.traffic_probe:
        return traffic_probe(args, psc3d, workloads, lib, torch)
    if args.workload == "moving":
        return bench_moving(args, psc3d, workloads, lib, torch, np)
    if args.workload == "windows":
        return bench_windows(args, workloads, torch, np)
    if args.workload == "adjoint":
        return bench_adjoint(args, workloads, torch, np)
    if args.workload == "bounce_back":
        return bench_bounce_back(args, workloads, torch, np)
    if args.workload == "coupling":
        return bench_coupling(args, workloads, torch, np)
    if args.workload == "moving_sphere":
        return bench_moving_sphere(args, psc3d, workloads, torch, np)
    n = args.n
    table = workloads.fixed_array(seed=0, n=n)
    total = args.warmup + args.steps
    N = n ** 3

    log(f"fp32 run: {args.warmup} warmup + {args.steps} timed steps at {n}^3", args.quiet)
    launch = tuple(int(v) for v in args.launch.split(",")) if args.launch else None
    sol32, k32, wall32, gpu_ms32, readout_ms = timed_run(psc3d, table, n, "fp32", args.layout, args.steps,
                                                         args.warmup, torch, launch, args.fuse_period)
    nonfinite = sol32.readout_sums()["nonfinite"]
    if nonfinite:
        raise RuntimeError(f"the FP32 state holds {nonfinite} non-finite cells")
    porosity = 1.0 - float(sol32.solid.mean())
    fused = sol32.loads_mode == "fused"
    inplace = fused and (sol32._inplace_tuned if sol32.inplace is None else bool(sol32.inplace))
    per_call = None
    if args.per_call_steps > 0:
        calls = -(-args.per_call_steps // 60) * 60
        us, wall_us = per_call_rate(sol32, calls, 60, torch)
        sol32.defer_loads = False
        us_plain, wall_plain = per_call_rate(sol32, calls, 60, torch)
        sol32.defer_loads = None
        per_call = {"calls": calls, "read_every": 60, "gpu_us_per_step": us, "wall_us_per_step": wall_us,
                    "undeferred_gpu_us_per_step": us_plain, "undeferred_wall_us_per_step": wall_plain,
                    "note": "collide_stream_psc() in a Python loop reading the diag dict every 60 calls, after "
                            "the timed run; default: loads reduced in the next step's launch"
                            + (", steps launched as in-place pairs" if inplace else "")
                            + "; undeferred: defer_loads=False, a step launch and two load launches per call"}
        log(f"fp32 one call at a time: {us:.1f} us/step ({wall_us:.1f} us wall), undeferred {us_plain:.1f} us/step",
            args.quiet)
    # the plain step kernel's own best launch (the timed loop's launch is tuned for the fused,
    # possibly in-place, form); with --launch both use the given one
    kern_launch = None
    if launch is None:
        kern_launch = importlib.import_module(PKG + ".autotune").tune(sol32)
    kern_ms = step_kernel_ms(sol32, args.kernel_repeats, torch, lib, launch_cfg=kern_launch)
    lbytes, nnz = loads_bytes(sol32)
    ibytes, nsolid = imb_bytes(sol32._csum)
    # fused: every timed step is one k_step_f32_loads launch (in place: one k_step_f32_aa
    # launch, the gather-scatter and local forms alternating), so the HIP-event time of the
    # timed region over the step count is that kernel's average launch duration (it also
    # holds the once-per-run pending loads of the last two steps, ~15 us)
    fused_ms = gpu_ms32 / args.steps if fused else None
    mlups = N * args.steps / wall32 / 1e6
    log(f"fp32: {mlups:.0f} MLUPS (wall), loads {sol32.loads_mode}{' (in-place step pairs)' if inplace else ''}, "
        f"step kernel alone {kern_ms*1e3:.1f} us, "
        f"fused step kernel {fused_ms*1e3 if fused else float('nan'):.1f} us, k={k32:.9e}", args.quiet)

    hbm_check = None
    if not args.no_large_check and n < 256 and launch is None:
        hbm_check = hbm_only_check(psc3d, workloads, torch, lib, args.kernel_repeats // 4, args.quiet)

    fp64 = None
    if not args.no_fp64:
        sol64, k64, wall64, gpu_ms64, _ = timed_run(psc3d, table, n, "fp64", args.layout, args.steps, args.warmup,
                                                    torch)
        kern64 = step_kernel_ms(sol64, max(50, args.kernel_repeats // 4), torch, lib)
        mlups64 = N * args.steps / wall64 / 1e6
        fp64 = {"MLUPS": mlups64, "ms_per_step": wall64 * 1e3 / args.steps, "k": k64,
                "step_kernel_us": kern64 * 1e3, "launch": [sol64._block[0], sol64._block[1], sol64._flags],
                "step_kernel_GBps": BYTES_F64_STATIC * N / (kern64 * 1e-3) / 1e9}
        if args.per_call_steps > 0:
            calls = -(-args.per_call_steps // 60) * 60
            us64, wall_us64 = per_call_rate(sol64, calls, 60, torch)
            fp64["per_call"] = {"calls": calls, "read_every": 60, "gpu_us_per_step": us64,
                                "wall_us_per_step": wall_us64}
            log(f"fp64 one call at a time: {us64:.1f} us/step ({wall_us64:.1f} us wall)", args.quiet)
        log(f"fp64: {mlups64:.0f} MLUPS, step kernel {kern64*1e3:.1f} us, k={k64:.9e}", args.quiet)
        del sol64
    copy_gbps = None
    if not args.no_copy:
        copy_gbps = copy_bandwidth(torch, lib)
        log(f"copy kernel: {copy_gbps:.0f} GB/s", args.quiet)

    cpu = None
    if not args.no_cpu_baseline:
        host = f"host {platform.processor() or platform.machine()}, os.cpu_count()={os.cpu_count()}"
        sizes = []
        # the workload itself: cpu_big_steps timed steps (after one untimed) of the n^3 lattice
        bs = max(1, args.cpu_big_steps)
        big_mlups, big_dt, big_vox = cpu_baseline(table, n, bs, args.quiet)
        big_dt /= bs
        sizes.append({"n": n, "particles": int(table.shape[0]), "steps": bs, "s_per_step": big_dt,
                      "MLUPS": big_mlups, "voxelize_s": big_vox})
        # smaller boxes: same sphere size and volume fraction (cpu_n^3), and the 20^3 box of the
        # moving-sphere benchmark with one centred r=2.5 sphere
        cn = args.cpu_n
        count = max(1, int(round(table.shape[0] * (cn / n) ** 3)))
        ctable = workloads.fixed_array(seed=0, n=cn, count=count)
        for bn, btab in ((cn, ctable), (20, np.array([[10.0, 10.0, 10.0, 2.5, 0, 0, 0, 0, 0, 0]]))):
            m, dt, vx = cpu_baseline(btab, bn, args.cpu_steps, args.quiet)
            sizes.append({"n": bn, "particles": int(btab.shape[0]), "steps": args.cpu_steps,
                          "s_per_step": dt / args.cpu_steps, "MLUPS": m, "voxelize_s": vx})
        cpu = {"value": big_mlups, "unit": "MLUPS", "cores": 1, "kind": "port",
               "sample": f"NumPy FP64 oracle (restatement of psc3d.py), {bs} timed static-geometry steps (after "
                         f"one untimed) of this workload ({n}^3, {table.shape[0]} spheres), {big_dt:.2f} s per "
                         f"step, one thread; {host}",
               "voxelize_s": big_vox, "sizes": sizes,
               "extrapolated_512_s_per_step": big_dt * (512 / n) ** 3}
        if args.cpu_procs > 1:
            procs = min(args.cpu_procs, os.cpu_count() or 1)
            agg, agg_dt = cpu_baseline_aggregate(ctable, cn, max(1, args.cpu_steps // 3), procs, args.quiet)
            cpu["aggregate"] = {"value": agg, "unit": "MLUPS", "cores": procs,
                                "sample": f"{procs} single-threaded processes, each {max(1, args.cpu_steps // 3)} "
                                          f"steps of its own {cn}^3 box ({count} spheres), slowest {agg_dt:.1f} s"}

    step_bytes = BYTES_F32_STATIC * N + ibytes
    if fused:
        # the dominant kernel of the timed step: collide+stream plus the loads of earlier steps
        launch_bytes = step_bytes + lbytes
        achieved = launch_bytes / (fused_ms * 1e-3) / 1e9
        run_launch = (sol32._block[0], sol32._block[1],
                      sol32._flags | (lib_flag("PSC_RUN_INPLACE") if inplace else 0), sol32.effective_fuse_period())
        prefix = "k_step_f32_aa" if inplace else "k_step_f32_loads"
        if inplace:
            roof_kernel, roof_us = ("k_step_f32_aa (static; step + loads of the two previous steps; in-place pairs "
                                    "of a gather-scatter and a local step on one state buffer)", fused_ms * 1e3)
        else:
            roof_kernel, roof_us = ("k_step_f32_loads (pull, static; step + loads of the two previous steps)",
                                    fused_ms * 1e3)
        traffic_detail = None
        if not args.no_traffic:
            traffic_detail = measure_traffic(args, run_launch[:3], run_launch[3], prefix, args.quiet)
        if traffic_detail is not None:
            traffic = traffic_detail["bytes_per_launch"]
            traffic_detail["source"] = "measured in this run: rocprofv3 --pmc over bench.py --traffic-probe, same launch"
        else:
            traffic = read_traffic(f"{prefix}_{n}_pmc.txt", run_launch)
            traffic_detail = {"source": "committed profile with the same launch" if traffic else "unmeasured"}
    else:
        launch_bytes = step_bytes
        achieved = launch_bytes / (kern_ms * 1e-3) / 1e9
        run_launch = (sol32._block[0], sol32._block[1], sol32._flags)
        roof_kernel, roof_us, traffic = ("k_step_f32 (pull, static)", kern_ms * 1e3,
                                         read_traffic(f"k_step_f32_{n}_pmc.txt", run_launch))
        traffic_detail = {"source": "committed profile with the same launch" if traffic else "unmeasured"}
    out = {
        "metric": METRIC,
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall32 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"fixed-array permeability: periodic {n}^3, 60 RSA spheres r=12 (seed 0), tau 0.8, "
                        "F=1e-6 along x, s=4, static geometry, FP32 deviation storage, "
                        f"{args.layout} streaming; step = collide+stream + particle loads "
                        f"({sol32.loads_mode}{', in-place step pairs' if inplace else ''})",
            "lattice": [n, n, n],
            "particles": int(table.shape[0]),
            "porosity": porosity,
            "total_steps": total,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "kernel": roof_kernel,
            "launch": {"block": list(sol32._block), "flags": sol32._flags,
                       "fuse_period": sol32.effective_fuse_period() if fused else None,
                       "inplace_pairs": inplace},
            "bytes_per_launch": launch_bytes,
            "bytes_per_cell": BYTES_F32_STATIC,
            "solid_cells": nsolid,
            "imb_bytes": ibytes,
            "loads_bytes_per_step": lbytes,
            "loads_entries": nnz,
            "kernel_us": roof_us,
            "step_kernel_alone_us": kern_ms * 1e3,
            "step_kernel_alone_launch": list(kern_launch) if kern_launch else [sol32._block[0], sol32._block[1],
                                                                               sol32._flags],
            "step_kernel_alone_GBps": step_bytes / (kern_ms * 1e-3) / 1e9,
            "traffic_detail": traffic_detail,
            "traffic_over_algorithmic": (traffic / launch_bytes) if traffic else None,
        },
        "cpu_baseline": cpu,
        "hbm_only_check": hbm_check,
        "k_fp32": k32,
        "fp64": fp64,
        "k_rel_err_fp32_vs_fp64": (abs(k32 - fp64["k"]) / abs(fp64["k"])) if fp64 else None,
        "fp32_over_fp64_MLUPS": (mlups / fp64["MLUPS"]) if fp64 else None,
        "copy_GBps": copy_gbps,
        "gpu_ms_timed_region": gpu_ms32,
        "gpu_us_per_step": gpu_ms32 * 1e3 / args.steps,
        "k_readout_ms": readout_ms,
        "per_call": per_call,
        "nonfinite_cells": int(nonfinite),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
