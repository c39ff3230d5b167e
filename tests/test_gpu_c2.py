"""The fixed-array permeability case at full size (GPU): periodic 128^3,
60 non-overlapping spheres of radius 12 by random sequential addition
(seed 0), tau 0.8, body force 1e-6 along x, 4^3 sub-samples, static
geometry -- the benchmarked workload (run_precision_validation.py:404-447,
k at :438; run_psc_permeability_resolution.py:144-150 for the readout).

* The FP64 kernel against the oracle for three steps: f bit-identical over
  the whole lattice, the solid field equal, loads within 1e-10 relative (tree
  sums against NumPy's), the permeability within 1e-12.
* The FP32 deviation path exactly as benchmarked (tuned launch, loads fused
  into the step kernels, in-place step pairs where the tuner picks them)
  against the FP64 kernel after 5000 steps: |k32 - k64| / k64 <= 1e-5, the
  accuracy target of this case, and the per-particle forces within 1e-4 of
  the largest.
"""

import numpy as np
import pytest

from oracle import psc_oracle as O
from tests.conftest import load_package

pytestmark = pytest.mark.gpu

N, TAU, S, F = 128, 0.8, 4, (1.0e-6, 0.0, 0.0)


def _table():
    return load_package("workloads").fixed_array(seed=0, n=N)


def _gpu(psc3d, table, precision):
    return psc3d.PscD3Q19(N, table.copy(), tau=TAU, subsamples=S, body_force=F, static_geometry=True,
                          precision=precision, device="cuda:0")


def test_fixed_array_fp64_matches_oracle_at_full_size(psc3d):
    table = _table()
    print(f"[c2 fp64 vs oracle] {N}^3, {len(table)} spheres, 3 steps", flush=True)
    o = O.OracleSolver(N, [O.Sphere(*row) for row in table], tau=TAU, subsamples=S, body_force=F,
                       static_geometry=True, layout="sparse")
    g = _gpu(psc3d, table, "fp64")
    for step in range(3):
        o.collide_stream_psc()
        g.collide_stream_psc()
        print(f"  step {step + 1} done", flush=True)
    assert np.array_equal(g.solid, o.solid)
    assert 0.75 < 1.0 - g.solid.mean() < 0.83  # 60 spheres of r = 12 fill ~21 % of the box
    f = g.f
    bad = int(np.count_nonzero(f != o.f))
    print(f"  cells x directions with f != oracle: {bad}", flush=True)
    assert bad == 0  # bit-exact: the reference's operation order, x-directed force included
    scale = np.max(np.abs(o.hydro_forces))
    ef = np.max(np.abs(g.hydro_forces - o.hydro_forces)) / scale
    et = np.max(np.abs(g.hydro_torques - o.hydro_torques)) / (scale * 14.0)
    print(f"  rel force err {ef:.2e}, torque err {et:.2e}", flush=True)
    assert ef <= 1e-10 and et <= 1e-10
    k_gpu = g.permeability(axis=0)
    k_ora = O.permeability(o.solid, o.ux, N, TAU, F[0])
    assert k_gpu == pytest.approx(k_ora, rel=1e-12)


def test_fixed_array_fp32_permeability_within_1e5_of_fp64_after_5000_steps(psc3d):
    table = _table()
    steps = 5000
    print(f"[c2 fp32 vs fp64] {N}^3, {steps} steps", flush=True)
    s32 = _gpu(psc3d, table, "fp32")
    s32.rebuild_geometry()
    s32.run(steps)
    k32 = s32.permeability(axis=0)
    print(f"  fp32 loads {s32.loads_mode}, launch {s32._block} {s32._flags:#x} "
          f"in place {bool(s32._inplace_tuned if s32.inplace is None else s32.inplace)}: k = {k32:.9e}", flush=True)
    s64 = _gpu(psc3d, table, "fp64")
    s64.rebuild_geometry()
    s64.run(steps)
    k64 = s64.permeability(axis=0)
    rel = abs(k32 - k64) / abs(k64)
    print(f"  fp64 k = {k64:.9e}; |k32 - k64| / k64 = {rel:.2e}", flush=True)
    assert k64 > 0
    assert rel <= 1e-5
    F32, F64 = s32.hydro_forces, s64.hydro_forces
    ef = np.max(np.abs(F32 - F64)) / np.max(np.abs(F64))
    print(f"  max force err / max |F| = {ef:.2e}", flush=True)
    assert ef <= 1e-4
    assert s32.readout_sums()["nonfinite"] == 0 and s64.readout_sums()["nonfinite"] == 0
