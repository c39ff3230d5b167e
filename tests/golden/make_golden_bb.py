"""Write the golden fixtures of the plain bounce-back D3Q19 solver and its
binary voxelizers from the reference implementation.

Run where the reference sources are available (the GPU tests only read the fixtures):

    python tests/golden/make_golden_bb.py [/path/to/reference]

It imports ``scripts/run_3d_window_lbm.py`` (``voxelize``, ``run_lbm``) and
``scripts/run_local_fines_lbm_permeability.py`` (``voxelize_variable``,
``run_force_set``) of wangjianfttt/particuology_Hydraulic-penalty- (NumPy,
FP64), drives them on small seeded cases and stores inputs and outputs as
``tests/golden/bb_*.npz``.  The step's in-place bounce-back loop and the
exact rounding of the voxel tests are what closed forms cannot pin.
Matplotlib, which those scripts import, is pointed at a temporary config
directory.
"""

from __future__ import annotations

import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")


def load_reference():
    os.environ.setdefault("MPLBACKEND", "Agg")
    os.environ.setdefault("MPLCONFIGDIR", tempfile.mkdtemp(prefix="mplconfig_"))
    sys.path.insert(0, str(REF))
    import scripts.run_3d_window_lbm as w  # noqa: E402  (reference module, read-only)
    import scripts.run_local_fines_lbm_permeability as f  # noqa: E402
    return w, f


def main():
    w, f = load_reference()
    rng = np.random.default_rng(7)

    # 1. voxelize: spheres of one radius in an anisotropic box around a centre
    n = 14
    center = np.array([0.51, -0.23, 0.37])
    size = np.array([1.1, 1.0, 1.25])
    points = center + rng.uniform(-0.7, 0.7, size=(18, 3)) * size
    radius = 0.26
    solid, dx = w.voxelize(points, center, size, radius, n)
    np.savez_compressed(HERE / "bb_voxelize.npz", points=points, center=center, size=size, radius=radius, n=n,
                        solid=solid, dx=dx)

    # 2. voxelize_variable: pebbles (type 1) and fines (type 2) in a 9 mm box, fine radii inflated
    box, nv = 0.009, 16
    P = 10
    rows = np.zeros((P, 6))
    rows[:, 0] = np.arange(1, P + 1)
    rows[:, 1] = np.where(np.arange(P) < 4, 1.0, 2.0)
    rows[:, 2:5] = rng.uniform(-0.0005, 0.0095, size=(P, 3))
    rows[:, 5] = np.where(rows[:, 1] == 1.0, rng.uniform(1.2e-3, 1.8e-3, P), rng.uniform(2.0e-4, 2.9e-4, P))
    sv = f.voxelize_variable(rows, box, nv, fine_radius_inflate_dx=0.5)
    sv0 = f.voxelize_variable(rows, box, nv)
    np.savez_compressed(HERE / "bb_voxelize_variable.npz", particles=rows, box_size=box, n=nv, inflate=0.5,
                        solid=sv, solid_no_inflate=sv0)

    # 3. run_lbm on the voxelized window, with fields
    tau, fz, steps = 0.72, 2.0e-5, 60
    res = w.run_lbm(solid, tau=tau, force_z=fz, steps=steps, return_fields=True)
    out = {k: np.asarray(v) for k, v in res.items()}
    np.savez_compressed(HERE / "bb_run_lbm.npz", solid=solid, tau=tau, force_z=fz, steps=steps,
                        **{"out_" + k: v for k, v in out.items()})

    # 4. run_force_set: three force levels on the variable-radius window (Darcy fit)
    result, slope, intercept, r2 = f.run_force_set(sv, 0.8, 1.0e-5, 25, True)
    np.savez_compressed(HERE / "bb_force_set.npz", solid=sv, tau=0.8, force_z=1.0e-5, steps=25, slope=slope,
                        intercept=intercept, r2=r2, superficial=result["superficial_uz"],
                        permeability=result["permeability_lu"])
    print("wrote bb_voxelize.npz bb_voxelize_variable.npz bb_run_lbm.npz bb_force_set.npz")
    print("solid fractions", solid.mean(), sv.mean(), "k", res["permeability_lu"], "slope", slope, "r2", r2)


if __name__ == "__main__":
    main()
