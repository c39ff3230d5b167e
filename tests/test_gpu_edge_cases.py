"""Degenerate lattices and geometries (GPU) against the oracle.

psc3d.py streams with np.roll and clamps sphere boxes to the lattice
(psc3d.py:60-94, :151-179, :234-241), so it is defined for any edge of at
least one cell and any sphere placement.  The cases here are the corners of
that: lattices of one, two and three cells per axis (a cell's upstream and
downstream neighbours coincide, or are the cell itself), no particles, a
sphere larger than the lattice (every cell solid), a sphere smaller than the
sub-sample spacing (no covered cell, zero loads), and spheres centred outside
the lattice (clamped boxes).  Each runs single calls, a native run() and
single calls again, in both class layouts.

Tolerances as in test_gpu_parity.py: FP64 f bit-identical to the oracle
(no cell here is covered by two spheres), loads to 1e-10 of the largest;
FP32 f within 1e-4 of the signal max|f_ref - w|, loads to 1e-3.  One
exception: on a one-cell lattice the (19, 1, 1, 1) state is contiguous along
axis 0, where NumPy's sum(axis=0) adds pairwise (eight partial sums) instead
of row by row, so the reference's own rho and moments there round in another
order than on any larger lattice (and than the kernel, which keeps the row
order); f then agrees to a few ulp.  test_numpy_sums_one_cell_pairwise
checks that premise on the CPU.
"""

import numpy as np
import pytest

from oracle import psc_oracle as O

pytestmark = pytest.mark.gpu

FORCE = (1.0e-5, -2.0e-5, 3.0e-5)
# name: (n, sphere rows (x, y, z, r[, U, omega]), static geometry)
CASES = {
    "empty6": (6, [], True),
    "one_cell": (1, [(0.5, 0.5, 0.5, 0.45)], True),
    "two_cells_moving": (2, [(1.0, 1.0, 1.0, 0.9, 1.0e-3, 0.0, -1.0e-3, 0.0, 2.0e-3, 0.0)], False),
    "all_solid3": (3, [(1.5, 1.5, 1.5, 10.0)], True),
    "tiny8": (8, [(4.1, 3.9, 4.2, 0.05), (1.5, 6.5, 1.5, 1.2)], True),
    "outside10": (10, [(-2.0, 5.0, 11.5, 4.0), (7.0, 12.0, 3.0, 3.0)], True),
}


def _steps(s):
    """6 single calls, a native run of 4, one more call: 11 steps."""
    for _ in range(6):
        s.collide_stream_psc()
    s.run(4)
    s.collide_stream_psc()


@pytest.mark.parametrize("layout", ["dense", "sparse"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(CASES))
def test_degenerate_case_matches_oracle(psc3d, name, precision, layout):
    n, rows, static = CASES[name]
    print(f"[edge] {name} n={n} {precision} {layout}", flush=True)
    cls = psc3d.PscD3Q19 if layout == "dense" else psc3d.SparsePscD3Q19
    s = cls(n, [psc3d.Particle3D(*r) for r in rows], tau=0.8, subsamples=3, body_force=FORCE,
            static_geometry=static, precision=precision, device="cuda:0")
    o = O.OracleSolver(n, [O.Sphere(*r) for r in rows], tau=0.8, subsamples=3, body_force=FORCE,
                       static_geometry=static, layout=layout)
    _steps(s)
    for _ in range(11):
        o.collide_stream_psc()

    np.testing.assert_allclose(s.solid, o.solid, rtol=0, atol=4e-16)
    if name == "all_solid3":
        assert np.all(o.solid == 1.0)
    if name == "tiny8":
        assert np.all(o.hydro_forces[0] == 0.0) and np.all(o.hydro_torques[0] == 0.0)
    signal = np.max(np.abs(o.f - O.W[:, None, None, None]))
    assert signal > 1e-6  # the force has moved the state away from rest
    err = np.max(np.abs(s.f - o.f))
    scale = max(np.max(np.abs(o.hydro_forces)), 1e-30) if len(rows) else 1.0
    ef = np.max(np.abs(s.hydro_forces - o.hydro_forces)) / scale if len(rows) else 0.0
    et = np.max(np.abs(s.hydro_torques - o.hydro_torques)) / (scale * (n + 10.0)) if len(rows) else 0.0
    print(f"  max|f - f_oracle| = {err:.3e} (signal {signal:.3e}), rel force err {ef:.3e}, torque {et:.3e}")
    assert s.hydro_forces.shape == o.hydro_forces.shape
    if precision == "fp64":
        if n == 1:
            assert err <= 8 * np.spacing(np.max(np.abs(o.f)))
        else:
            assert err == 0.0
        assert ef <= 1e-10 and et <= 1e-10
    else:
        assert err <= 1e-4 * signal
        assert ef <= 1e-3 and et <= 1e-3
    if name == "tiny8":
        assert np.all(s.hydro_forces[0] == 0.0) and np.all(s.hydro_torques[0] == 0.0)
