"""Dynamic geometry (static_geometry=False) voxelizes again only when the
particle table changed: the reference re-voxelizes on every call
(psc3d.py:209-210), and an unchanged table gives the same solid fraction,
solid velocity and entries.  A solver that skips those rebuilds must step
exactly like one forced to rebuild before every step, and an edit of the
caller's own particle array in place (as the drivers make,
run_precision_validation.py:257-258) must still reach the next step."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_unchanged_particles_skip_the_rebuild_with_the_same_results(psc3d, precision):
    n = 16
    table = np.array([[8.2, 7.9, 8.4, 3.1, 0.0, 0.0, 1.0e-4, 0.0, 2.0e-5, 0.0]])
    forced_table = table.copy()
    kw = dict(tau=0.8, subsamples=2, body_force=(0.0, 0.0, 1.0e-6), static_geometry=False, precision=precision,
              device="cuda:0")
    lazy = psc3d.SparsePscD3Q19(n, table, **kw)          # holds the caller's array itself
    forced = psc3d.SparsePscD3Q19(n, forced_table, **kw)
    solid_before = None
    for step in range(8):
        if step == 4:
            # in-place edits of both callers' arrays: a new centre and angular velocity
            table[0, 0] += 0.37
            table[0, 8] = -3.0e-5
            forced_table[0, 0] += 0.37
            forced_table[0, 8] = -3.0e-5
            solid_before = lazy.solid.copy()
        lazy.collide_stream_psc()
        forced.rebuild_geometry()   # a rebuild before every step, as the reference does
        forced.collide_stream_psc()
        print(f"[{precision}] step {step + 1}", flush=True)
    assert not np.array_equal(lazy.solid, solid_before), "the in-place edit did not reach the geometry"
    assert np.array_equal(lazy.solid, forced.solid)
    assert np.array_equal(lazy.f, forced.f)
    assert np.array_equal(lazy.hydro_forces, forced.hydro_forces)
    assert np.array_equal(lazy.hydro_torques, forced.hydro_torques)
