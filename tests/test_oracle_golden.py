"""The oracle against the reference's own outputs (CPU).

tests/golden/*.npz were written by tests/golden/make_golden.py, which drives
src/lbm_dem/psc3d.py of the reference.  The oracle must reproduce them bit
for bit: f, solid, per-step forces/torques and diag values.
"""

import numpy as np
import pytest

from oracle import psc_oracle as O
from tests import _golden as G


def _oracle(g):
    parts = [O.Sphere(*row) for row in g["particles"]]
    s = O.OracleSolver(int(g["n"]), parts, tau=float(g["tau"]), subsamples=int(g["subsamples"]),
                       body_force=tuple(float(v) for v in g["force"]), static_geometry=bool(g["static"]),
                       layout=str(g["layout"]))
    return s, parts


@pytest.mark.parametrize("name", ["smoke12", "moving20", "spin16", "sparse16"])
def test_trajectory_bit_exact(name):
    g = G.load(name)
    s, parts = _oracle(g)
    F, T, D = [], [], []
    for k in range(int(g["steps"])):
        if name == "sparse16":
            G.move_sparse16(parts, k)
        d = s.collide_stream_psc()
        F.append(s.hydro_forces.copy())
        T.append(s.hydro_torques.copy())
        D.append([d[key] for key in O_DIAG])
    assert np.array_equal(np.array(F), g["forces"])
    assert np.array_equal(np.array(T), g["torques"])
    assert np.array_equal(np.array(D), g["diag"])
    assert np.array_equal(s.solid, g["solid"])
    assert np.array_equal(s.uz, g["uz"])
    assert np.array_equal(np.stack([s.imb_mx, s.imb_my, s.imb_mz]), g["imb"])
    if "f" in g:
        assert np.array_equal(s.f, g["f"])
    else:
        assert np.array_equal(s.f.ravel()[g["f_idx"]], g["f_val"])
        assert np.array_equal(s.f.sum(axis=(1, 2, 3)), g["f_sum"])


O_DIAG = ("solid_volume_lu", "mean_uz", "momentum_residual", "max_force", "max_torque")


def test_sparse_lists_bit_exact():
    g = G.load("sparse16")
    s, _ = _oracle(g)
    s.rebuild_geometry()
    assert [len(a) for a in s.node_ids] == list(g["ids_len"])
    assert np.array_equal(np.concatenate(s.node_ids), g["ids"])
    assert np.array_equal(np.concatenate(s.alpha_vals), g["vals"])


def test_dense_and_sparse_layouts():
    g = G.load("dense_vs_sparse16")
    for layout in ("dense", "sparse"):
        parts = [O.Sphere(*row) for row in g["particles"]]
        s = O.OracleSolver(int(g["n"]), parts, tau=float(g["tau"]), subsamples=int(g["subsamples"]),
                           body_force=tuple(g["force"]), static_geometry=True, layout=layout)
        for _ in range(int(g["steps"])):
            s.collide_stream_psc()
        assert np.array_equal(s.hydro_forces, g[layout + "_forces"])
        assert np.array_equal(s.hydro_torques, g[layout + "_torques"])
        assert np.array_equal(s.solid, g[layout + "_solid"])
    # the two reference classes agree to summation-order round-off
    np.testing.assert_allclose(g["dense_forces"], g["sparse_forces"], rtol=1e-9, atol=1e-22)


def test_permeability_trace():
    g = G.load("perm24")
    s, _ = _oracle(g)
    n, tau, fz = int(g["n"]), float(g["tau"]), float(g["force"][2])
    ks = []
    for step in range(int(g["steps"])):
        s.collide_stream_psc()
        if (step + 1) % 20 == 0:
            ks.append(O.permeability(s.solid, s.uz, n, tau, fz))
    assert np.array_equal(np.array(ks), g["k_trace"])
    assert np.array_equal(s.f.ravel()[g["f_idx"]], g["f_val"])
