"""Permeability probe of a dumped window on the GPU against the oracle, and
the convergence merge of probe rows (GPU)."""

import numpy as np
import pytest

from oracle import psc_oracle as O
from tests.conftest import load_package
from tests.test_io_convergence_cpu import DUMP

pytestmark = pytest.mark.gpu


def _oracle_row(rows, n, steps, tau, fz, s):
    gio = load_package("geometry_io")
    t = gio.to_table(rows, n)
    ora = O.OracleSolver(n, [O.Sphere(*r[:4]) for r in t], tau=tau, subsamples=s, body_force=(0.0, 0.0, fz),
                         static_geometry=True, layout="sparse")
    for _ in range(steps):
        d = ora.collide_stream_psc()
    w = 1.0 - ora.solid
    wuz = ora.uz * w
    fsum = ora.hydro_forces.sum(axis=0)
    return {"porosity_psc": float(np.sum(w)) / n ** 3, "mean_uz_fluid": float(np.sum(wuz) / float(np.sum(w))),
            "superficial_uz": float(np.sum(wuz) / n ** 3), "max_uz": float(np.max(ora.uz)),
            "fluid_imb_sum_z_lu": float(np.sum(ora.imb_mz)), "particle_force_sum_z_lu": float(fsum[2]),
            "momentum_residual": d["momentum_residual"]}


def test_permeability_probe_matches_oracle(tmp_path):
    gio = load_package("geometry_io")
    cv = load_package("convergence")
    p = tmp_path / "w_traj_9900.dump"
    p.write_text(DUMP)
    rows = gio.parse_dump(p)
    n, steps, tau, fz = 24, 30, 0.72, 2.0e-5
    print(f"[probe] {n}^3 x {steps} vs oracle", flush=True)
    row = gio.permeability_probe(rows, "baseline", n, steps=steps, tau=tau, force_z=fz, source_dump=str(p),
                                 device="cuda:0")
    ref = _oracle_row(rows, n, steps, tau, fz, 2)
    for k in ("porosity_psc", "mean_uz_fluid", "superficial_uz", "max_uz", "fluid_imb_sum_z_lu"):
        assert row[k] == ref[k], k  # fields are bit-identical, the host sums identical
    assert row["particle_force_sum_z_lu"] == pytest.approx(ref["particle_force_sum_z_lu"], rel=1e-10)
    assert row["permeability_lu"] == (tau - 0.5) / 3.0 * row["superficial_uz"] / fz
    # three resolutions of two cases -> convergence tables
    for case, filt in (("skeleton", "type==1"), ("baseline", "all")):
        for res in (16, 20, 24):
            gio.write_probe(tmp_path, gio.permeability_probe(rows, case, res, steps=10, tau=tau, force_z=fz,
                                                             particle_filter=filt, device="cuda:0"))
    raw, ext = cv.merge(tmp_path)
    assert len(raw) == 6 and [e["case"] for e in ext] == ["baseline", "skeleton"]
    assert all(r["K_over_skeleton_same_resolution"] != "" for r in raw)
    assert ext[1]["K_over_skeleton_extrapolated"] == 1.0
