"""Geometry input and convergence post-processing, host side (CPU).

* ``convergence.merge`` on the reference's own per-resolution probe rows
  (tests/golden/convergence/<case>_nXXX.csv, data files of the reference's
  convergence/permeability/ directory) reproduces the reference's merged
  tables in the same directory (permeability_convergence_raw.csv,
  permeability_richardson.csv) value for value;
* the observed-order bisection recovers a planted order;
* the dump readers on a LIGGGHTS custom dump written here.
"""

from __future__ import annotations

import csv
import math
import shutil
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import load_package

cv = load_package("convergence")
gio = load_package("geometry_io")
GOLD = Path(__file__).resolve().parent / "golden" / "convergence"


def _read(path):
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


def test_merge_reproduces_reference_tables(tmp_path):
    for p in GOLD.glob("*_n*.csv"):
        shutil.copy(p, tmp_path / p.name)
    rows, ext = cv.merge(tmp_path)
    raw, ref_raw = _read(tmp_path / "permeability_convergence_raw.csv"), _read(GOLD / "permeability_convergence_raw.csv")
    assert len(raw) == len(ref_raw) == 26
    for a, b in zip(raw, ref_raw):
        assert a == b
    rich, ref_rich = _read(tmp_path / "permeability_richardson.csv"), _read(GOLD / "permeability_richardson.csv")
    assert len(rich) == len(ref_rich) == 3
    for a, b in zip(rich, ref_rich):
        assert a.keys() == b.keys()
        for k in a:
            if k == "case" or a[k] == "" or "resolution" in k:
                assert a[k] == b[k]
            else:
                assert float(a[k]) == float(b[k]), k


def test_observed_order_recovers_planted_order():
    p, k_inf, c = 1.7, 2.5, 40.0
    n = (64, 96, 128)
    h = [1.0 / v for v in n]
    k = [k_inf + c * hi ** p for hi in h]
    got = cv.observed_order(*h, *k)
    assert got == pytest.approx(p, abs=1e-12)
    rows = [{"case": "x", "resolution": str(v), "permeability_lu": repr(ki)} for v, ki in zip(n, k)]
    r = cv.richardson_for_case(rows, "x")
    assert r["permeability_extrapolated_lu"] == pytest.approx(k_inf, rel=1e-10)
    # oscillating differences fall back to order 2
    assert cv.observed_order(*h, 1.0, 2.0, 1.5) == 2.0


DUMP = """ITEM: TIMESTEP
9900
ITEM: NUMBER OF ATOMS
4
ITEM: BOX BOUNDS pp pp ff
0 0.009
0 0.009
0 0.009
ITEM: ATOMS id type x y z vx vy vz fx fy fz radius c_contacts
7 1 0.0021 0.0030 0.0044 0 0 0 0 0 0 0.0015 3
3 2 0.0050 0.0061 0.0012 1e-3 0 0 0 0 1e-9 0.00025 1
12 1 0.0070 0.0011 0.0080 0 0 0 0 0 0 0.0016 2
9 2 0.0043 0.0047 0.0051 0 0 0 0 0 0 0.00028 0
"""


def test_dump_readers(tmp_path):
    p = tmp_path / "w_traj_9900.dump"
    p.write_text(DUMP)
    rows = gio.parse_dump(p)
    assert rows.shape == (4, 6)
    assert np.array_equal(rows[:, 0], [7, 3, 12, 9]) and np.array_equal(rows[:, 1], [1, 2, 1, 2])
    assert np.array_equal(rows[:, 5], [0.0015, 0.00025, 0.0016, 0.00028])
    assert np.array_equal(gio.read_last_dump(p), rows)
    sk = gio.filter_particles(rows, "type==1")
    assert np.array_equal(sk[:, 0], [7, 12])
    t = gio.to_table(rows, 64)
    assert np.allclose(t[:, 0:4] * 0.009 / 64, rows[:, 2:6], rtol=1e-15, atol=0) and np.all(t[:, 4:] == 0)
    bad = tmp_path / "bad.dump"
    bad.write_text(DUMP.replace(" radius", " diameter"))
    with pytest.raises(ValueError):
        gio.parse_dump(bad)
    two = tmp_path / "two.dump"
    two.write_text(DUMP + DUMP)
    assert gio.read_last_dump(two).shape == (8, 6)  # every frame of the file
    assert np.array_equal(gio.parse_dump(two), rows)  # the first frame only
    gap = tmp_path / "gap.dump"
    gap.write_text(DUMP.replace("3 2 0.0050", "\n3 2 0.0050"))
    assert np.array_equal(gio.parse_dump(gap), rows)  # blank lines inside a section are skipped
    w = gio.windows_from_dumps([p], 32, "type==1")
    assert w[0][0] == 32 and w[0][1].shape == (2, 10)


def test_study_edge_rules(tmp_path):
    """A case without a skeleton row at its resolution gets a blank ratio; a
    case with fewer than three resolutions gets a blank Richardson row; the
    skeleton's own ratios are 1; a probe file must hold exactly one row."""
    def probe(case, n, k):
        cv.write_csv(tmp_path / f"{case}_n{n:03d}.csv", ["case", "resolution", "permeability_lu"],
                     [{"case": case, "resolution": str(n), "permeability_lu": repr(k)}])
    for n, k in ((32, 1.0), (48, 1.2), (64, 1.3)):
        probe("skeleton", n, k)
    probe("baseline", 32, 0.9)
    probe("baseline", 96, 1.1)  # no skeleton at 96
    raw, rich = cv.merge(tmp_path)
    by = {(r["case"], r["resolution"]): r["K_over_skeleton_same_resolution"] for r in raw}
    assert by[("baseline", "32")] == repr(0.9) and by[("baseline", "96")] == ""
    assert by[("skeleton", "48")] == repr(1.0)
    row = {r["case"]: r for r in rich}
    assert all(v == "" for k, v in row["baseline"].items() if k != "case")
    assert row["skeleton"]["K_over_skeleton_extrapolated"] == 1.0
    assert row["skeleton"]["fine_resolution"] == 64
    (tmp_path / "broken_n128.csv").write_text("case,resolution,permeability_lu\nx,128,1.0\nx,128,2.0\n")
    with pytest.raises(ValueError, match="one data row"):
        cv.merge(tmp_path)


def test_bisection_helper_brackets_the_root():
    got = cv.bisect_rising(lambda x: x ** 3, 2.0, 0.0, 4.0, 60)
    assert got == pytest.approx(2.0 ** (1 / 3), abs=1e-15)
