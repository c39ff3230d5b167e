"""Grid-convergence study of the permeability probes (host side).

A permeability probe at one resolution leaves one CSV file,
``<case>_n<NNN>.csv``, holding a single data row (``geometry_io.write_probe``
writes them; the reference's campaign writes the same files).  A
``ConvergenceStudy`` loads a directory of such files and derives

* the raw table: every probe row, ordered by case then resolution, with the
  permeability over that of the skeleton case at the same resolution;
* the Richardson table: per case, the extrapolation h -> 0 (h = 1/n) from its
  three finest resolutions, with the observed order p solving
  (h1^p - h2^p) / (h2^p - h3^p) = |(k1 - k2) / (k2 - k3)|,
  and the extrapolated permeability over the skeleton's.

Specification followed (scripts/merge_permeability_convergence.py:93-146):
the two file names and their columns; the ratio is written as ``repr`` of
the float; p is bracketed in [0.1, 8] and halved 80 times, and taken as 2
when the two differences vanish or change sign; k_inf = (k3 h2^p - k2 h3^p)
/ (h2^p - h3^p), or k3 when the denominator vanishes; a ratio that cannot be
formed (no skeleton row, an unreadable or zero value) is left blank.  On the
reference's own probe rows the tables come out identical to its merged
tables (tests/test_io_convergence_cpu.py).
"""

from __future__ import annotations

import csv
from dataclasses import dataclass, field
from pathlib import Path
import re
from typing import Callable, Iterable, Sequence

SUMMARY_FIELDS = [
    "case", "resolution", "steps", "tau", "force_z", "subsamples", "geometry_workers", "source_dump",
    "particle_filter", "n_particles", "porosity_psc", "mean_uz_fluid", "superficial_uz", "max_uz", "permeability_lu",
    "K_over_skeleton_same_resolution", "momentum_residual", "particle_force_sum_x_lu", "particle_force_sum_y_lu",
    "particle_force_sum_z_lu", "particle_force_norm_sum_lu", "fluid_imb_sum_x_lu", "fluid_imb_sum_y_lu",
    "fluid_imb_sum_z_lu", "particle_fluid_momentum_closure_lu", "domain_body_force_z_lu",
    "fluid_volume_body_force_z_lu",
]
EXTRAP_FIELDS = [
    "case", "coarse_resolution", "mid_resolution", "fine_resolution", "richardson_order",
    "permeability_extrapolated_lu", "fine_grid_permeability_lu", "relative_fine_to_extrapolated",
    "K_over_skeleton_extrapolated",
]
RAW_NAME = "permeability_convergence_raw.csv"
RICHARDSON_NAME = "permeability_richardson.csv"
REFERENCE_CASE = "skeleton"
PROBE_FILE = re.compile(r"^(?P<case>.+)_n(?P<n>\d{3})\.csv$")

ORDER_BRACKET = (0.1, 8.0)
ORDER_HALVINGS = 80
ORDER_FALLBACK = 2.0
TINY = 1.0e-300


def write_csv(path, columns: Iterable[str], rows: Iterable[dict]) -> None:
    """A CSV table of ``rows`` restricted to ``columns`` (missing ones blank)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", newline="") as fh:
        out = csv.DictWriter(fh, fieldnames=list(columns), extrasaction="ignore")
        out.writeheader()
        out.writerows(rows)


# Name kept for geometry_io, which writes single probe rows with it.
write_rows = write_csv


def bisect_rising(fn: Callable[[float], float], target: float, lo: float, hi: float, halvings: int) -> float:
    """Midpoint of the bracket [lo, hi] after ``halvings`` halvings towards
    fn(x) = target, for fn rising on the bracket (the lower half is kept
    while fn(mid) >= target)."""
    for _ in range(halvings):
        mid = 0.5 * (lo + hi)
        if fn(mid) < target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def observed_order(h1, h2, h3, k1, k2, k3) -> float:
    """Convergence order p of k(h) from three spacings h1 > h2 > h3."""
    d_coarse, d_fine = k1 - k2, k2 - k3
    if d_coarse == 0.0 or d_fine == 0.0 or d_coarse * d_fine <= 0.0:
        return ORDER_FALLBACK  # no monotone convergence to fit

    def spacing_ratio(p):
        return (h1 ** p - h2 ** p) / max(h2 ** p - h3 ** p, TINY)

    return bisect_rising(spacing_ratio, abs(d_coarse / d_fine), *ORDER_BRACKET, ORDER_HALVINGS)


@dataclass
class Extrapolation:
    """Richardson extrapolation of one case from resolutions n1 < n2 < n3."""

    case: str
    resolutions: tuple = ()
    order: float | None = None
    k_inf: float | None = None
    k_fine: float | None = None
    over_skeleton: float | str = ""

    @classmethod
    def from_points(cls, case: str, points: Sequence[tuple[int, float]]) -> "Extrapolation":
        """``points``: (resolution, permeability) of the case, ascending; the
        three finest are used, fewer give an empty row."""
        if len(points) < 3:
            return cls(case)
        (n1, k1), (n2, k2), (n3, k3) = points[-3:]
        h1, h2, h3 = 1.0 / n1, 1.0 / n2, 1.0 / n3
        p = observed_order(h1, h2, h3, k1, k2, k3)
        den = h2 ** p - h3 ** p
        k_inf = (k3 * h2 ** p - k2 * h3 ** p) / den if abs(den) > TINY else k3
        return cls(case, (n1, n2, n3), p, k_inf, k3)

    def row(self) -> dict:
        if self.order is None:
            out = dict.fromkeys(EXTRAP_FIELDS, "")
            out["case"] = self.case
            return out
        n1, n2, n3 = self.resolutions
        return {
            "case": self.case, "coarse_resolution": n1, "mid_resolution": n2, "fine_resolution": n3,
            "richardson_order": self.order, "permeability_extrapolated_lu": self.k_inf,
            "fine_grid_permeability_lu": self.k_fine,
            "relative_fine_to_extrapolated": abs(self.k_fine - self.k_inf) / max(abs(self.k_inf), TINY),
            "K_over_skeleton_extrapolated": self.over_skeleton,
        }


def _ratio(num, den):
    """num / den as floats, or None when either cannot be read or den is 0."""
    try:
        return float(num) / float(den)
    except (TypeError, ValueError, ZeroDivisionError):
        return None


@dataclass
class ConvergenceStudy:
    """The probe rows of one convergence campaign (dicts of CSV strings)."""

    rows: list = field(default_factory=list)

    @classmethod
    def from_directory(cls, directory) -> "ConvergenceStudy":
        rows = []
        for path in sorted(Path(directory).glob("*_n*.csv")):
            if PROBE_FILE.match(path.name) is None:
                continue  # the merged tables and other files
            with path.open(newline="") as fh:
                data = list(csv.DictReader(fh))
            if len(data) != 1:
                raise ValueError(f"probe file {path} must hold one data row, it holds {len(data)}")
            rows.append(data[0])
        rows.sort(key=lambda r: (r["case"], int(r["resolution"])))
        return cls(rows)

    def cases(self) -> list:
        return sorted({r["case"] for r in self.rows})

    def points(self, case: str) -> list:
        """(resolution, permeability) of a case, ascending in resolution."""
        pts = [(int(r["resolution"]), float(r["permeability_lu"])) for r in self.rows if r.get("case") == case]
        return sorted(pts, key=lambda t: t[0])

    def raw_table(self) -> list:
        """Probe rows with K_over_skeleton_same_resolution filled in."""
        skeleton = {r["resolution"]: float(r["permeability_lu"]) for r in self.rows
                    if r.get("case") == REFERENCE_CASE}
        table = []
        for r in self.rows:
            q = _ratio(r["permeability_lu"], skeleton[r["resolution"]]) if r["resolution"] in skeleton else None
            table.append({**r, "K_over_skeleton_same_resolution": "" if q is None else repr(q)})
        return table

    def richardson_table(self) -> list:
        """One Extrapolation per case, with K_over_skeleton_extrapolated."""
        ext = [Extrapolation.from_points(c, self.points(c)) for c in self.cases()]
        skel = next((e for e in ext if e.case == REFERENCE_CASE), None)
        if skel is not None and skel.k_inf is not None:
            for e in ext:
                q = _ratio(e.k_inf, skel.k_inf)
                e.over_skeleton = "" if q is None else q
        return ext


def merge(input_dir, output_dir=None):
    """Write the raw and Richardson tables of the probe files in
    ``input_dir`` into ``output_dir`` (default: ``input_dir``); return
    (raw rows, Richardson rows) as written."""
    out = Path(output_dir) if output_dir is not None else Path(input_dir)
    study = ConvergenceStudy.from_directory(input_dir)
    raw = study.raw_table()
    rich = [e.row() for e in study.richardson_table()]
    write_csv(out / RAW_NAME, SUMMARY_FIELDS, raw)
    write_csv(out / RICHARDSON_NAME, EXTRAP_FIELDS, rich)
    return raw, rich


def richardson_for_case(rows, case: str) -> dict:
    """The Richardson row of ``case`` from probe rows (dicts with case,
    resolution and permeability_lu)."""
    return Extrapolation.from_points(case, ConvergenceStudy(list(rows)).points(case)).row()
