"""Permeability grid-convergence post-processing (host side).

The reference's convergence campaign writes one ``<case>_nXXX.csv`` row per
(case, resolution) permeability probe and merges them
(scripts/merge_permeability_convergence.py:93-146): a raw table with the
permeability of each case over the skeleton's at the same resolution, and a
Richardson extrapolation per case from its three finest resolutions, with the
observed order found by bisection of (h1^p - h2^p)/(h2^p - h3^p) = |e12/e23|
on [0.1, 8] (80 halvings; order 2 when the differences vanish or change sign).
This module produces the same two tables, with the same columns, from the
same files (``permeability_probe`` in ``geometry_io`` writes them).
"""

from __future__ import annotations

import csv
import re
from pathlib import Path
from typing import Iterable

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
_NAME = re.compile(r"^(?P<case>.+)_n(?P<n>\d{3})\.csv$")


def write_rows(path, fieldnames: Iterable[str], rows) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with path.open("w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(fieldnames), extrasaction="ignore")
        w.writeheader()
        for r in rows:
            w.writerow(r)


def collect_rows(input_dir) -> list[dict]:
    """The single data row of every ``<case>_nXXX.csv`` in ``input_dir``,
    sorted by (case, resolution)."""
    rows = []
    for path in sorted(Path(input_dir).glob("*_n*.csv")):
        if not _NAME.match(path.name):
            continue
        with path.open(newline="") as fh:
            data = list(csv.DictReader(fh))
        if len(data) != 1:
            raise ValueError(f"{path} should contain exactly one data row")
        rows.append(data[0])
    return sorted(rows, key=lambda r: (r["case"], int(r["resolution"])))


def add_raw_ratios(rows) -> None:
    """K_over_skeleton_same_resolution = k / k(skeleton, same resolution),
    blank where there is no skeleton row (or it cannot be formed)."""
    skel = {r["resolution"]: float(r["permeability_lu"]) for r in rows if r.get("case") == "skeleton"}
    for r in rows:
        try:
            r["K_over_skeleton_same_resolution"] = repr(float(r["permeability_lu"]) / skel[r["resolution"]])
        except (KeyError, ValueError, ZeroDivisionError):
            r["K_over_skeleton_same_resolution"] = ""


def observed_order(h1, h2, h3, k1, k2, k3) -> float:
    e12, e23 = k1 - k2, k2 - k3
    if e12 == 0.0 or e23 == 0.0 or e12 * e23 <= 0.0:
        return 2.0
    target = abs(e12 / e23)
    lo, hi = 0.1, 8.0
    for _ in range(80):
        p = 0.5 * (lo + hi)
        if (h1 ** p - h2 ** p) / max(h2 ** p - h3 ** p, 1.0e-300) < target:
            lo = p
        else:
            hi = p
    return 0.5 * (lo + hi)


def richardson_for_case(rows, case) -> dict:
    """Extrapolation from the case's three finest resolutions (h = 1/n)."""
    cr = sorted([r for r in rows if r.get("case") == case], key=lambda r: int(r["resolution"]))
    if len(cr) < 3:
        out = {k: "" for k in EXTRAP_FIELDS}
        out["case"] = case
        return out
    n1, n2, n3 = (int(r["resolution"]) for r in cr[-3:])
    k1, k2, k3 = (float(r["permeability_lu"]) for r in cr[-3:])
    h1, h2, h3 = 1.0 / n1, 1.0 / n2, 1.0 / n3
    p = observed_order(h1, h2, h3, k1, k2, k3)
    den = h2 ** p - h3 ** p
    k_inf = (k3 * h2 ** p - k2 * h3 ** p) / den if abs(den) > 1.0e-300 else k3
    return {
        "case": case, "coarse_resolution": n1, "mid_resolution": n2, "fine_resolution": n3,
        "richardson_order": p, "permeability_extrapolated_lu": k_inf, "fine_grid_permeability_lu": k3,
        "relative_fine_to_extrapolated": abs(k3 - k_inf) / max(abs(k_inf), 1.0e-300),
        "K_over_skeleton_extrapolated": "",
    }


def add_extrapolated_ratios(rows) -> None:
    skel = next((r for r in rows if r["case"] == "skeleton"), None)
    if not skel:
        return
    try:
        ks = float(skel["permeability_extrapolated_lu"])
    except (TypeError, ValueError):
        return
    for r in rows:
        try:
            r["K_over_skeleton_extrapolated"] = float(r["permeability_extrapolated_lu"]) / ks
        except (TypeError, ValueError, ZeroDivisionError):
            r["K_over_skeleton_extrapolated"] = ""


def merge(input_dir, output_dir=None):
    """Write ``permeability_convergence_raw.csv`` and
    ``permeability_richardson.csv`` into ``output_dir`` (default: the input
    directory); return (raw rows, extrapolation rows)."""
    out = Path(output_dir) if output_dir is not None else Path(input_dir)
    rows = collect_rows(input_dir)
    add_raw_ratios(rows)
    write_rows(out / "permeability_convergence_raw.csv", SUMMARY_FIELDS, rows)
    ext = [richardson_for_case(rows, c) for c in sorted({r["case"] for r in rows})]
    add_extrapolated_ratios(ext)
    write_rows(out / "permeability_richardson.csv", EXTRAP_FIELDS, ext)
    return rows, ext
