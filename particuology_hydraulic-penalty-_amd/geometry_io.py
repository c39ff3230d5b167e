"""Geometry input: LIGGGHTS dump readers and the permeability probe of a
window (host side; the probe itself runs on the GPU solver).

* ``atoms_frames`` -- the ATOMS sections of a LIGGGHTS custom dump;
* ``parse_dump`` -- the first section as rows (id, type, x, y, z, radius)
  (the probe driver's reader, run_psc_permeability_resolution.py:44-68);
* ``read_last_dump`` -- the rows of every section of the file
  (analyze_local_fines_geometry.py:27-45);
* ``filter_particles`` ("all" or "type==1", the skeleton) and ``to_particles``
  / ``to_table`` (SI metres -> lattice units at a resolution, box 9 mm)
  (run_psc_permeability_resolution.py:71-82);
* ``permeability_probe`` -- the probe of run_psc_permeability_resolution.py:
  118-196 on the GPU solver: a static-geometry SparsePscD3Q19 run, then the
  same readouts and CSV row, which ``convergence.merge`` consumes;
* ``windows_from_dumps`` -- real windows for the batched solver.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

BOX_SIZE_M = 9.0e-3
REQUIRED = ("id", "type", "x", "y", "z", "radius")


def atoms_frames(path) -> list:
    """Every ``ITEM: ATOMS`` section of a LIGGGHTS/LAMMPS text dump, in file
    order, as (column names, (rows, columns) float array).  A section holds
    the non-blank lines up to the next ``ITEM:`` line."""
    frames = []
    names, body = None, []

    def close():
        if names is not None:
            vals = np.array(" ".join(body).split(), dtype=float) if body else np.zeros(0)
            frames.append((names, vals.reshape(-1, len(names)) if len(names) else vals.reshape(0, 0)))

    for line in Path(path).read_text().splitlines():
        if line.startswith("ITEM:"):
            close()
            names = tuple(line.split()[2:]) if line.startswith("ITEM: ATOMS") else None
            body = []
        elif names is not None and line.strip():
            body.append(line)
    close()
    if not frames:
        raise ValueError(f"No ITEM: ATOMS section in {path}")
    return frames


def _select(names, values, path) -> np.ndarray:
    """The (id, type, x, y, z, radius) columns of one frame."""
    missing = [c for c in REQUIRED if c not in names]
    if missing:
        raise ValueError(f"{path} missing dump columns: {missing}")
    return values[:, [names.index(c) for c in REQUIRED]] if len(values) else np.zeros((0, 6))


def parse_dump(path) -> np.ndarray:
    """(P, 6) rows (id, type, x, y, z, radius) of the first ATOMS section,
    id and type truncated to integers as the probe driver reads them."""
    names, values = atoms_frames(path)[0]
    out = _select(names, values, path)
    out[:, 0:2] = np.trunc(out[:, 0:2])
    return out


def read_last_dump(path) -> np.ndarray:
    """(P, 6) rows of every ATOMS section of the file, one after the other
    (the campaign's dumps hold one frame per file)."""
    return np.concatenate([_select(names, values, path) for names, values in atoms_frames(path)])


def filter_particles(rows, particle_filter: str = "all") -> np.ndarray:
    if particle_filter == "type==1":
        return rows[rows[:, 1] == 1]
    if particle_filter != "all":
        raise ValueError(f"unknown particle filter {particle_filter!r}")
    return rows


def to_table(rows, resolution: int, box_size: float = BOX_SIZE_M) -> np.ndarray:
    """(P, 10) stationary particle table in lattice units."""
    scale = resolution / box_size
    t = np.zeros((len(rows), 10))
    t[:, 0:4] = np.asarray(rows, dtype=float)[:, 2:6] * scale
    return t


def to_particles(rows, resolution: int, box_size: float = BOX_SIZE_M):
    from .psc3d import Particle3D
    return [Particle3D(*r) for r in to_table(rows, resolution, box_size)[:, 0:4].tolist()]


def permeability_probe(rows, case: str, resolution: int, steps: int = 360, tau: float = 0.72,
                       force_z: float = 2.0e-7, subsamples: int = 2, source_dump: str = "",
                       particle_filter: str = "all", geometry_workers: int = 1, precision: str = "fp64",
                       device=None) -> dict:
    """One permeability probe: ``steps`` steps of a static-geometry sparse
    solver from rest on the filtered rows, then the readouts of the
    reference's probe row (mean and superficial u_z over the (1 - solid)
    weight, max u_z, k, load/imb closure, body-force totals)."""
    from .psc3d import SparsePscD3Q19
    rows = filter_particles(np.asarray(rows, dtype=float), particle_filter)
    solver = SparsePscD3Q19(n=resolution, particles=to_particles(rows, resolution), tau=tau, subsamples=subsamples,
                            body_force_z=force_z, static_geometry=True, precision=precision, device=device)
    if steps > 1:
        solver.rebuild_geometry()
        solver.run(steps - 1)
    last = solver.collide_stream_psc() if steps >= 1 else None
    residual = float(last["momentum_residual"]) if last is not None else float("nan")
    n = solver.n
    w = 1.0 - solver.solid
    fluid_volume = float(np.sum(w))
    wuz = solver.uz * w
    mean_uz_fluid = float(np.sum(wuz) / max(fluid_volume, 1.0e-300))
    superficial = float(np.sum(wuz) / n ** 3)
    nu = (tau - 0.5) / 3.0
    fsum = np.sum(solver.hydro_forces, axis=0) if len(solver.hydro_forces) else np.zeros(3)
    imb = np.array([float(np.sum(solver.imb_mx)), float(np.sum(solver.imb_my)), float(np.sum(solver.imb_mz))])
    return {
        "case": case, "resolution": resolution, "steps": steps, "tau": tau, "force_z": force_z,
        "subsamples": subsamples, "geometry_workers": geometry_workers,
        "source_dump": str(Path(source_dump).resolve()) if source_dump else "",
        "particle_filter": particle_filter, "n_particles": len(rows),
        "porosity_psc": fluid_volume / n ** 3, "mean_uz_fluid": mean_uz_fluid, "superficial_uz": superficial,
        "max_uz": float(np.max(solver.uz)), "permeability_lu": float(nu * superficial / max(force_z, 1.0e-300)),
        "K_over_skeleton_same_resolution": "", "momentum_residual": residual,
        "particle_force_sum_x_lu": float(fsum[0]), "particle_force_sum_y_lu": float(fsum[1]),
        "particle_force_sum_z_lu": float(fsum[2]), "particle_force_norm_sum_lu": float(np.linalg.norm(fsum)),
        "fluid_imb_sum_x_lu": float(imb[0]), "fluid_imb_sum_y_lu": float(imb[1]),
        "fluid_imb_sum_z_lu": float(imb[2]),
        "particle_fluid_momentum_closure_lu": float(np.linalg.norm(fsum + imb)),
        "domain_body_force_z_lu": float(force_z * n ** 3),
        "fluid_volume_body_force_z_lu": float(force_z * fluid_volume),
    }


def write_probe(outdir, row) -> Path:
    """``<case>_nXXX.csv`` with the probe row (the convergence input)."""
    from .convergence import SUMMARY_FIELDS, write_rows
    path = Path(outdir) / f"{row['case']}_n{int(row['resolution']):03d}.csv"
    write_rows(path, SUMMARY_FIELDS, [row])
    return path


def windows_from_dumps(paths, resolution: int, particle_filter: str = "all", box_size: float = BOX_SIZE_M):
    """[(resolution, table), ...] of dump files, for BatchedPscD3Q19."""
    return [(resolution, to_table(filter_particles(read_last_dump(p), particle_filter), resolution, box_size))
            for p in paths]
