"""Write the golden fixtures of the PSC step from the reference implementation.

Run where the reference sources are available (the GPU tests only read the fixtures):

    python tests/golden/make_golden.py [/path/to/reference]

It imports ``src/lbm_dem/psc3d.py`` of wangjianfttt/particuology_Hydraulic-penalty-
(NumPy, FP64), drives it on small seeded cases and stores inputs and outputs as
``tests/golden/*.npz``.  Only data is stored: particle tables, parameters and
the resulting arrays.  The cases cover what closed forms cannot pin: the
multi-step trajectories of ``collide_stream_psc`` with every quirk combined,
the sparse geometry lists, dense-vs-sparse loads and a permeability trace.
"""

from __future__ import annotations

import dataclasses
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")


def load_reference():
    sys.path.insert(0, str(REF))
    from src.lbm_dem import psc3d  # noqa: E402  (reference module, read-only)
    return psc3d


def ptable(particles):
    return np.array([[getattr(p, k) for k in ("x", "y", "z", "r", "vx", "vy", "vz", "wx", "wy", "wz")]
                     for p in particles], dtype=np.float64).reshape(-1, 10)


def diag_row(d):
    return [d["solid_volume_lu"], d["mean_uz"], d["momentum_residual"], d["max_force"], d["max_torque"]]


def sample_idx(rng, size, k):
    return np.sort(rng.choice(size, size=min(k, size), replace=False))


def case_smoke12(ref):
    """test_psc3d_smoke.py:6-22 inputs; one step; full state."""
    p = ref.Particle3D(x=6.0, y=6.0, z=6.0, r=2.0)
    s = ref.PscD3Q19(n=12, particles=[p], tau=0.72, subsamples=2, body_force_z=1.0e-7, static_geometry=True)
    d = s.collide_stream_psc()
    return dict(n=12, tau=0.72, subsamples=2, force=np.array([0, 0, 1e-7]), static=1, layout="dense",
                particles=ptable([p]), steps=1, f=s.f, solid=s.solid,
                imb=np.stack([s.imb_mx, s.imb_my, s.imb_mz]), uz=s.uz, ux=s.ux, rho=s.rho,
                forces=s.hydro_forces[None], torques=s.hydro_torques[None], diag=np.array([diag_row(d)]))


def run_traj(ref, solver, particles, steps, move=None, rng_seed=7, nsample=4096):
    forces, torques, diags = [], [], []
    for k in range(steps):
        if move is not None:
            move(particles, k)
        d = solver.collide_stream_psc()
        forces.append(solver.hydro_forces.copy())
        torques.append(solver.hydro_torques.copy())
        diags.append(diag_row(d))
    rng = np.random.default_rng(rng_seed)
    idx = sample_idx(rng, solver.f.size, nsample)
    return dict(forces=np.array(forces), torques=np.array(torques), diag=np.array(diags),
                f_idx=idx, f_val=solver.f.ravel()[idx], f_sum=np.array(solver.f.sum(axis=(1, 2, 3))),
                solid=solver.solid, uz=solver.uz, ux=solver.ux,
                imb=np.stack([solver.imb_mx, solver.imb_my, solver.imb_mz]))


def case_moving20(ref):
    """20^3 dense solver, tau 0.8, s=2, one r=2.5 sphere at a seeded centre
    moving at U=(1e-4,0,0); dynamic geometry; 100 steps."""
    rng = np.random.default_rng(0)
    c = rng.uniform(0.0, 20.0, size=3)
    p = ref.Particle3D(float(c[0]), float(c[1]), float(c[2]), 2.5, vx=1.0e-4)
    s = ref.PscD3Q19(n=20, particles=[p], tau=0.8, subsamples=2, body_force_z=0.0)
    out = run_traj(ref, s, [p], 100)
    out.update(n=20, tau=0.8, subsamples=2, force=np.zeros(3), static=0, layout="dense",
               particles=ptable([p]), steps=100)
    return out


def case_spin16(ref):
    """16^3 dense, centred sphere translating and rotating, Fz, 30 steps."""
    p = ref.Particle3D(8.3, 7.6, 8.1, 3.2, vx=2e-4, vy=-1e-4, vz=5e-5, wx=1e-5, wy=-2e-5, wz=3e-5)
    s = ref.PscD3Q19(n=16, particles=[p], tau=0.9, subsamples=3, body_force_z=3.0e-7)
    out = run_traj(ref, s, [p], 30)
    out.update(n=16, tau=0.9, subsamples=3, force=np.array([0, 0, 3e-7]), static=0, layout="dense",
               particles=ptable([p]), steps=30)
    return out


def _three(ref):
    return [
        ref.Particle3D(1.2, 7.5, 8.0, 2.6, vx=1e-4, wz=2e-5),            # truncated by the clamp
        ref.Particle3D(9.0, 8.4, 7.7, 3.1, vy=-5e-5, wx=1e-5),            # overlaps the next one
        ref.Particle3D(12.5, 9.1, 9.3, 2.2, vz=8e-5, wy=-3e-5),
    ]


def case_sparse16(ref):
    """16^3 sparse solver, three spheres (one clamped at x=0, one overlap pair),
    s=3, tau 0.72, Fz 2e-7, dynamic geometry; the caller moves the spheres in
    place before every step (as run_precision_validation.py:257-258 does)."""
    parts = _three(ref)
    s = ref.SparsePscD3Q19(n=16, particles=parts, tau=0.72, subsamples=3, body_force_z=2.0e-7)
    s.rebuild_geometry()
    ids0 = [np.asarray(a) for a in s.node_ids]
    vals0 = [np.asarray(a) for a in s.alpha_vals]
    start = ptable(parts)

    def move(ps, k):
        for j, q in enumerate(ps):
            q.x += 0.013 * (j + 1)
            q.z -= 0.007 * (j + 1)

    out = run_traj(ref, s, parts, 12, move=move)
    out.update(n=16, tau=0.72, subsamples=3, force=np.array([0, 0, 2e-7]), static=0, layout="sparse",
               particles=start, steps=12, move_dx=np.array([0.013, 0.0, -0.007]),
               ids_len=np.array([len(a) for a in ids0]), ids=np.concatenate(ids0),
               vals=np.concatenate(vals0))
    return out


def case_dense_vs_sparse(ref):
    """Same overlapping geometry through the dense and the sparse classes."""
    res = {}
    for name, cls, kw in (("dense", ref.PscD3Q19, {}), ("sparse", ref.SparsePscD3Q19, {})):
        parts = _three(ref)
        s = cls(n=16, particles=parts, tau=0.72, subsamples=3, body_force_z=2.0e-7, static_geometry=True, **kw)
        for _ in range(5):
            s.collide_stream_psc()
        res[name + "_forces"] = s.hydro_forces.copy()
        res[name + "_torques"] = s.hydro_torques.copy()
        res[name + "_solid"] = s.solid.copy()
    res.update(n=16, tau=0.72, subsamples=3, force=np.array([0, 0, 2e-7]), static=1,
               particles=ptable(_three(ref)), steps=5)
    return res


def case_perm24(ref):
    """24^3 sparse static permeability probe (run_psc_permeability_resolution.py
    semantics): 4 spheres of r=3, s=2, tau 0.72, Fz 2e-7, 300 steps, k every 20."""
    rng = np.random.default_rng(3)
    parts = []
    while len(parts) < 4:
        c = rng.uniform(4.5, 19.5, size=3)
        if all(np.linalg.norm(c - np.array([q.x, q.y, q.z])) > 6.5 for q in parts):
            parts.append(ref.Particle3D(float(c[0]), float(c[1]), float(c[2]), 3.0))
    tau, fz, n = 0.72, 2.0e-7, 24
    s = ref.SparsePscD3Q19(n=n, particles=parts, tau=tau, subsamples=2, body_force_z=fz, static_geometry=True)
    ks = []
    for step in range(300):
        s.collide_stream_psc()
        if (step + 1) % 20 == 0:
            w = 1.0 - s.solid
            sup = float(np.sum(s.uz * w) / n ** 3)
            ks.append((tau - 0.5) / 3.0 * sup / fz)
    rng2 = np.random.default_rng(11)
    idx = sample_idx(rng2, s.f.size, 4096)
    return dict(n=n, tau=tau, subsamples=2, force=np.array([0, 0, fz]), static=1, layout="sparse",
                particles=ptable(parts), steps=300, k_trace=np.array(ks), f_idx=idx, f_val=s.f.ravel()[idx],
                forces=s.hydro_forces.copy(), torques=s.hydro_torques.copy(), solid=s.solid)


def main():
    ref = load_reference()
    for name, fn in (("smoke12", case_smoke12), ("moving20", case_moving20), ("spin16", case_spin16),
                     ("sparse16", case_sparse16), ("dense_vs_sparse16", case_dense_vs_sparse),
                     ("perm24", case_perm24)):
        data = fn(ref)
        np.savez_compressed(HERE / f"{name}.npz", **{k: np.asarray(v) for k, v in data.items()})
        print("wrote", name, sorted(data))


if __name__ == "__main__":
    main()
