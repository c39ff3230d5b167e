"""DEM coupling adapter on the GPU solver (GPU).

* ``converge`` with check_every=1 is the reference's lbm_converge loop
  (couple_psc_liggghts.py:284-296): same step count and residual as that loop
  run on the oracle, and the same loads;
* check_every=K stops within K steps of it, on a state equal to stepping that
  many steps directly;
* a three-exchange loop against an in-process fake DEM: particle list
  reassignment with a changing particle count, SI load transfer to fines only.
"""

import math

import numpy as np
import pytest

from oracle import psc_oracle as O
from tests.conftest import load_package

pytestmark = pytest.mark.gpu

N, TAU, S, FZ = 16, 0.72, 2, 2.0e-5
SPHERES = [(5.3, 6.1, 7.2, 3.5), (11.2, 10.4, 8.9, 3.0), (8.4, 12.6, 3.1, 1.4)]


def _solver(psc3d, precision="fp64"):
    parts = [psc3d.Particle3D(*s) for s in SPHERES]
    return psc3d.SparsePscD3Q19(N, parts, tau=TAU, subsamples=S, body_force_z=FZ, precision=precision,
                                device="cuda:0")


def _oracle_converge(min_steps, max_steps, tol):
    """The reference's loop (couple_psc_liggghts.py:284-296) on the oracle."""
    ora = O.OracleSolver(N, [O.Sphere(*s) for s in SPHERES], tau=TAU, subsamples=S, body_force=(0.0, 0.0, FZ),
                         layout="sparse")
    prev, residual = None, math.inf
    for it in range(1, max_steps + 1):
        d = ora.collide_stream_psc()
        cur = d["mean_uz"]
        if prev is not None:
            residual = abs(cur - prev) / max(abs(cur), 1.0e-30)
        prev = cur
        if it >= min_steps and residual < tol:
            return it, residual, ora
    return max_steps, residual, ora


def test_converge_is_the_reference_loop():
    psc3d = load_package("psc3d")
    cp = load_package("coupling")
    print("[converge K=1] 16^3, 3 spheres", flush=True)
    sol = _solver(psc3d)
    it, res, ok, diag = cp.converge(sol, min_steps=20, max_steps=600, tol=5e-3)
    it_ref, res_ref, ora = _oracle_converge(20, 600, 5e-3)
    print(f"  converged after {it} steps (oracle {it_ref}), residual {res:.3e}")
    assert ok and it == it_ref
    assert abs(res - res_ref) <= 1e-9 * abs(res_ref)
    assert np.max(np.abs(sol.hydro_forces - ora.hydro_forces)) <= 1e-10 * np.max(np.abs(ora.hydro_forces))
    assert np.array_equal(sol.f, ora.f)


def test_converge_every_k_steps():
    psc3d = load_package("psc3d")
    cp = load_package("coupling")
    print("[converge K=8] 16^3", flush=True)
    it1, _, _, _ = cp.converge(_solver(psc3d), min_steps=20, max_steps=600, tol=5e-3)
    sol = _solver(psc3d)
    itk, res, ok, diag = cp.converge(sol, min_steps=20, max_steps=600, tol=5e-3, check_every=8)
    assert ok and it1 <= itk < it1 + 8 and res < 5e-3
    ref = _solver(psc3d)
    for _ in range(itk):
        ref.collide_stream_psc()
    assert np.array_equal(sol.f, ref.f)
    assert diag["mean_uz"] == ref.mean_uz()


class FakeDEM:
    """Explicit-Euler point particles in a 9 mm periodic box (SI units)."""

    def __init__(self, seed=0):
        rng = np.random.default_rng(seed)
        self.ids = np.array([11, 12, 13, 14, 15])
        self.x = np.array([[2.6e-3, 3.0e-3, 3.6e-3], [5.6e-3, 5.2e-3, 4.5e-3], [4.2e-3, 6.3e-3, 1.5e-3],
                           [7.4e-3, 2.2e-3, 6.6e-3], [1.2e-3, 7.5e-3, 7.1e-3]])
        self.r = np.array([1.7e-3, 1.5e-3, 2.7e-4, 2.6e-4, 2.9e-4])
        self.v = rng.normal(0.0, 1e-4, (5, 3))
        self.w = np.zeros((5, 3))
        self.drag = np.zeros((5, 3))
        self.dt = 1e-6

    def state(self):
        cp = load_package("coupling")
        return cp.ParticleState(ids=self.ids.copy(), x=self.x.copy(), v=self.v.copy(), radius=self.r.copy(),
                                omega=self.w.copy())

    def set_dragforces(self, ids, forces):
        for a, f in zip(ids, forces):
            self.drag[np.nonzero(self.ids == a)[0][0]] = f

    def run(self, n):
        m = 4.0 / 3.0 * np.pi * self.r ** 3 * 2500.0
        self.v += n * self.dt * self.drag / m[:, None]
        self.x = np.mod(self.x + n * self.dt * self.v, 0.009)

    def drop(self, atom_id):
        keep = self.ids != atom_id
        for k in ("ids", "x", "r", "v", "w", "drag"):
            setattr(self, k, getattr(self, k)[keep])


def test_exchange_loop_with_fake_dem():
    psc3d = load_package("psc3d")
    cp = load_package("coupling")
    print("[exchange loop] fake DEM, 3 macro steps", flush=True)
    res = 32
    sc = cp.si_scaling(res, TAU)
    dem = FakeDEM()
    solver = None
    for macro in range(3):
        if macro == 2:
            dem.drop(14)  # a fine leaves the window
        st = dem.state()
        fines = cp.fine_mask(st)
        parts = cp.psc_particles_from_state(st, res, sc["velocity_lu_per_m_s"], sc["angular_lu_per_rad_s"])
        if solver is None:
            solver = psc3d.SparsePscD3Q19(res, parts, tau=TAU, subsamples=2, body_force_z=FZ, device="cuda:0")
        else:
            cp.update_solver(solver, parts)
        it, r, ok, diag = cp.converge(solver, min_steps=10, max_steps=60, tol=1e-12, check_every=16)
        assert it == 60 and not ok
        forces, torques = cp.transfer_loads(solver, fines, sc["force_scale_N_per_lu"], sc["torque_scale_Nm_per_lu"])
        assert forces.shape == (len(st.ids), 3) and solver.hydro_forces.shape == (len(st.ids), 3)
        assert np.all(forces[~fines] == 0.0) and np.all(np.any(forces[fines] != 0.0, axis=1))
        assert np.array_equal(forces[fines], sc["force_scale_N_per_lu"] * solver.hydro_forces[fines])
        dem.set_dragforces(st.ids, forces)
        dem.run(10)
        porosity = 1.0 - diag["solid_volume_lu"] / res ** 3
        assert 0.0 < porosity < 1.0
