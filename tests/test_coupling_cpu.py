"""DEM coupling adapter, host side (no GPU): SI map invariants, state
conversion, and the vectorised load transfer against an in-process fake of the
LIGGGHTS library (the real libliggghts.so is absent)."""

from __future__ import annotations

import ctypes
import math
import re
from types import SimpleNamespace

import numpy as np
import pytest

from tests.conftest import load_package

cp = load_package("coupling")


def test_si_scaling_invariants():
    s = cp.si_scaling(40, 0.72)
    # ideal gas, lattice viscosity mapped onto the physical one, unit chain
    assert math.isclose(2.0e5, s["rho_he_kg_m3"] * 2077.1 * 773.15, rel_tol=1e-15)
    assert math.isclose(s["nu_lu"] * s["dx_m"] ** 2 / s["dt_s"], s["nu_phys_m2_s"], rel_tol=1e-14)
    assert math.isclose(s["nu_lu"], (0.72 - 0.5) / 3.0, rel_tol=1e-15)
    assert math.isclose(s["dx_m"], 0.009 / 40, rel_tol=0)
    assert math.isclose(s["u_scale_m_s"] * s["velocity_lu_per_m_s"], 1.0, rel_tol=1e-15)
    assert s["angular_lu_per_rad_s"] == s["dt_s"]
    # a lattice force of 1 is rho dx^3 (mass of a cell) times dx/dt^2 (acceleration)
    assert math.isclose(s["force_scale_N_per_lu"], s["rho_he_kg_m3"] * s["dx_m"] ** 3 * s["dx_m"] / s["dt_s"] ** 2,
                        rel_tol=1e-14)
    assert math.isclose(s["torque_scale_Nm_per_lu"], s["force_scale_N_per_lu"] * s["dx_m"], rel_tol=1e-15)
    # doubling the resolution: dx halves, dt quarters, force scale is unchanged times 2^-4 * 2^4... = 1
    s2 = cp.si_scaling(80, 0.72)
    assert math.isclose(s2["dt_s"], s["dt_s"] / 4.0, rel_tol=1e-14)
    assert math.isclose(s2["force_scale_N_per_lu"], s["force_scale_N_per_lu"], rel_tol=1e-14)


def _args(use):
    return SimpleNamespace(resolution=40, tau=0.72, temperature_k=773.15, pressure_pa=2.0e5, mu_he=3.7e-5,
                           use_si_scaling=use, psc_force_scale=1.0, psc_torque_scale=1.0, velocity_lu_scale=0.0,
                           angular_lu_scale=0.0, velocity_feedback_factor=1.0, angular_feedback_factor=1.0)


def test_apply_si_scaling_in_place():
    a = _args(True)
    out = cp.apply_si_scaling(a)
    s = cp.si_scaling(40, 0.72)
    assert a.psc_force_scale == s["force_scale_N_per_lu"] == out["applied_force_scale"]
    assert a.velocity_lu_scale == s["velocity_lu_per_m_s"] == out["velocity_lu_per_m_s"]
    b = _args(False)
    out = cp.apply_si_scaling(b)
    assert b.psc_force_scale == 1.0 and out["velocity_lu_per_m_s"] == 0.0 and out["applied_force_scale"] == 1.0


def _state(P=7, seed=0):
    rng = np.random.default_rng(seed)
    r = np.where(rng.random(P) < 0.5, rng.uniform(1e-4, 2.9e-4, P), rng.uniform(1.0e-3, 1.5e-3, P))
    return cp.ParticleState(ids=rng.permutation(np.arange(1, P + 1)), x=rng.uniform(0, 0.009, (P, 3)),
                            v=rng.normal(0, 1e-3, (P, 3)), radius=r, omega=rng.normal(0, 1.0, (P, 3)))


def test_state_conversion():
    st = _state()
    rows = cp.particle_rows(st)
    assert np.array_equal(rows[:, 1] == 2.0, st.radius < 3e-4)
    assert np.array_equal(rows[:, 0], st.ids.astype(float))
    t = cp.particle_table(st, 40, 2.5, 0.5)
    assert np.allclose(t[:, 0:3] * 0.009 / 40, st.x, rtol=1e-15, atol=0)
    assert np.allclose(t[:, 3] * 0.009 / 40, st.radius, rtol=1e-15, atol=0)
    assert np.array_equal(t[:, 4:7], st.v * 2.5) and np.array_equal(t[:, 7:10], st.omega * 0.5)


# ---------------------------------------------------------------------------
# an in-process stand-in for libliggghts.so: per-atom arrays in a shuffled
# local order, the C entry points the adapter uses, and `set atom` parsing
# ---------------------------------------------------------------------------
class FakeLib:
    def __init__(self, ids, with_fix=True, with_string=True):
        self.ids = np.ascontiguousarray(np.asarray(ids, dtype=np.int32))
        n = len(self.ids)
        self.arrays = {"dragforce": np.zeros((n, 3)), "hdtorque": np.zeros((n, 3))}
        self._rows = {}
        for name, a in self.arrays.items():
            rows = (ctypes.POINTER(ctypes.c_double) * max(1, n))()
            for k in range(n):
                rows[k] = ctypes.cast(a.ctypes.data + 24 * k, ctypes.POINTER(ctypes.c_double))
            self._rows[name] = rows
        self.commands = 0
        self.lammps_extract_atom = lambda h, name: self.ids.ctypes.data if name == b"id" else None
        if with_fix:
            self.lammps_extract_fix = lambda h, name, style, typ, i, j: (
                ctypes.addressof(self._rows[name.decode()]) if style == 1 and typ == 2 else None)
        if with_string:
            self.lammps_commands_string = lambda h, text: [self.run(line) for line in text.decode().splitlines()]

    def run(self, line):
        m = re.match(r"set atom (\d+) property/atom (\w+) (\S+) (\S+) (\S+)", line)
        k = int(np.nonzero(self.ids == int(m.group(1)))[0][0])
        self.arrays[m.group(2)][k] = [float(m.group(i)) for i in (3, 4, 5)]
        self.commands += 1


class FakeLmp:
    def __init__(self, lib):
        self.lib = lib
        self.handle = ctypes.c_void_p(1)

    def natoms(self):
        return len(self.lib.ids)

    def command(self, text):
        self.lib.run(text)


@pytest.mark.parametrize("mode", ["array", "commands_string", "command"])
def test_vector_force_coupler_matches_per_atom_commands(mode):
    rng = np.random.default_rng(1)
    ids = rng.permutation(np.arange(10, 30))
    local = rng.permutation(ids)  # the library's own local order
    forces = rng.normal(size=(len(ids), 3))
    torques = rng.normal(size=(len(ids), 3))
    lib = FakeLib(local, with_fix=(mode == "array"), with_string=(mode != "command"))
    c = cp.VectorForceCoupler(FakeLmp(lib))
    c.set_dragforces(ids, forces)
    c.set_hdtorques(ids, torques)
    assert c.mode == mode
    # the reference's path: one `set atom` command per particle
    ref = FakeLib(local, with_fix=False, with_string=False)
    for a, f in zip(ids, forces):
        ref.run(f"set atom {int(a)} property/atom dragforce {f[0]:.16e} {f[1]:.16e} {f[2]:.16e}")
    for a, t in zip(ids, torques):
        ref.run(f"set atom {int(a)} property/atom hdtorque {t[0]:.16e} {t[1]:.16e} {t[2]:.16e}")
    assert np.array_equal(lib.arrays["dragforce"], ref.arrays["dragforce"])
    assert np.array_equal(lib.arrays["hdtorque"], ref.arrays["hdtorque"])
    if mode == "array":
        assert lib.commands == 0


def test_vector_force_coupler_leaves_unlisted_atoms():
    lib = FakeLib(np.array([5, 3, 9, 7]))
    lib.arrays["dragforce"][:] = -1.0
    cp.VectorForceCoupler(FakeLmp(lib)).set_dragforces(np.array([9, 3]), np.array([[1.0, 2, 3], [4, 5, 6]]))
    assert np.array_equal(lib.arrays["dragforce"], [[-1, -1, -1], [4, 5, 6], [1, 2, 3], [-1, -1, -1]])


class FakeSolver:
    """A solver whose mean u_z relaxes as 1 - exp(-t/50): enough to drive the
    convergence loops without a GPU."""

    def __init__(self):
        self.t = 0
        self.static_geometry = False
        self.rebuilds = 0
        self.readouts = 0

    def _m(self):
        return 1.0 - math.exp(-self.t / 50.0)

    def rebuild_geometry(self):
        self.rebuilds += 1

    def collide_stream_psc(self):
        self.t += 1
        self.readouts += 1
        return {"mean_uz": self._m(), "solid_volume_lu": 0.0}

    def run(self, n):
        self.t += int(n)

    def mean_uz(self):
        return self._m()


def _reference_loop(min_steps, max_steps, tol):
    """couple_psc_liggghts.py:284-296 on the fake solver."""
    s = FakeSolver()
    prev, residual = None, math.inf
    for it in range(1, max_steps + 1):
        cur = s.collide_stream_psc()["mean_uz"]
        if prev is not None:
            residual = abs(cur - prev) / max(abs(cur), 1.0e-30)
        prev = cur
        if it >= min_steps and residual < tol:
            return it, residual, True
    return max_steps, residual, False


@pytest.mark.parametrize("K", [1, 2, 5, 16, 64])
@pytest.mark.parametrize("tol", [1e-3, 1e-4, 1e-12])
def test_converge_block_logic(K, tol):
    it_ref, res_ref, ok_ref = _reference_loop(20, 600, tol)
    s = FakeSolver()
    it, res, ok, diag = cp.converge(s, min_steps=20, max_steps=600, tol=tol, check_every=K)
    assert s.static_geometry is False and s.rebuilds == 1  # voxelized once, flag restored
    assert ok == ok_ref and s.t == it
    if K == 1:
        assert (it, res) == (it_ref, res_ref)
    elif ok:
        assert it_ref <= it < it_ref + K and res < tol
        assert s.readouts <= 2 * (-(-it // K)) + 2  # two readouts per block of K
    else:
        assert it == 600


def test_transient_advance_steps_and_residual():
    s = FakeSolver()
    s.t = 10
    prev = s.mean_uz()
    n, res, ok, diag = cp.transient_advance(s, 7)
    assert n == 7 and ok and s.t == 17 and s.readouts == 1
    assert res == abs(diag["mean_uz"] - prev) / abs(diag["mean_uz"])
