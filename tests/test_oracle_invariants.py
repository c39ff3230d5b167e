"""Closed forms and invariants that pin the oracle independently of the
reference (CPU).  Each check is chosen so that a plausible slip (dropped
half-force, wrong opposite index, transposed axis, off-by-one sample offset)
fails it."""

from fractions import Fraction
import math

import numpy as np
import pytest

from oracle import psc_oracle as O


# ---------------------------------------------------------------- a1 lattice
def test_lattice_moments():
    c = O.C.astype(float)
    assert O.W.sum() == pytest.approx(1.0, abs=1e-15)
    np.testing.assert_allclose(O.W @ c, 0.0, atol=1e-16)
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", O.W, c, c), O.CS2 * np.eye(3), atol=1e-16)
    # fourth moment isotropy of D3Q19: sum w c_a^2 c_b^2 = cs^4 (1 + 2 delta_ab)
    m4 = np.einsum("i,ia,ib->ab", O.W, c * c, c * c)
    np.testing.assert_allclose(m4, O.CS2 ** 2 * (np.ones((3, 3)) + 2 * np.eye(3)), atol=1e-15)
    assert np.array_equal(O.OPP[O.OPP], np.arange(19))
    assert np.array_equal(O.C[O.OPP], -O.C)
    assert len({tuple(v) for v in O.C}) == 19


# ------------------------------------------------------- a6/a7 eq and forcing
def test_equilibrium_moments():
    rng = np.random.default_rng(1)
    rho = 1.0 + 0.01 * rng.standard_normal(50)
    u = 0.02 * rng.standard_normal((3, 50))
    feq = O.equilibrium(rho, *u)
    c = O.C.astype(float)
    np.testing.assert_allclose(feq.sum(0), rho, rtol=1e-14)
    np.testing.assert_allclose(np.einsum("ia,in->an", c, feq), rho * u, rtol=1e-12, atol=1e-17)
    pi = np.einsum("ia,ib,in->abn", c, c, feq)
    expect = rho * (O.CS2 * np.eye(3)[:, :, None] + u[:, None, :] * u[None, :, :])
    np.testing.assert_allclose(pi, expect, rtol=1e-12, atol=1e-16)


def test_guo_moments():
    rng = np.random.default_rng(2)
    u = 0.02 * rng.standard_normal((3, 40))
    F = (3e-6, -1e-6, 2e-6)
    tau = 0.8
    g = O.guo_force(*F, *u, tau)
    c = O.C.astype(float)
    np.testing.assert_allclose(g.sum(0), 0.0, atol=1e-20)
    mom = np.einsum("ia,in->an", c, g)
    np.testing.assert_allclose(mom, (1 - 0.5 / tau) * np.array(F)[:, None] * np.ones((1, 40)), rtol=1e-12)


# ------------------------------------------------------------------ a8 step
def test_rest_state_is_fixed_point():
    s = O.OracleSolver(8, [], tau=0.72, subsamples=2, static_geometry=True)
    f0 = s.f.copy()
    for _ in range(3):
        s.collide_stream_psc()
    np.testing.assert_allclose(s.f, f0, rtol=0, atol=1e-16)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_uniform_body_force_momentum_double_half_shift(axis):
    """No solids: every step adds F (1 + 1/(2 tau)) of momentum per cell --
    the reference's two half-force shifts (u carries F/2, u_eff another F/2)
    give F/tau through the BGK relaxation plus (1 - 1/(2 tau)) F from Guo."""
    n, tau, t = 6, 0.8, 7
    F = [0.0, 0.0, 0.0]
    F[axis] = 1e-5
    s = O.OracleSolver(n, [], tau=tau, subsamples=2, body_force=F, static_geometry=True)
    for _ in range(t):
        s.collide_stream_psc()
    j = O.axis_moment(s.f, axis).sum()
    assert j == pytest.approx(t * n ** 3 * F[axis] * (1 + 0.5 / tau), rel=1e-10)
    # the stored velocity is (j + F/2)/rho: uniform field
    u = [s.ux, s.uy, s.uz][axis]
    rho = s.rho
    np.testing.assert_allclose(u * rho, t * F[axis] * (1 + 0.5 / tau) + 0.5 * F[axis], rtol=1e-9)


def test_streaming_is_a_permutation():
    rng = np.random.default_rng(3)
    f = rng.random((19, 5, 6, 7))
    g = O.push_stream(f)
    for i in range(19):
        assert np.array_equal(np.sort(g[i].ravel()), np.sort(f[i].ravel()))
        cx, cy, cz = O.C[i]
        assert g[i][(1 + cx) % 5, (2 + cy) % 6, (3 + cz) % 7] == f[i][1, 2, 3]


def test_fully_solid_cell_collision():
    """alpha = 1: B = 1/(1+1e-14); the post-collision state is
    f + (1-B) Omega_bgk + B (f_opp - f + feq_s - feq_f,opp) + Guo."""
    n, tau = 8, 0.72
    s = O.OracleSolver(n, [O.Sphere(4.0, 4.0, 4.0, 3.0)], tau=tau, subsamples=2, static_geometry=True,
                       layout="dense")
    rng = np.random.default_rng(4)
    s.f = s.f + 1e-4 * rng.standard_normal(s.f.shape)
    s.rebuild_geometry()
    assert s.solid[4, 4, 4] == 1.0
    f = s.f[:, 4, 4, 4].copy()
    rho = f.sum()
    u = O.C.T.astype(float) @ f / rho
    feq_f = O.equilibrium(np.array(rho), *u)
    feq_s = O.equilibrium(np.array(rho), 0.0, 0.0, 0.0)
    B = 1.0 / (1.0 + 1e-14)
    post = f + (1 - B) * (-(f - feq_f) / tau) + B * (f[O.OPP] - f + feq_s - feq_f[O.OPP])
    s.collide_stream_psc()
    # read the post-collision value back from the streamed destination
    got = np.array([s.f[i][(4 + O.C[i, 0]) % n, (4 + O.C[i, 1]) % n, (4 + O.C[i, 2]) % n] for i in range(19)])
    np.testing.assert_allclose(got, post, rtol=1e-13, atol=1e-19)


def test_mirror_symmetry_of_moving_sphere():
    """A centred sphere translating along x: no transverse force, no torque."""
    n = 12
    s = O.OracleSolver(n, [O.Sphere(6.0, 6.0, 6.0, 2.7, vx=1e-4)], tau=0.8, subsamples=3, layout="dense")
    for _ in range(10):
        s.collide_stream_psc()
    F = s.hydro_forces[0]
    T = s.hydro_torques[0]
    assert F[0] < 0  # resists the motion
    assert abs(F[1]) < 1e-12 * abs(F[0]) and abs(F[2]) < 1e-12 * abs(F[0])
    assert np.max(np.abs(T)) < 1e-10 * abs(F[0])
    assert s.momentum_residual < 1e-12 * abs(F[0])


def test_axis_permutation_equivariance():
    """Force and motion along x on a geometry equal the z case with x and z
    swapped (pins the vector body-force generalisation)."""
    n = 10
    pz = O.Sphere(4.3, 5.1, 5.6, 2.2, vz=2e-4, wx=1e-5)
    px = O.Sphere(5.6, 5.1, 4.3, 2.2, vx=2e-4, wz=-1e-5)  # omega is a pseudo-vector
    sz = O.OracleSolver(n, [pz], tau=0.9, subsamples=2, body_force=(0, 0, 3e-6), layout="sparse")
    sx = O.OracleSolver(n, [px], tau=0.9, subsamples=2, body_force=(3e-6, 0, 0), layout="sparse")
    for _ in range(6):
        sz.collide_stream_psc()
        sx.collide_stream_psc()
    np.testing.assert_allclose(sx.hydro_forces[0][::-1], sz.hydro_forces[0], rtol=1e-9, atol=1e-20)
    # torque is a pseudo-vector: swapping x and z flips its sign
    np.testing.assert_allclose(-sx.hydro_torques[0][::-1], sz.hydro_torques[0], rtol=1e-8, atol=1e-20)
    np.testing.assert_allclose(np.transpose(sx.ux, (2, 1, 0)), sz.uz, rtol=1e-9, atol=1e-20)


# ------------------------------------------------------------ a3 voxelization
def _exact_count(ix, iy, iz, p, s):
    """Exact rational count (ties resolved exactly)."""
    r2 = Fraction(p.r) ** 2
    cnt = 0
    for a in range(s):
        for b in range(s):
            for c in range(s):
                dx = ix + Fraction(2 * a + 1, 2 * s) - Fraction(p.x)
                dy = iy + Fraction(2 * b + 1, 2 * s) - Fraction(p.y)
                dz = iz + Fraction(2 * c + 1, 2 * s) - Fraction(p.z)
                cnt += (dx * dx + dy * dy + dz * dz) <= r2
    return cnt


def test_voxel_counts_match_exact_arithmetic_away_from_ties():
    rng = np.random.default_rng(6)
    for trial in range(6):
        p = O.Sphere(*rng.uniform(3.0, 7.0, 3), rng.uniform(0.8, 2.6))
        s = int(rng.integers(1, 5))
        ix, iy, iz, counts = O.sphere_counts(p, 11, s)
        for a, x in enumerate(ix):
            for b, y in enumerate(iy):
                for c, z in enumerate(iz):
                    assert counts[a, b, c] == _exact_count(int(x), int(y), int(z), p, s)


def test_voxel_volume_converges_and_interior_is_full():
    p = O.Sphere(10.37, 9.81, 10.12, 5.3)
    vols = []
    for s in (2, 4, 8):
        ids, vals, _ = O.sparse_alpha(p, 21, s)
        vols.append(vals.sum())
    exact = 4 / 3 * math.pi * p.r ** 3
    errs = [abs(v - exact) / exact for v in vols]
    assert errs[-1] < 0.01 and errs[-1] < errs[0]
    ix, iy, iz, counts = O.sphere_counts(p, 21, 3)
    X, Y, Z = np.meshgrid(ix, iy, iz, indexing="ij")
    far = np.sqrt((np.maximum(np.abs(X - p.x), np.abs(X + 1 - p.x))) ** 2
                  + (np.maximum(np.abs(Y - p.y), np.abs(Y + 1 - p.y))) ** 2
                  + (np.maximum(np.abs(Z - p.z), np.abs(Z + 1 - p.z))) ** 2)
    assert np.all(counts[far < p.r] == 27)


def test_clamped_box_truncates_spheres():
    p = O.Sphere(0.4, 5.0, 5.0, 2.0)
    lo = O.sphere_box(p, 10)
    assert lo[0] == 0 and lo[1] == 4
    ids, vals, _ = O.sparse_alpha(p, 10, 2)
    # periodic boxes wrap the same sphere onto x = 7..9
    ids_p, vals_p, pos = O.sparse_alpha(p, 10, 2, periodic=True)
    assert vals_p.sum() > vals.sum()
    assert np.any(pos[0] < 0)


# ---------------------------------------------------- a5 solid velocity, a10
def test_solid_velocity_rigid_body():
    p = O.Sphere(6.2, 5.7, 6.4, 2.9, vx=1e-4, vy=-2e-4, vz=3e-4, wx=1e-5, wy=2e-5, wz=-3e-5)
    s = O.OracleSolver(12, [p], subsamples=2, layout="sparse")
    s.rebuild_geometry()
    us = s.solid_velocity()
    X, Y, Z = np.indices((12, 12, 12)) + 0.5
    r = np.stack([X - p.x, Y - p.y, Z - p.z])
    w = np.array([p.wx, p.wy, p.wz])[:, None, None, None]
    v = np.array([p.vx, p.vy, p.vz])[:, None, None, None] + np.cross(w, r, axis=0)
    mask = s.solid > 0
    for a in range(3):
        np.testing.assert_allclose(us[a][mask], v[a][mask], rtol=1e-12, atol=1e-18)
        assert np.all(us[a][~mask] == 0.0)


def test_load_closure_and_owner_split():
    parts = [O.Sphere(5.0, 6.0, 6.0, 2.5, vx=1e-4), O.Sphere(8.5, 6.2, 5.9, 2.5, vy=-1e-4)]
    s = O.OracleSolver(14, parts, tau=0.8, subsamples=3, body_force=(0, 0, 1e-6), layout="sparse")
    for _ in range(4):
        s.collide_stream_psc()
    total = s.hydro_forces.sum(axis=0)
    fluid = np.array([s.imb_mx.sum(), s.imb_my.sum(), s.imb_mz.sum()])
    np.testing.assert_allclose(total, -fluid, rtol=1e-12, atol=1e-22)
    assert s.momentum_residual < 1e-13 * np.max(np.abs(total))


# ---------------------------------------------------------------- a9 readouts
def test_published_permeability_reproduced_from_its_csv_columns():
    """convergence/permeability/baseline_n128.csv of the reference: k_lu from
    superficial_uz, tau 0.72 and Fz 2e-7 (values copied as data)."""
    import json
    import os
    row = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published_k.json")))
    nu = (row["tau"] - 0.5) / 3.0
    assert nu * row["superficial_uz"] / row["force_z"] == pytest.approx(row["permeability_lu"], rel=1e-15)


def test_hasimoto_small_lambda():
    assert O.hasimoto_factor(1e-6, 1.0) == pytest.approx(1.0 + 2.837297e-6, rel=1e-9)


# ------------------------------------------------------- periodic geometry mode
def _periodic_run(centre, shift, layout, steps=3, n=16):
    """A rotating, translating sphere in periodic mode with a body force,
    started from a seeded perturbed state; the whole system is translated by
    the integer vector ``shift`` (centre and initial state)."""
    rng = np.random.default_rng(8)
    f0 = O.equilibrium(np.ones((n,) * 3), *np.zeros((3,) + (n,) * 3)) + 1e-5 * rng.standard_normal((19,) + (n,) * 3)
    c = np.mod(np.asarray(centre) + np.asarray(shift), n)
    p = O.Sphere(*c, 3.3, vx=2e-4, vy=-1e-4, vz=1.5e-4, wx=3e-5, wy=-2e-5, wz=4e-5)
    s = O.OracleSolver(n, [p], tau=0.8, subsamples=3, body_force=(0.0, 2e-6, 1e-6), layout=layout,
                       periodic_geometry=True)
    s.f = np.roll(f0, tuple(int(v) for v in shift), axis=(1, 2, 3))
    for _ in range(steps):
        s.collide_stream_psc()
    return s


@pytest.mark.parametrize("layout", ["sparse", "dense"])
@pytest.mark.parametrize("shift", [(9, 0, 0), (0, -8, 0), (3, 5, 9)])
def test_periodic_mode_is_translation_invariant(layout, shift):
    """Periodic geometry: translating the system by whole cells so that the
    sphere straddles the boundary rolls alpha and f with it and leaves the
    force and torque unchanged.  A box wrapped wrongly changes alpha; using
    the wrapped instead of the unwrapped cell position as lever arm changes
    the solid velocity (hence f) and the torque by O(n)."""
    centre = (8.31, 7.62, 8.93)
    a = _periodic_run(centre, (0, 0, 0), layout)
    b = _periodic_run(centre, shift, layout)
    p = b.particles[0]
    lo = O.sphere_box(p, b.n, periodic=True)
    assert min(lo[0], lo[2], lo[4]) < 0 or max(lo[1], lo[3], lo[5]) > b.n  # the shifted box wraps
    sh = tuple(int(v) for v in shift)
    assert np.array_equal(b.solid, np.roll(a.solid, sh, axis=(0, 1, 2)))
    np.testing.assert_allclose(b.f, np.roll(a.f, sh, axis=(1, 2, 3)), rtol=0, atol=1e-17)
    scale = np.max(np.abs(a.hydro_forces))
    assert scale > 1e-9
    np.testing.assert_allclose(b.hydro_forces, a.hydro_forces, rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(b.hydro_torques, a.hydro_torques, rtol=0, atol=1e-11 * scale * 4.3)
    assert b.momentum_residual <= 1e-12 * scale


def test_periodic_solid_velocity_uses_the_unwrapped_lever_arm():
    """u_s = U + omega x r on every covered cell of a sphere that wraps around
    the x and z faces, with r the offset of the cell's unwrapped position
    (the wrapped one differs by n)."""
    n = 12
    p = O.Sphere(0.7, 6.2, 11.4, 2.9, vx=1e-4, vy=-2e-4, vz=3e-4, wx=1e-5, wy=2e-5, wz=-3e-5)
    for layout in ("sparse", "dense"):
        s = O.OracleSolver(n, [p], subsamples=2, layout=layout, periodic_geometry=True)
        s.rebuild_geometry()
        us = s.solid_velocity()
        X, Y, Z = np.indices((n,) * 3) + 0.5
        r = [X - p.x, Y - p.y, Z - p.z]
        r = np.stack([v - n * np.round(v / n) for v in r])  # nearest image of the cell centre
        w = np.array([p.wx, p.wy, p.wz])[:, None, None, None]
        v = np.array([p.vx, p.vy, p.vz])[:, None, None, None] + np.cross(w, r, axis=0)
        mask = s.solid > 0
        assert mask[0].any() and mask[-1].any() and mask[:, :, 0].any()  # covers cells on both sides
        for a in range(3):
            np.testing.assert_allclose(us[a][mask], v[a][mask], rtol=1e-12, atol=1e-18)


def test_periodic_dense_and_sparse_layouts_agree_on_wrapping_overlaps():
    """The two layouts implement periodic mode separately (nearest-image
    offsets over the full lattice vs unwrapped box coordinates): on
    overlapping, rotating spheres that wrap, loads and state agree to
    round-off, with the force closure of both."""
    n = 14
    parts = [O.Sphere(0.9, 7.1, 6.6, 3.1, vx=1e-4, wz=2e-5), O.Sphere(12.2, 7.4, 6.9, 2.6, vy=-1e-4, wx=-3e-5),
             O.Sphere(6.3, 13.5, 0.4, 2.2, vz=2e-4)]
    res = {}
    for layout in ("sparse", "dense"):
        s = O.OracleSolver(n, [O.Sphere(**vars(p)) for p in parts], tau=0.9, subsamples=3,
                           body_force=(1e-6, 0, 0), layout=layout, periodic_geometry=True)
        for _ in range(3):
            s.collide_stream_psc()
        res[layout] = s
        scale = np.max(np.abs(s.hydro_forces))
        assert s.momentum_residual <= 1e-12 * scale
    a, b = res["sparse"], res["dense"]
    assert np.max(a.alpha_sum) > 1.0  # the pair overlaps
    np.testing.assert_allclose(a.f, b.f, rtol=0, atol=1e-16)
    scale = np.max(np.abs(a.hydro_forces))
    np.testing.assert_allclose(a.hydro_forces, b.hydro_forces, rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(a.hydro_torques, b.hydro_torques, rtol=0, atol=1e-10 * scale * 4)


# ------------------------------------------------------------ readout closed forms
def test_fluid_statistics_closed_forms():
    """(intrinsic, superficial, porosity) of run_precision_validation.py:142-147:
    a uniform velocity U gives intrinsic = U and superficial = porosity * U;
    for any field superficial = porosity * intrinsic; solid is clipped to
    [0, 1] first (overlaps above 1 count as fully solid)."""
    rng = np.random.default_rng(9)
    n = 6
    solid = rng.uniform(-0.2, 1.4, (n,) * 3)
    w = 1.0 - np.clip(solid, 0.0, 1.0)
    por = float(w.sum()) / n ** 3
    assert 0.2 < por < 0.8
    U = 3.7e-5
    i, s, p = O.fluid_statistics(solid, np.full((n,) * 3, U), n)
    assert p == pytest.approx(por, rel=1e-15)
    assert i == pytest.approx(U, rel=1e-14)
    assert s == pytest.approx(por * U, rel=1e-14)
    u = rng.standard_normal((n,) * 3)
    i, s, p = O.fluid_statistics(solid, u, n)
    assert s == pytest.approx(p * i, rel=1e-12)
    assert i == pytest.approx(float((w * u).sum() / w.sum()), rel=1e-13)
    assert O.fluid_weighted_mean(solid, u) == pytest.approx(i, rel=1e-14)
    # an all-solid lattice: no fluid, the intrinsic mean is 0 (the reference's 1e-30 guard)
    assert O.fluid_statistics(np.ones((n,) * 3), u, n) == (0.0, 0.0, 0.0)
    assert O.fluid_weighted_mean(np.full((n,) * 3, 2.0), u) == 0.0


def test_permeability_readout_closed_form():
    """k = nu sum((1 - solid) u) / n^3 / F with nu = (tau - 1/2)/3
    (run_psc_permeability_resolution.py:144-150): a uniform flow U through a
    lattice without solids gives k = nu U / F; half-solid cells count half."""
    n, tau, F, U = 5, 0.8, 2e-7, 4e-4
    nu = (tau - 0.5) / 3.0
    assert O.permeability(np.zeros((n,) * 3), np.full((n,) * 3, U), n, tau, F) == pytest.approx(nu * U / F,
                                                                                                   rel=1e-14)
    assert O.permeability(np.full((n,) * 3, 0.5), np.full((n,) * 3, U), n, tau, F) == pytest.approx(
        0.5 * nu * U / F, rel=1e-14)


def test_numpy_sums_one_cell_pairwise():
    """The premise of the one-cell tolerance of the GPU edge cases: NumPy's
    sum(axis=0) over a (19, 1, 1, 1) state (contiguous along the axis) does
    not add row by row, while over a (19, 2, 2, 2) state it does."""
    rng = np.random.default_rng(0)
    differ = 0
    for _ in range(200):
        v = rng.uniform(0.0, 0.4, 19)
        seq = v[0]
        for x in v[1:]:
            seq = seq + x
        differ += v.reshape(19, 1, 1, 1).sum(axis=0)[0, 0, 0] != seq
        wide = np.broadcast_to(v[:, None, None, None], (19, 2, 2, 2)).copy()
        assert np.all(wide.sum(axis=0) == seq)
    assert differ > 20
