"""Host-side geometry tables and workload generators (CPU)."""

import numpy as np
import pytest

from oracle import psc_oracle as O
from tests.conftest import load_package


@pytest.fixture(scope="module")
def geo():
    return load_package("geometry")


def test_boxes_follow_reference_clamp(geo):
    rng = np.random.default_rng(0)
    n = 17
    table = np.zeros((30, 10))
    table[:, 0:3] = rng.uniform(-3, n + 3, (30, 3))
    table[:, 3] = rng.uniform(0.2, 5.0, 30)
    tabs = geo.build_tables(table, (n, n, n), periodic=False, chunk_cells=100)
    for p in range(30):
        s = O.Sphere(*table[p])
        x0, x1, y0, y1, z0, z1 = O.sphere_box(s, n)
        b = tabs.boxes[p]
        assert (b[0], b[1], b[2]) == (x0, y0, z0)
        assert (b[3], b[4], b[5]) == (max(0, x1 - x0), max(0, y1 - y0), max(0, z1 - z0))
    vols = tabs.boxes[:, 3].astype(int) * tabs.boxes[:, 4] * tabs.boxes[:, 5]
    assert tabs.n_entries == vols.sum()
    assert np.array_equal(tabs.boxes[:, 6], np.concatenate([[0], np.cumsum(vols)[:-1]]))
    # chunks tile every box exactly once
    for p in range(30):
        cs = tabs.chunks[tabs.chunk_first[p]:tabs.chunk_first[p + 1]]
        assert np.all(cs[:, 0] == p)
        assert np.array_equal(cs[:, 1], np.arange(0, vols[p], 100))
    assert not tabs.moving


def test_periodic_boxes_unclamped_and_checked(geo):
    table = np.array([[0.5, 6.0, 11.7, 2.0] + [0] * 6], dtype=float)
    tabs = geo.build_tables(table, (12, 12, 12), periodic=True)
    assert tabs.boxes[0, 0] < 0 and tabs.boxes[0, 2] + tabs.boxes[0, 5] > 12
    with pytest.raises(ValueError):
        geo.build_tables(np.array([[5.0, 5, 5, 6.0] + [0] * 6]), (12, 12, 12), periodic=True)


def test_pack_particles_and_moving_flag(geo):
    psc3d = load_package("psc3d")
    ps = [psc3d.Particle3D(1, 2, 3, 4), psc3d.Particle3D(5, 6, 7, 8, wz=1e-5)]
    t = geo.pack_particles(ps)
    assert t.shape == (2, 10) and t[1, 9] == 1e-5
    assert geo.build_tables(t, (20, 20, 20)).moving
    assert geo.pack_particles([]).shape == (0, 10)
    empty = geo.build_tables(geo.pack_particles([]), (8, 8, 8))
    assert empty.n_entries == 0 and empty.chunks.shape == (0, 2) and list(empty.chunk_first) == [0]


def test_workloads_are_seeded_and_valid():
    wl = load_package("workloads")
    a = wl.fixed_array(seed=0)
    b = wl.fixed_array(seed=0)
    assert np.array_equal(a, b) and a.shape == (60, 10)
    c = a[:, :3]
    d2 = ((c[:, None] - c[None]) ** 2).sum(-1) + np.eye(60) * 1e9
    assert d2.min() >= 24.0 ** 2
    assert c.min() >= 13.0 and c.max() <= 115.0
    m = wl.moving_sphere(seed=0)
    assert m.shape == (1, 10) and m[0, 4] == 1e-4 and m[0, 3] == 2.5
    rng = np.random.default_rng(0)
    assert np.array_equal(m[0, :3], rng.uniform(0.0, 20.0, size=3))
