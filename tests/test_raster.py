"""The host library's cell rasteriser (ems_raster_cells) against a numpy
restatement of its rule (a cell belongs to a solid when its centre, the
midpoints of its node lines, lies inside: box lo <= c < hi, sphere and
cylinder by squared distance), bit for bit on seeded graded grids with
boxes, spheres and cylinders along each axis; against closed forms (the
volume of a sphere and a cylinder on a fine grid, a box between node lines
holding exactly the cells between them, the index fill it replaces); and
its refusals."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd.raster import Solid, cell_mask, raster_cells


def graded(seed, n=(23, 19, 17)):
    rng = np.random.default_rng(seed)
    return E.Grid([np.concatenate([[0.0], np.cumsum(0.5e-3 + 1e-3 * rng.random(k))]) for k in n])


def restated(grid, s):
    c = [0.5 * (l[:-1] + l[1:]) for l in grid.lines]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    C = (X, Y, Z)
    p = s.p
    if s.kind == 0:
        m = np.ones(X.shape, dtype=bool)
        for a in range(3):
            m &= (p[a] <= C[a]) & (C[a] < p[3 + a])
        return m
    if s.kind == 1:
        dx, dy, dz = X - p[0], Y - p[1], Z - p[2]
        return dx * dx + dy * dy + dz * dz < p[3] * p[3]
    ax = s.axis
    r0, r1 = [a for a in range(3) if a != ax]
    d0, d1 = C[r0] - p[0], C[r1] - p[1]
    return (d0 * d0 + d1 * d1 < p[2] * p[2]) & (p[3] <= C[ax]) & (C[ax] < p[4])


def random_solid(rng, grid):
    ext = [l[-1] for l in grid.lines]
    kind = int(rng.integers(3))
    if kind == 0:
        lo = [rng.uniform(-0.1, 0.8) * e for e in ext]
        return Solid.box(lo, [lo[a] + rng.uniform(0.05, 0.6) * ext[a] for a in range(3)])
    if kind == 1:
        return Solid.sphere([rng.uniform(0.0, 1.0) * e for e in ext], rng.uniform(0.05, 0.5) * min(ext))
    ax = int(rng.integers(3))
    o = [a for a in range(3) if a != ax]
    lo = rng.uniform(-0.1, 0.6) * ext[ax]
    return Solid.cylinder(ax, [rng.uniform(0.0, 1.0) * ext[o[0]], rng.uniform(0.0, 1.0) * ext[o[1]]],
                          rng.uniform(0.05, 0.5) * min(ext), lo, lo + rng.uniform(0.1, 0.8) * ext[ax])


@pytest.mark.parametrize("seed", range(8))
def test_matches_restatement(seed):
    rng = np.random.default_rng(500 + seed)
    g = graded(seed)
    ids = np.zeros(g.cells(), dtype=np.uint16)
    want = np.zeros(g.cells(), dtype=np.uint16)
    for q in range(12):
        s = random_solid(rng, g)
        n = raster_cells(g, s, ids, q + 1)
        m = restated(g, s)
        assert n == int(m.sum())
        want[m] = q + 1
        assert np.array_equal(ids, want), (q, s)


def test_volumes():
    h = 0.1e-3
    g = E.Grid.uniform([200, 200, 200], [h] * 3)
    r = 7e-3
    v = cell_mask(g, Solid.sphere([10e-3, 10e-3, 10e-3], r)).sum() * h ** 3
    assert abs(v / (4.0 / 3.0 * math.pi * r ** 3) - 1.0) < 2e-3
    for ax in range(3):
        v = cell_mask(g, Solid.cylinder(ax, [9e-3, 11e-3], 6e-3, 2e-3, 17e-3)).sum() * h ** 3
        assert abs(v / (math.pi * 36e-6 * 15e-3) - 1.0) < 3e-3


@pytest.mark.parametrize("seed", range(4))
def test_box_between_node_lines_is_the_index_fill(seed):
    rng = np.random.default_rng(seed)
    g = graded(10 + seed)
    nc = g.cells()
    fr4 = E.Material.dielectric(4.3).with_kappa(0.02)
    a_ = E.CellMaterials.vacuum(g)
    b_ = E.CellMaterials.vacuum(g)
    for _ in range(5):
        lo = [int(rng.integers(0, nc[a] - 1)) for a in range(3)]
        hi = [int(rng.integers(lo[a] + 1, nc[a] + 1)) for a in range(3)]
        a_.fill_box(list(zip(lo, hi)), fr4)
        n = b_.fill_solid(Solid.box([g.lines[a][lo[a]] for a in range(3)],
                                    [g.lines[a][hi[a]] for a in range(3)]), fr4)
        assert n == (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2])
    assert np.array_equal(a_.ids, b_.ids)


def test_refusals_and_outside():
    g = graded(3)
    ids = np.zeros(g.cells(), dtype=np.uint16)
    with pytest.raises(E.EmsError):
        raster_cells(g, Solid(3, 0, (0.0,) * 6), ids, 1)
    with pytest.raises(E.EmsError):
        raster_cells(g, Solid.cylinder(3, [0.0, 0.0], 1.0, 0.0, 1.0), ids, 1)
    with pytest.raises(E.EmsError):
        raster_cells(g, Solid.sphere([float("nan"), 0.0, 0.0], 1.0), ids, 1)
    with pytest.raises(E.EmsError):
        raster_cells(g, Solid.sphere([0.0, 0.0, 0.0], 1.0), ids.astype(np.int32), 1)
    assert raster_cells(g, Solid.sphere([1.0, 1.0, 1.0], 1e-3), ids, 1) == 0
    assert not ids.any()
