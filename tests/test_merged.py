"""Small cut faces without their own current (ems_amd/merged.py and the
reduced-current operator in the oracle), on the CPU, by the reference's own
cases (src/engine/conformal.rs tests, :3116-3570): the sliver face merged
with the neighbour across its full edge, or interpolated from a host and
its neighbours; the energy invariant held over 5000 steps with
energy() equal to step_with_energy(); the cell-wise bound equal to the
Courant number squared for cubic cells; small faces kept where flat cells
make the time step short; affine-exact weights; the envelope Cholesky
solve against a dense solve; plain Dey-Mittra diverging where the merge
scheme is stable."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import conformal as K
from ems_amd import merged as M
from oracle.bridge import OracleEngine


def _engine(s, dz=1.0):
    g = E.Grid.uniform([12, 12, 12], [1e-3, 1e-3, dz * 1e-3])
    dt = E.TimeStep.courant(s).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    return g, dt, lanes


def _sliver(centroid=False, fraction=0.02, options=None):
    """A z face keeping its full x edge at y = 5 and `fraction` of its area
    (the y edges that open, the far x edge closed)."""
    c = K.ConformalPec(options)
    f = K.ConformalFace(2, (5, 5, 5), fraction)
    if centroid:
        c.push_face_with_centroid(f, (0.5, fraction / 2))
    else:
        c.push_face(f)
    for q in [(5, 5, 5), (6, 5, 5)]:
        c.push_edge(K.ConformalEdge(1, q, fraction))
    c.push_edge(K.ConformalEdge(0, (5, 6, 5), 0.0))
    return c


def _energy_run(g, dt, lanes, pec, sets):
    o = OracleEngine(g, dt, lanes, threads=1)
    rep = pec.apply(o)
    for axis, node, v in sets:
        u = o.u_component(axis)
        u[node] = v
        o.set_component(0, axis, u)
    o.run(1)
    w0 = o.energy()
    dev = 0.0
    for _ in range(50):
        o.run(100)
        dev = max(dev, abs(o.energy() / w0 - 1.0))
    a = o.energy()
    b = o.step_with_energy()
    o.close()
    return rep, dev, a, b


def test_sliver_face_is_merged():
    g, dt, lanes = _engine(0.99)
    changed, rep, op = _sliver().build(lanes, g, dt)
    assert (rep.merged, rep.groups, rep.floored) == (2, 1, 0)
    assert changed[11][5, 5, 5] == 0.0 and changed[11][5, 4, 5] == 0.0   # merged across y - 1
    assert changed[3][5, 6, 5] == 0.0
    assert abs(float(changed[4][5, 5, 5]) / (float(lanes[4][5, 5, 5]) * 0.02) - 1.0) < 1e-6
    assert op.faces[1] == (2, (5, 5, 5), [(0, 1.0), (1, 1.0)])


def test_sliver_face_is_interpolated():
    g, dt, lanes = _engine(0.99)
    changed, rep, op = _sliver(centroid=True).build(lanes, g, dt)
    assert (rep.interpolated, rep.hosts, rep.merged) == (1, 3, 0)
    for p in [(5, 5, 5), (5, 4, 5), (5, 3, 5)]:
        assert changed[11][p] == 0.0, p
    # the small face reads its host (y - 1) and the host's neighbour away from
    # it (y - 2) with the linear weights at its centroid (y offset 0.49 of a
    # cell below the host's centre), plus its deviation current
    w = dict(op.faces[-1][2])
    ids = {f[1]: f[2][0][0] for f in op.faces[:-1]}
    assert op.faces[-1][0:2] == (2, (5, 5, 5))
    assert math.isclose(w[ids[(5, 4, 5)]], 1.51, rel_tol=1e-12)
    assert math.isclose(w[ids[(5, 3, 5)]], -0.51, rel_tol=1e-12)


@pytest.mark.parametrize("centroid", [False, True], ids=["merged", "interpolated"])
def test_reduced_schemes_conserve_energy(centroid):
    g, dt, lanes = _engine(0.99)
    sets = [(0, (5, 5, 5), 1.0), (1, (3, 7, 4), -0.5)]
    if centroid:
        sets += [(1, (4, 4, 4), -0.5), (1, (3, 7, 4), 0.25)]
    rep, dev, a, b = _energy_run(g, dt, lanes, _sliver(centroid), sets)
    assert dev < 1e-5, dev
    assert a == b  # the in-step energy agrees with the look-ahead


def test_plain_dey_mittra_sliver_diverges():
    """Without the reduction (threshold 0) the 2 % face's own current has a
    resonance far above the Courant limit: the run blows up."""
    g, dt, lanes = _engine(0.99)
    pec = _sliver(options=K.ConformalOptions(merge_below=0.0, keep_stable=False))
    o = OracleEngine(g, dt, lanes, threads=1)
    rep = pec.apply(o)
    assert (rep.merged, rep.interpolated, rep.faces) == (0, 0, 1)
    u = o.u_component(0)
    u[5, 5, 5] = 1.0
    o.set_component(0, 0, u)
    o.run(1)
    w0 = o.energy()
    o.run(3000)
    assert not (abs(o.energy()) < 1e6 * w0)
    o.close()


@pytest.mark.parametrize("s", [0.5, 0.8, 0.95])
def test_cell_bound_is_the_courant_number(s):
    g, dt, lanes = _engine(s)
    co = lanes.reshape(12, -1)
    assert abs(M.cell_stability(co, g, dt, (5, 5, 5), {}) - s * s) < 1e-6


@pytest.mark.parametrize("size,s", [((1e-3,) * 3, 0.9), ((1e-3, 0.5e-3, 0.2e-3), 0.7)])
def test_cell_bound_of_cubes_and_bricks(size, s):
    """conformal.rs:3329-3351: for cubes or bricks of a uniform grid the cell
    bound is c^2 dt^2 (1/dx^2 + 1/dy^2 + 1/dz^2) over its limit, the Courant
    factor squared, in every cell; a face of a fifth of its area raises it;
    the batched form agrees."""
    g = E.Grid.uniform([6, 6, 6], list(size))
    dt = E.TimeStep.courant(s).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    co = lanes.reshape(12, -1)
    cells = [(2, 3, 2), (1, 1, 1), (4, 2, 3)]
    for q in cells:
        assert abs(M.cell_stability(co, g, dt, q, {}) - s * s) < 1e-5, q
    got = M.cell_bounds(co, g, dt, np.array(cells), {})
    assert np.allclose(got, s * s, atol=1e-5)
    cut = {(2, M.index(g.nodes(), (2, 3, 2))): 0.2}
    assert M.cell_stability(co, g, dt, (2, 3, 2), cut) > s * s * 1.05


def test_small_faces_keep_their_current_where_the_time_step_allows():
    # cubes at the Courant limit: the face loses its current
    g, dt, lanes = _engine(0.99)
    pec = _sliver(centroid=True, fraction=0.2)
    _, rep, _ = pec.build(lanes, g, dt)
    assert rep.kept == 0 and rep.interpolated + rep.merged > 0 and not rep.bounded
    # cells a tenth as high as wide: the face keeps its current and area
    g, dt, lanes = _engine(0.99, dz=0.1)
    changed, rep, op = _sliver(centroid=True, fraction=0.2).build(lanes, g, dt)
    assert (rep.kept, rep.interpolated, rep.merged, rep.floored) == (1, 0, 0, 0)
    assert rep.bounded and op is None
    assert abs(float(changed[11][5, 5, 5]) / float(lanes[11][5, 5, 5]) - 5.0) < 1e-5
    rep, dev, a, b = _energy_run(g, dt, lanes, _sliver(centroid=True, fraction=0.2),
                                 [(1, (5, 5, 5), 1.0), (0, (5, 5, 5), -0.5)])
    assert dev < 1e-5, dev
    # switched off: as on the cubes
    pec = _sliver(centroid=True, fraction=0.2, options=K.ConformalOptions(keep_stable=False))
    _, rep, _ = pec.build(lanes, g, dt)
    assert rep.kept == 0 and rep.interpolated + rep.merged > 0


def test_weights_are_affine_exact():
    for d in ([(0.0, 0.0)], [(0.6, 0.0), (1.6, 0.0)], [(0.6, 0.2), (1.6, 0.25), (0.55, -0.8)]):
        w = M.interpolation_weights(d)
        assert abs(sum(w) - 1.0) < 1e-12
        for ax in range(2):
            assert abs(sum(wi * di[ax] for wi, di in zip(w, d))) < 1e-12
    assert M.interpolation_weights([(0.7, 0.1)]) is None
    assert M.interpolation_weights([(0.6, 0.1), (1.6, 0.0)]) is None
    assert M.interpolation_weights([(5.0, 0.0), (6.0, 0.0)]) is None
    assert M.interpolation_weights([(1.0, 0.0), (2.0, 0.0), (3.0, 0.0)]) is None


def test_envelope_cholesky_solves():
    n = 7
    a = [[(i, 4.0 + i)] for i in range(n)]

    def link(i, j, v):
        a[i].append((j, v))
        a[j].append((i, v))
    for i in range(n - 1):
        link(i, i + 1, -1.0)
    link(0, n - 1, 0.5)
    f = M.Envelope.factor(a)
    x = [math.sin(i) + 0.3 for i in range(n)]
    b = [sum(v * x[j] for j, v in row) for row in a]
    f.solve(b)
    assert max(abs(p - q) for p, q in zip(b, x)) < 1e-12
    dense = np.zeros((n, n))
    for i, row in enumerate(a):
        for j, v in row:
            dense[i, j] = v
    L = np.zeros((n, n))
    for i in range(n):
        L[i, f.first[i]:i + 1] = f.rows[i]
    assert np.allclose(L @ L.T, dense, rtol=0, atol=1e-12)
    assert M.Envelope.factor([[(0, 1.0), (1, 2.0)], [(0, 2.0), (1, 1.0)]]) is None


def test_rcm_order_keeps_a_path_banded():
    """A path graph given in scrambled order comes out as the path: the
    search starts at the end of smaller number (both ends have the minimum
    degree), 2, and the reversed order ends there."""
    perm = [3, 0, 5, 1, 4, 2]
    adj = {perm[i]: [] for i in range(6)}
    for i in range(5):
        adj[perm[i]].append(perm[i + 1])
        adj[perm[i + 1]].append(perm[i])
    order = M.rcm_order(sorted(adj), lambda q: sorted(adj[q]))
    assert order == perm


def test_rejects_bad_input():
    g, dt, lanes = _engine(0.9)
    c = K.ConformalPec()
    c.push_edge(K.ConformalEdge(0, (12, 0, 0), 0.5))
    with pytest.raises(E.EmsError):
        c.build(lanes, g, dt)
    c = K.ConformalPec()
    c.push_face(K.ConformalFace(0, (0, 0, 0), 1.5))
    with pytest.raises(E.EmsError):
        c.build(lanes, g, dt)
    c = K.ConformalPec()
    c.push_face_with_centroid(K.ConformalFace(2, (3, 3, 3), 0.3), (0.5, 1.2))
    with pytest.raises(E.EmsError):
        c.build(lanes, g, dt)


def test_cell_bounds_at_once_equal_the_cell_by_cell_form():
    """The batched bound of many cells (graded grid, cut edges and faces,
    boundary cells) against cell_stability one cell at a time."""
    rng = np.random.default_rng(3)
    lines = [np.cumsum(np.r_[0.0, rng.uniform(0.5e-3, 1.5e-3, 9)]) for _ in range(3)]
    g = E.Grid(lines)
    dt = E.TimeStep.courant(0.9).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    co = M.Overlay(lanes.reshape(12, -1))
    n = g.nodes()
    for _ in range(40):
        a = int(rng.integers(3))
        p = [int(rng.integers(n[d] - 1)) for d in range(3)]
        k = M.index(n, p)
        co[M.CB + a, k] = np.float32(float(co[M.CB + a, k]) * rng.uniform(0.1, 1.0))
    fr = {(int(rng.integers(3)), M.index(n, [int(rng.integers(n[d] - 1)) for d in range(3)])):
          float(rng.uniform(0.05, 1.0)) for _ in range(40)}
    cells = np.array([[i, j, k] for i in range(n[0]) for j in range(0, n[1], 3)
                      for k in range(0, n[2], 2)])
    got = M.cell_bounds(co, g, dt, cells, fr)
    want = np.array([M.cell_stability(co, g, dt, tuple(c), fr) for c in cells])
    assert np.allclose(got, want, rtol=1e-10, atol=0), np.abs(got - want).max()


def test_columns_give_the_row_solve_bit_for_bit():
    """The device's backward substitution by columns (the factor's column
    lists of the descriptor, decreasing rows) performs the row-by-row
    solve's operations in the same order: equal to Envelope.solve bit for
    bit on the blocks of a cut cylinder."""
    g = E.Grid.uniform([30, 30, 8], [1e-3] * 3)
    m = E.CellMaterials.vacuum(g)
    pec = K.cylinder_z(g, (15.2e-3, 14.9e-3), 9.3e-3, mats=m, metal="inside", z_nodes=(1, 7))
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, m, [E.FaceCondition.Pec] * 6, dt)
    _, rep, op = pec.build(lanes, g, dt)
    d = op.desc_arrays()
    rng = np.random.default_rng(5)
    big = 0
    for b, blk in enumerate(op.blocks):
        r0, nr = int(d["blk_start"][b]), len(blk.members)
        big = max(big, nr)
        y = list(rng.standard_normal(nr))
        want = list(y)
        blk.factor.solve(want)
        x = list(y)
        for i in range(nr):   # forward, as the row form
            f, row = blk.factor.first[i], blk.factor.rows[i]
            s = x[i]
            for j in range(f, i):
                s -= row[j - f] * x[j]
            x[i] = s / row[i - f]
        for j in range(nr - 1, -1, -1):
            s = x[j]
            for q in range(int(d["col_start"][r0 + j]), int(d["col_start"][r0 + j + 1])):
                s -= float(d["col_val"][q]) * x[int(d["col_row"][q])]
            x[j] = s / blk.factor.rows[j][j - blk.factor.first[j]]
        assert np.array_equal(np.array(x).view(np.uint64), np.array(want).view(np.uint64)), b
    assert big > 10 and rep.interpolated > 0


def test_energy_lookahead_includes_the_pml():
    """The sliver inside an x+ CPML layer (conformal.rs:3294-3327): merged
    (no interpolation in layers), its group applies the layer's H-side
    correction, and the look-ahead update from the circulations after the
    E half-step is exactly the next step's currents."""
    g, dt, lanes = _engine(0.99)
    o = OracleEngine(g, dt, lanes, threads=1)
    cpml = E.CpmlBoundary.new(o, [(E.Face.XMax, E.Cpml.with_cells(6))])
    pec = K.ConformalPec()
    f = K.ConformalFace(2, (8, 5, 5), 0.02)
    pec.push_face_with_centroid(f, (0.5, 0.01))
    for q in [(8, 5, 5), (9, 5, 5)]:
        pec.push_edge(K.ConformalEdge(1, q, 0.02))
    pec.push_edge(K.ConformalEdge(0, (8, 6, 5), 0.0))
    pec.set_cpml(cpml)
    rep = pec.apply(o)
    assert rep.groups == 1 and rep.interpolated == 0
    o.add_operator(cpml)
    op = [x for x in o.ops if type(x).__name__ == "MergedFaces"][0]
    assert len(op.pml) == 2 and all(t.other == 1 and t.sign == 1.0 for t in op.pml)
    for axis, node, v in [(0, (8, 5, 5), 1.0), (1, (7, 7, 4), -0.5)]:
        u = o.u_component(axis)
        u[node] = v
        o.set_component(0, axis, u)
    for _ in range(5):
        o.run(7)
        cur, circ = op.currents()
        ahead = op.advance(list(cur), list(circ))
        o.run(1)
        now, _ = op.currents()
        assert np.array_equal(np.array(ahead), now)
        assert np.any(now != 0.0)
    # the layer's correction is in the circulations: they differ from the
    # plain sums of the groups' terms over u
    _, circ = op.currents()
    u = [o.u_component(a) for a in range(3)]
    plain = [sum(w * float(u[a][p]) for a, p, w in gr.terms) for gr in op.groups]
    assert any(abs(c - q) > 1e-6 * (abs(q) + 1e-30) for c, q in zip(circ, plain))
    o.close()


def test_uncut_grid_stability_is_the_courant_factor():
    """The power iteration on an uncut 12^3 PEC box at Courant factor 0.9
    (conformal.rs:3126-3133): its highest mode is slightly below the
    infinite grid's limit."""
    g, dt, lanes = _engine(0.9)
    s = M.stability(list(lanes.reshape(12, -1)), g.nodes(), None, [], 400)
    assert 0.85 < s < 0.9, s


def test_stability_check_of_the_sliver():
    """conformal.rs:3160-3189, 3232-3246: plain Dey-Mittra (threshold 0,
    one round) reads unstable with a warning; merged and interpolated read
    stable, a little below the limit."""
    g, dt, lanes = _engine(0.99)
    raw = _sliver(options=K.ConformalOptions(merge_below=0.0, max_rounds=1))
    _, r, _ = raw.build(lanes, g, dt)
    assert not r.verified and r.stability > 1.5 and r.warning() and r.merged == 0
    for cen in (False, True):
        _, r, _ = _sliver(centroid=cen).build(lanes, g, dt)
        assert r.verified and 0.9 < r.stability < 0.999 and not r.proven, r
        assert r.warning() is None


def test_flat_cells_are_proven_by_the_cell_bound():
    """conformal.rs:3432-3455: with the face kept on flat cells (no reduced
    currents) the grid-wide cell bound stands in for the estimate, and the
    power iteration finds no more than it; on cubes the face is
    interpolated and the estimate is made."""
    g, dt, lanes = _engine(0.99, dz=0.1)
    changed, r, op = _sliver(centroid=True, fraction=0.2).build(lanes, g, dt)
    assert r.proven and r.verified and r.stability < 0.999 and op is None, r
    full = lanes.reshape(12, -1).copy()
    for l, v in changed.items():
        full[l] = v.reshape(-1)
    est = M.stability(list(full), g.nodes(), None, [(2, M.index(g.nodes(), (5, 5, 5)))], 300)
    assert est <= r.stability, (est, r.stability)
    g, dt, lanes = _engine(0.99)
    _, r, _ = _sliver(centroid=True, fraction=0.2).build(lanes, g, dt)
    assert not r.proven and r.verified


def test_stability_rejects_invalid_tables():
    """ems_conformal_stability checks every index it follows: a seed edge
    outside the grid is refused, not read."""
    g, dt, lanes = _engine(0.9)
    with pytest.raises(ValueError):
        M.stability(list(lanes.reshape(12, -1)), g.nodes(), None, [(1, 10 ** 9)], 10)
    with pytest.raises(ValueError):
        M.stability(list(lanes.reshape(12, -1)), g.nodes(), None, [(5, 0)], 10)
