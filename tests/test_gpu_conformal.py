"""Parity of conformal PEC on the device with the oracle: the coefficient
patches and the reduced currents of small cut faces (merge groups,
interpolation hosts, deviation currents, envelope-Cholesky blocks) on a
circular cavity, a PEC rod in a graded CPML box, and the reference's
sliver face merged and interpolated; fields, probes, reduced currents and
energy bit for bit. Engines with reduced currents run the unfused
schedule whatever is asked."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import conformal as CF

import cases as K
from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=[(0, 0, 0), (1, 7, 0), (2, 3, 16)],
                ids=lambda s: "mode%d_tj%d_xc%d" % s)
def schedule(request):
    f, t, c = request.param
    old = dict(K.SCHEDULE)
    K.SCHEDULE.update(fused=f, tile_rows=t, x_chunks=c)
    yield
    K.SCHEDULE.update(old)


def test_conformal_cavity():
    def build(kind):
        g = E.Grid.uniform([17, 17, 3], [1e-3] * 3)
        m = E.CellMaterials.vacuum(g)
        pec = CF.cylinder_z(g, (8.5e-3, 8.5e-3), 7.3e-3, mats=m, metal="outside")
        e = K.make(kind, g, m, threads=2)
        rep = pec.apply(e)
        assert rep.faces > 0 and rep.interpolated > 0 and rep.hosts > 0
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(15e9, 10e9)).edge(E.Axis.Z, [7, 8, 1], 1.0))
        p = e.add_operator(E.Probe.voltage(E.Axis.Z, [8, 8, 1]))
        return e, p, pec
    (g, a, pg), (o, b, po) = build("gpu"), build("oracle")
    for n in (1, 300, 1200):
        g.run(n)
        o.run(n)
        wg, wo = g.step_with_energy(), o.step_with_energy()
        assert np.float64(wg).tobytes() == np.float64(wo).tobytes(), (wg, wo)
    assert np.array_equal(bits(a.values()), bits(o.probe_values(b)))
    same_fields(g, o)
    op_g = [x for x in g.setup if x[0] == "merged"][0][1]
    op_o = [x for x in o.ops if type(x).__name__ == "MergedFaces"][0]
    for x, y in zip(op_g.currents(), op_o.currents()):
        assert np.array_equal(bits(x), bits(y))
    assert np.abs(op_g.currents()[0]).max() > 0.0


@pytest.mark.parametrize("centroid", [False, True], ids=["merged", "interpolated"])
def test_sliver_face(centroid):
    """The reference's sliver (2 % of a z face along its full x edge) with
    fields set on both sides of it: merged across y - 1, or interpolated
    from three hosts with a deviation current."""
    def build(kind):
        g = E.Grid.uniform([12, 12, 12], [1e-3] * 3)
        e = K.make(kind, g, None)
        pec = CF.ConformalPec()
        f = CF.ConformalFace(2, (5, 5, 5), 0.02)
        if centroid:
            pec.push_face_with_centroid(f, (0.5, 0.01))
        else:
            pec.push_face(f)
        for q in [(5, 5, 5), (6, 5, 5)]:
            pec.push_edge(CF.ConformalEdge(1, q, 0.02))
        pec.push_edge(CF.ConformalEdge(0, (5, 6, 5), 0.0))
        rep = pec.apply(e)
        assert (rep.merged, rep.interpolated) == ((0, 1) if centroid else (2, 0))
        for axis, node, v in [(0, (5, 5, 5), 1.0), (1, (3, 7, 4), -0.5), (1, (4, 4, 4), 0.25)]:
            u = e.u_component(axis)
            u[node] = v
            e.set_component(0, axis, u)
        op = [x for x in (e.setup if kind == "gpu" else [(0, y) for y in e.ops])
              if type(x[1]).__name__ == "MergedFaces"][0][1]
        return e, op
    (g, pg), (o, po) = build("gpu"), build("oracle")
    for n in (1, 7, 400):
        g.run(n)
        o.run(n)
        wg, wo = g.step_with_energy(), o.step_with_energy()
        assert np.float64(wg).tobytes() == np.float64(wo).tobytes(), (n, wg, wo)
        eg, eo = g.energy(), o.energy()
        assert np.float64(eg).tobytes() == np.float64(eo).tobytes(), (n, eg, eo)
        for x, y in zip(pg.currents(), po.currents()):
            assert np.array_equal(bits(x), bits(y)), n
    same_fields(g, o)
    assert np.abs(pg.currents()[0]).max() > 0.0


def test_sliver_face_in_cpml():
    """The sliver inside an x+ CPML layer: merged, its group applying the
    layer's H-side correction (committed after H, evaluated for the energy
    look-ahead after E)."""
    def build(kind):
        g = E.Grid.uniform([12, 12, 12], [1e-3] * 3)
        e = K.make(kind, g, None)
        cpml = E.CpmlBoundary.new(e, [(E.Face.XMax, E.Cpml.with_cells(6))])
        pec = CF.ConformalPec()
        pec.push_face_with_centroid(CF.ConformalFace(2, (8, 5, 5), 0.02), (0.5, 0.01))
        for q in [(8, 5, 5), (9, 5, 5)]:
            pec.push_edge(CF.ConformalEdge(1, q, 0.02))
        pec.push_edge(CF.ConformalEdge(0, (8, 6, 5), 0.0))
        pec.set_cpml(cpml)
        rep = pec.apply(e)
        assert rep.groups == 1
        e.add_operator(cpml)
        for axis, node, v in [(0, (8, 5, 5), 1.0), (1, (7, 7, 4), -0.5)]:
            u = e.u_component(axis)
            u[node] = v
            e.set_component(0, axis, u)
        op = [x for x in (e.setup if kind == "gpu" else [(0, y) for y in e.ops])
              if type(x[1]).__name__ == "MergedFaces"][0][1]
        assert len(op.pml) == 2
        return e, op
    (g, pg), (o, po) = build("gpu"), build("oracle")
    for n in (1, 7, 300):
        g.run(n)
        o.run(n)
        wg, wo = g.step_with_energy(), o.step_with_energy()
        assert np.float64(wg).tobytes() == np.float64(wo).tobytes(), (n, wg, wo)
        for x, y in zip(pg.currents(), po.currents()):
            assert np.array_equal(bits(x), bits(y)), n
    same_fields(g, o)
    assert np.abs(pg.currents()[0]).max() > 0.0


def test_conformal_rod_in_cpml_box():
    def build(kind):
        lines = []
        for n in (28, 26, 20):
            x, l = 0.0, []
            for i in range(n + 1):
                l.append(x)
                x += 1e-3 * (1.0 + 0.25 * math.sin(0.7 * i))
            lines.append(l)
        g = E.Grid(lines)
        m = E.CellMaterials.vacuum(g)
        m.fill_box([(3, 10), (4, 12), (5, 15)], E.Material.dielectric(3.0).with_kappa(0.01))
        xc, yc = g.lines[0][17] + 0.37e-3, g.lines[1][13] - 0.21e-3
        pec = CF.cylinder_z(g, (xc, yc), 3.6e-3, mats=m, metal="inside")
        e = K.make(kind, g, m, threads=2)
        pec.apply(e)
        e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(5)) for f in E.Face]))
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(12e9, 8e9)).edge(E.Axis.Z, [8, 8, 10], 1.0))
        p1 = e.add_operator(E.Probe.voltage(E.Axis.Z, [22, 13, 10]))
        p2 = e.add_operator(E.Probe.current(E.Axis.Y, [16, 13, 10]))
        return e, p1, p2
    (g, a1, a2), (o, b1, b2) = build("gpu"), build("oracle")
    for n in (1, 200, 250):
        g.run(n)
        o.run(n)
    assert np.array_equal(bits(a1.values()), bits(o.probe_values(b1)))
    assert np.array_equal(bits(a2.values()), bits(o.probe_values(b2)))
    same_fields(g, o)


def _synthetic_merged(grid, dt, sizes, seed=3):
    """A MergedFaces operator built by hand: one merge group, then one block
    per entry of `sizes` with a full (dense) factor, so that the device runs
    its general LDS path (a small dense block), its global-scratch path (a
    block whose image exceeds 64 KiB) and its narrow path (a tridiagonal
    block); circulation terms on random edges, driven faces with two weights."""
    from ems_amd import merged as MF
    rng = np.random.default_rng(seed)
    n = grid.nodes()

    def edge():
        a = int(rng.integers(3))
        p = [int(rng.integers(n[d] - (1 if d == a else 0))) for d in range(3)]
        return a, tuple(p)
    groups = []
    total = 1 + sum(abs(s) for s in sizes)
    for q in range(total):
        terms = sorted({edge(): float(rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 1.5))
                        for _ in range(4)}.items(), key=lambda t: (t[0][0], MF.index(n, t[0][1])))
        terms = [(a, p, w) for (a, p), w in terms]
        groups.append(MF.Group.new(terms, 1e-9 * (1 + q % 3), 0.0, dt))
    blocks, first = [], 1
    for sz in sizes:
        m = abs(sz)
        rows = []
        for i in range(m):
            row = [(i, 4.0e-9 + 1e-10 * i)]
            for j in range(m):
                if j != i and (sz > 0 or abs(i - j) == 1):
                    row.append((j, -1.0e-9 / m if sz > 0 else -1.0e-9))
            rows.append(sorted(row))
        fac = MF.Envelope.factor(rows)
        assert fac is not None
        blocks.append(MF.Block(list(range(first, first + m)), fac, rows, rows, dt))
        first += m
    faces, used = [], set()
    while len(faces) < 12:
        a = int(rng.integers(3))
        b, c = MF.others(a)
        p = [int(rng.integers(n[d] - (1 if d in (b, c) else 0))) for d in range(3)]
        if (a, tuple(p)) in used:
            continue
        used.add((a, tuple(p)))
        faces.append((a, tuple(p), [(int(rng.integers(total)), 0.7), (0, 0.3)]))
    return MF.MergedFaces(tuple(n), dt, groups, 1, blocks, faces, [])


def test_reduced_current_block_paths():
    """Every device path of the block solve against the oracle, bit for bit:
    a dense 20-row block (general LDS path of the small-block launch), a
    dense 40-row block (general LDS path of the large-block launch), a
    dense 140-row block (image above 64 KiB: global scratch), a tridiagonal
    300-row block (narrow path of the large-block launch)."""
    def build(kind):
        g = E.Grid.uniform([10, 9, 8], [1e-3] * 3)
        e = K.make(kind, g, None)
        op = _synthetic_merged(g, e.dt, [20, 40, 140, -300])
        e.add_operator(op)
        rng = np.random.default_rng(11)
        for a in range(3):
            e.set_component(0, a, rng.standard_normal(g.nodes()).astype(np.float32))
        return e, op
    (g, pg), (o, po) = build("gpu"), build("oracle")
    for n in (1, 3, 6):
        g.run(n)
        o.run(n)
        for x, y in zip(pg.currents(), po.currents()):
            assert np.array_equal(bits(x), bits(y)), n
        wg, wo = g.energy(), o.energy()
        assert np.float64(wg).tobytes() == np.float64(wo).tobytes(), (n, wg, wo)
    same_fields(g, o)
    assert np.abs(pg.currents()[0]).max() > 0.0
