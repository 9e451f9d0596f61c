"""The host library's setup producers (ems_dispersive_fill_cells,
ems_dispersive_plan, ems_lumped_plan: Dispersive::fill_cells and install,
LumpedElement::install of the reference) against their Python
restatements, bit for bit, on seeded graded grids: dispersive cell boxes
of every pole kind and both fields, plans in both schemes with PEC edges
skipped, and lumped elements over boxes and edge lists in both topologies
with every combination of R, L and C."""
import itertools

import numpy as np
import pytest

import ems_amd as E
from ems_amd.dispersive import Dispersive, Field, Pole, Scheme
from ems_amd.lumped import EdgeValues, LumpedBox, LumpedElement, Topology

import setup_restatement as R


def graded(seed, n=(9, 8, 10)):
    rng = np.random.default_rng(seed)
    return E.Grid([np.concatenate([[0.0], np.cumsum(0.5e-3 + 1e-3 * rng.random(k))]) for k in n])


def bits(x):
    return np.float64(x).tobytes()


def random_poles(rng):
    out = []
    for kind in rng.permutation(R.KINDS)[: 1 + rng.integers(3)]:
        if kind == "lorentz":
            out.append(Pole.lorentz(1e10 * (1 + rng.random()), 2e10 * rng.random(), 1e9 * rng.random()))
        elif kind == "drude":
            out.append(Pole.drude(1e10 * (1 + rng.random()), 1e9 * rng.random()))
        else:
            out.append(Pole.debye(3.0 * rng.random() + 0.1, 1e-11 * (1 + rng.random())))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_dispersive_fill_and_plan(seed):
    rng = np.random.default_rng(100 + seed)
    g = graded(seed)
    dt = 0.9 * E.courant_limit(g)
    mats = E.CellMaterials.vacuum(g)
    mats.fill_pec_box([(2, 5), (3, 4), (4, 5)])  # PEC edges: their poles are skipped
    lanes = E.coefficients(g, mats, [E.FaceCondition.Pec] * 6, dt)
    nc = g.cells()
    d = Dispersive.new().with_scheme(Scheme.Leapfrog if seed % 2 else Scheme.Trapezoidal)
    ref = []
    for _ in range(3):
        lo = [int(rng.integers(0, nc[a])) for a in range(3)]
        hi = [int(rng.integers(lo[a] + 1, nc[a] + 1)) for a in range(3)]
        if seed == 0 and not ref:  # an empty box adds nothing
            d.fill_cells(g, [(lo[0], lo[0]), (lo[1], hi[1]), (lo[2], hi[2])], Field.Electric,
                         [Pole.debye(1.0, 1e-11)])
        box = list(zip(lo, hi))
        field = Field(int(rng.integers(2)))
        poles = random_poles(rng)
        d.fill_cells(g, box, field, poles)
        ref += R.fill_cells(g, box, int(field), [(p.kind, p.p) for p in poles])
    got = [(f, a, n, p.kind, p.p) for f, a, n, p in d.entries]
    assert len(got) == len(ref) > 0
    for x, y in zip(got, ref):
        assert x[:4] == y[:4] and all(bits(u) == bits(v) for u, v in zip(x[4], y[4]))
    plan = d._plan(g, dt, lanes)
    want = R.dispersive_plan(ref, seed % 2 == 1, g, dt, lanes)
    assert plan["skipped"] == want["skipped"]
    for f in (0, 1):
        a, b = plan["sides"][f], want["sides"][f]
        assert a["axis"] == b["axis"] and a["nodes"] == b["nodes"]
        assert a["pole_end"] == b["pole_end"] and a["cidx"] == b["cidx"]
        assert [bits(v) for v in a["sum_inv_z"]] == [bits(v) for v in b["sum_inv_z"]]
    assert len(plan["consts"]) == len(want["consts"])
    for x, y in zip(plan["consts"], want["consts"]):
        assert [bits(v) for v in x[:4]] == [bits(v) for v in y[:4]] and x[4] == y[4]


def test_plan_skips_pec_edges_and_counts_them():
    g = graded(7)
    dt = 0.9 * E.courant_limit(g)
    mats = E.CellMaterials.vacuum(g)
    mats.fill_pec_box([(0, 9), (0, 8), (3, 4)])
    lanes = E.coefficients(g, mats, [E.FaceCondition.Pec] * 6, dt)
    d = Dispersive.new()
    d.fill_cells(g, [(1, 4), (1, 4), (2, 5)], Field.Electric, [Pole.debye(2.0, 1e-11)])
    plan = d._plan(g, dt, lanes)
    assert plan["skipped"] > 0
    assert plan["skipped"] == R.dispersive_plan(
        [(f, a, n, p.kind, p.p) for f, a, n, p in d.entries], False, g, dt, lanes)["skipped"]


CASES = [(topo, present, explicit) for topo in (Topology.Parallel, Topology.Series)
         for present in itertools.product([False, True], repeat=3) if any(present)
         for explicit in (False, True)]


@pytest.mark.parametrize("topo,present,explicit", CASES)
@pytest.mark.parametrize("form", ["box", "edges"])
def test_lumped_plan(topo, present, explicit, form):
    g = graded(11)
    dt = 0.9 * E.courant_limit(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    rng = np.random.default_rng(64 * int(topo) + 32 * present[0] + 16 * present[1] + 8 * present[2]
                                + 2 * explicit + (form == "box"))
    vals = [50.0 * (1 + rng.random()), 1e-9 * (1 + rng.random()), 1e-12 * (1 + rng.random())]
    r, l, c = [v if p else None for v, p in zip(vals, present)]
    axis = E.Axis.Y
    if form == "box":
        box = LumpedBox(axis, [2, 2, 3], [4, 5, 5])
        el = LumpedElement.new(box).topology(topo).explicit_inductor(explicit)
        for v, f in ((r, el.r), (l, el.l), (c, el.c)):
            if v is not None:
                f(v)
        edges = box.edges()
        parts, hb = R.lumped_parts(True, box.series_edges(), box.columns(), r, l, c,
                                   [None] * len(edges), topo == Topology.Parallel)
    else:
        edges = [(2, 2, 3), (2, 3, 3), (5, 2, 6), (6, 4, 2)]
        ev = []
        for q, nd in enumerate(edges):
            s = 1.0 + 0.1 * q
            ev.append(EdgeValues(nd, *(None if v is None else v * s for v in (r, l, c))))
        el = LumpedElement.from_edges(axis, topo, ev).explicit_inductor(explicit)
        parts, hb = R.lumped_parts(False, 0, 0, None, None, None,
                                   [(e.r, e.l, e.c) for e in ev], topo == Topology.Parallel)
    el.record_current()
    plan = el._plan(g, dt, lanes)
    a = int(axis)
    ca, cb = lanes[a].reshape(-1).copy(), lanes[3 + a].reshape(-1).copy()
    want = R.lumped_plan(g.nodes(), a, edges, parts, hb, topo == Topology.Parallel, explicit,
                         dt, ca, cb)
    assert np.array_equal(plan["ca"].reshape(-1).view(np.uint32), ca.view(np.uint32))
    assert np.array_equal(plan["cb"].reshape(-1).view(np.uint32), cb.view(np.uint32))
    assert plan["branch"] == want["branch"] and plan["src_in_branch"] == want["src_in_branch"]
    for k in ("cb_e", "g_src", "rec_g", "rec_c"):
        assert [bits(v) for v in plan[k]] == [bits(v) for v in want[k]], k
    if want["branch"]:
        for k in ("two_l_dt", "inv_c", "z"):
            assert [bits(v) for v in plan[k]] == [bits(v) for v in want[k]], k
    else:
        assert plan["two_l_dt"] is None


def test_lumped_plan_refuses_pec_and_outside():
    g = graded(12)
    dt = 0.9 * E.courant_limit(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    with pytest.raises(E.EmsError):  # y = 0 is the PEC face: its z edges are PEC
        LumpedElement.resistor(LumpedBox(E.Axis.Z, [2, 0, 2], [2, 0, 4]), 50.0)._plan(g, dt, lanes)
    with pytest.raises(E.EmsError):
        LumpedElement.resistor(LumpedBox(E.Axis.Z, [2, 2, 2], [2, 2, 40]), 50.0)._plan(g, dt, lanes)
