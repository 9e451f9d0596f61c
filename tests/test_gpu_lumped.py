"""Parity of the device lumped elements and ports with the oracle: fields,
port V/I, terminal currents, branch state and energies, bit for bit, on
the unfused schedule (u^n captured in the mid gap) and the fused ones (u^n
read from the previous buffer)."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import post
from ems_amd.lumped import EdgeValues, LumpedBox, LumpedElement, LumpedPort, Topology

import cases as K
from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu

SCHEDULES = [(0, 0, 0), (1, 7, 0), (1, 3, 4), (3, 7, 0)]


@pytest.fixture(autouse=True, params=SCHEDULES, ids=lambda s: "mode%d_tj%d_xc%d" % s)
def schedule(request):
    f, t, c = request.param
    old = dict(K.SCHEDULE)
    K.SCHEDULE.update(fused=f, tile_rows=t, x_chunks=c)
    yield
    K.SCHEDULE.update(old)


def same(a, b):
    assert len(a) == len(b) and np.array_equal(bits(np.asarray(a, dtype=np.float64)),
                                               bits(np.asarray(b, dtype=np.float64)))


def test_port_with_series_rl_load():
    """Port (Gaussian Thevenin source) and a series R+L branch on the loop."""
    load = lambda b: LumpedElement.new(b).r(20.0).l(5e-9).topology(Topology.Series)  # noqa: E731
    (g, pg), (o, po) = K.run_loop("gpu", load, steps=600), K.run_loop("oracle", load, steps=600)
    same(pg.voltage(), po.voltage())
    same(pg.current(), po.current())
    assert pg.voltage_t0() == po.voltage_t0()
    same_fields(g, o)
    rg, ro = pg.analyze(K.FREQS), po.analyze(K.FREQS)
    assert rg.s11() == ro.s11()


def test_parallel_rlc_explicit_inductor_and_ramp_source():
    """Explicit (leapfrog) inductor, folded R and C, and a callable ramp
    source whose table grows while running."""
    def build(kind):
        g = E.Grid.uniform([20, 20, 20], [1e-3] * 3)
        e = K.make(kind, g, K.loop_materials(g), threads=2)
        p = e.add_operator(LumpedPort(K.loop_column(K.LOOP["port_x"]), 50.0).excite(K.ramp(0.3e-9)))
        h = e.add_operator(LumpedElement.new(K.loop_column(K.LOOP["load_x"])).r(200.0).l(8e-9)
                           .c(0.5e-12).explicit_inductor(True).record_current())
        return e, p, h
    (g, pg, hg), (o, po, ho) = build("gpu"), build("oracle")
    for n in (1, 700, 900):
        g.run(n)
        o.run(n)
    same(pg.voltage(), po.voltage())
    same(pg.current(), po.current())
    same(hg.current(), ho.current())
    for a, b in zip(hg.branch_state(), ho.branch_state()):
        same(a, b)
    same_fields(g, o)


def test_two_port_network():
    """Two ports, a lumped capacitor and a series RC shunt (tests/lumped.rs
    two_port), 2000 steps."""
    (g, g1, g2), (o, o1, o2) = K.two_port("gpu", 1, steps=2000), K.two_port("oracle", 1, steps=2000)
    for a, b in ((g1, o1), (g2, o2)):
        same(a.voltage(), b.voltage())
        same(a.current(), b.current())
    same_fields(g, o)
    assert post.s_param(g2.analyze(K.FREQS), g1.analyze(K.FREQS)) == \
        post.s_param(o2.analyze(K.FREQS), o1.analyze(K.FREQS))


def test_lc_cavity_energy_with_branches():
    """Branch energies enter step_with_energy and energy (before the step)."""
    (g, hg, sg), (o, ho, so) = K.lc_cavity("gpu"), K.lc_cavity("oracle")
    for _ in range(40):
        assert np.float64(g.energy()).tobytes() == np.float64(o.energy()).tobytes()
        assert np.float64(g.step_with_energy()).tobytes() == \
            np.float64(o.step_with_energy()).tobytes()
        g.run(3)
        o.run(3)
    assert hg.branch_energy() == ho.branch_energy()
    assert sg.branch_current() == so.branch_current()
    same_fields(g, o)


def test_edge_list_drive_source_with_h_source_and_current_probe():
    """An edge-list series branch with per-edge and branch-drive sources,
    next to an H source (the mid gap runs between H and E) and a current
    probe."""
    def build(kind):
        g = E.Grid.uniform([9, 8, 7], [1e-3] * 3)
        e = K.make(kind, g, threads=2)
        e.add_operator(E.EdgeSource(lambda t: math.sin(t * 2e10), horizon=400)
                       .kind(E.SourceKind.SoftH).edge(E.Axis.X, [3, 3, 3], 1.0))
        edges = [EdgeValues([4, 4, z], r=30.0, l=2e-9, c=1e-12) for z in (2, 3, 4)]
        el = (LumpedElement.from_edges(E.Axis.Z, Topology.Series, edges)
              .edge_source([0.5, 0.25, 0.25], lambda t: math.cos(t * 1.3e10))
              .branch_drive_source([1.0, 0.0, -1.0], lambda t: 0.1 * math.sin(t * 7e9))
              .record_current())
        h = e.add_operator(el)
        p = e.add_operator(E.Probe.current(E.Axis.Y, [4, 3, 3]).interval(2))
        return e, h, p
    (g, hg, pg), (o, ho, po) = build("gpu"), build("oracle")
    g.run(350)
    o.run(350)
    same(hg.current(), ho.current())
    same(pg.values(), o.probe_values(po))
    same_fields(g, o)
