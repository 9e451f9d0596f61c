"""openEMS's PML_n layers on the device against the oracle, bit for bit:
the openEMS fixture ``pml_face`` (and its probe files, to their printed
digits), and a graded box with three overlapping layers (edges and a
corner), a lossy dielectric-magnetic block reaching into them, E and H
sources, in both curl orders; the layer state as well as the fields and
probes; the registration rules and the batch refusal."""
import math

import numpy as np
import pytest

import ems_amd as E
from oracle.bridge import OracleEngine

import openems_pml_case as PC
from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu


def same_state(g, o, gop, oop):
    for side in (0, 1):
        for a, b in zip(g.upml_state(gop, side), o.upml_state(oop, side)):
            assert np.array_equal(bits(a), bits(b)), f"layer state side {side}"


def test_pml_face_fixture_on_device():
    gr, dt = PC.grid(), PC.dt()
    g = E.Engine(gr, dt, PC.lanes(gr, dt), curl_order=1)
    o = OracleEngine(gr, dt, PC.lanes(gr, dt), curl_order=1, threads=2)
    gop, gp = PC.build(g, gr)
    oop, opr = PC.build(o, gr)
    for n in (1, 99, 400, 500):
        g.run(n)
        o.run(n)
    same_fields(g, o)
    same_state(g, o, gop, oop)
    for name in PC.FILES:
        for a, b in zip(gp[name], opr[name]):
            assert np.array_equal(bits(a.values()), bits(o.probe_values(b))), name
    cols = PC.columns(gp, lambda p: p.values())
    for name in PC.FILES:
        e = PC.rel_l2_worst(PC.read_rows(name)[:, 1:], cols[name])
        assert e < 1e-10, f"{name}: relative L2 {e:.3e}"
    g.close()
    o.close()


def graded_box(kind, curl, materials, h_in_layer=False):
    cells = (14, 12, 13)

    def lines(n, s, w):
        return np.array([s * n * (q / n + w * math.sin(5.0 * q / n) / 5.0) for q in range(n + 1)])
    gr = E.Grid([lines(cells[0], 1e-3, 0.15), lines(cells[1], 1.1e-3, 0.1),
                 lines(cells[2], 0.9e-3, 0.12)])
    dt = E.TimeStep.courant(0.95).resolve(gr)
    mats = E.CellMaterials.vacuum(gr)
    block = E.Material(eps_r=(3.1, 2.9, 3.3), mu_r=(1.4, 1.3, 1.5), kappa=(0.05, 0.04, 0.06),
                       sigma_m=(200.0, 0.0, 100.0))
    mats.fill_box([(8, 14), (0, 5), (7, 13)], block)
    faces = [E.FaceCondition.Pec] * 6
    faces[3] = E.FaceCondition.Pmc
    lanes = E.coefficients(gr, mats, faces, dt)
    if kind == "gpu":
        e = E.Engine(gr, dt, lanes, curl_order=curl)
    else:
        e = OracleEngine(gr, dt, lanes, curl_order=curl, threads=3)
    layer_faces = [(E.Face.XMax, 6), (E.Face.YMin, 4), (E.Face.ZMax, 5)]
    if materials:
        n = gr.nodes()
        rng = np.random.default_rng(7)
        eps = 1.0 + rng.random((3,) + n)
        mu = 1.0 + 0.5 * rng.random((3,) + n)
        op = E.OpenEmsPml.with_materials(e, layer_faces, eps, mu)
    else:
        op = E.OpenEmsPml.new(e, layer_faces)
    e.add_operator(op)
    tau = 2.0e-11

    def pulse(t):
        x = (t - 4.0 * tau) / tau
        return math.exp(-x * x)
    e.add_operator(E.EdgeSource(pulse, horizon=3000).edge(E.Axis.Z, [6, 6, 6], 1e-3)
                   .edge(E.Axis.X, [5, 7, 6], 3e-4).edge(E.Axis.Y, [11, 2, 10], 5e-4))
    hs = E.EdgeSource(pulse, horizon=3000).kind(E.SourceKind.SoftH).edge(E.Axis.Y, [4, 5, 5], 2e-3)
    if h_in_layer:  # H_x at j = 1: inside the y_min layer
        hs.edge(E.Axis.X, [6, 1, 6], 1e-3)
    e.add_operator(hs)
    probes = [e.add_operator(E.Probe.voltage(E.Axis.Z, [12, 2, 11])),
              e.add_operator(E.Probe.current(E.Axis.X, [12, 1, 11])),
              e.add_operator(E.Probe.voltage(E.Axis.Y, [7, 6, 6]))]
    return e, op, probes


@pytest.mark.parametrize("curl", [0, 1], ids=["paired", "openems"])
@pytest.mark.parametrize("materials", [False, True], ids=["coefficients", "materials"])
def test_graded_overlapping_layers_match_oracle(curl, materials):
    g, gop, gp = graded_box("gpu", curl, materials)
    o, oop, op = graded_box("oracle", curl, materials)
    assert gop.edge_counts() == oop.edge_counts()
    assert min(gop.edge_counts()) > 500
    for n in (1, 20, 179, 300):
        g.run(n)
        o.run(n)
        same_fields(g, o)
        same_state(g, o, gop, oop)
    for a, b in zip(gp, op):
        assert np.array_equal(bits(a.values()), bits(o.probe_values(b)))
    assert np.max(np.abs(gp[0].values())) > 0.0  # the probe in the corner overlap moved
    w_g, w_o = g.step_with_energy(), o.step_with_energy()
    assert np.float64(w_g).tobytes() == np.float64(w_o).tobytes()
    g.close()
    o.close()


def test_layers_force_the_unfused_schedule():
    """A fused schedule asked for runs unfused with the layers, with the
    same bits."""
    g, gop, _ = graded_box("gpu", 0, False)
    h, hop, _ = graded_box("gpu", 0, False)
    h.set_schedule(2, 7, 0)
    g.run(60)
    h.run(60)
    same_fields(g, h)
    g.close()
    h.close()


def test_registration_rules():
    gr = E.Grid.uniform([10, 8, 8], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(gr)
    lanes = E.coefficients(gr, None, [E.FaceCondition.Pec] * 6, dt)
    e = E.Engine(gr, dt, lanes)
    e.add_operator(E.MurBoundary.new(e, E.Face.XMin, 299792458.0))
    with pytest.raises(E.EmsError, match="before"):
        e.add_operator(E.OpenEmsPml.with_materials(e, [(E.Face.XMax, 3)], 1.0, 1.0))
    e.close()
    e = E.Engine(gr, dt, lanes)
    e.add_operator(E.CpmlBoundary.new(e, [(E.Face.XMin, E.Cpml.with_cells(3))]))
    with pytest.raises(E.EmsError, match="CPML"):
        e.add_operator(E.OpenEmsPml.with_materials(e, [(E.Face.XMax, 3)], 1.0, 1.0))
    e.close()
    e = E.Engine(gr, dt, lanes)
    op = E.OpenEmsPml.with_materials(e, [(E.Face.XMax, 3)], 1.0, 1.0)
    e.add_operator(op)
    with pytest.raises(E.EmsError, match="CPML"):
        e.add_operator(E.CpmlBoundary.new(e, [(E.Face.XMin, E.Cpml.with_cells(3))]))
    with pytest.raises(E.EmsError, match="second"):
        e.add_operator(E.OpenEmsPml.with_materials(e, [(E.Face.XMin, 3)], 1.0, 1.0))
    b = E.EngineBatch([e])
    with pytest.raises(E.EmsError, match="unfused"):
        b.run(2)
    b.close()
    e.close()


def test_edge_listed_twice_is_refused():
    gr = E.Grid.uniform([10, 8, 8], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(gr)
    e = E.Engine(gr, dt, E.coefficients(gr, None, [E.FaceCondition.Pec] * 6, dt))
    op = E.OpenEmsPml.with_materials(e, [(E.Face.XMax, 3)], 1.0, 1.0)
    ax, nd, co = op.e
    op.e = (np.concatenate([ax, ax[:1]]), np.concatenate([nd, nd[:1]]),
            np.concatenate([co, co[:1]]))
    with pytest.raises(E.EmsError, match="twice"):
        e.add_operator(op)
    e.close()
