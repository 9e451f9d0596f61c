"""Parity of the device TF/SF plane wave with the oracle: fields, probes
and the auxiliary grid bit for bit (a plane wave runs unfused; the fused
request falls back)."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import planewave as PW

import cases as K
from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=[(0, 0, 0), (1, 7, 0)], ids=lambda s: "mode%d_tj%d_xc%d" % s)
def schedule(request):
    f, t, c = request.param
    old = dict(K.SCHEDULE)
    K.SCHEDULE.update(fused=f, tile_rows=t, x_chunks=c)
    yield
    K.SCHEDULE.update(old)


@pytest.mark.parametrize("d,pol", [([1.0, 0.0, 0.0], [0.0, 0.0, 1.0]),
                                   ([1.0, 1.0, 1.0], [1.0, -1.0, 0.0]),
                                   ([-0.3, 0.5, -1.0], [1.0, 0.6, 0.0])])
def test_planewave_in_cpml_box(d, pol):
    def build(kind):
        g = E.Grid.uniform([24] * 3, [1e-3] * 3)
        e = K.make(kind, g, E.CellMaterials(g, E.Material.dielectric(2.0)), threads=2)
        e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(6)) for f in E.Face]))
        p1 = e.add_operator(E.Probe.voltage(E.Axis.Z, [12, 12, 12]))
        spec = PW.PlaneWaveSpec.new([8] * 3, [16] * 3, d, pol).with_background(2.0, 1.0) \
            .with_match_frequency(8e9)
        pw = e.add_operator(PW.PlaneWave.new(e, spec, K.sine_pulse(8e9, 4e9)))
        p2 = e.add_operator(E.Probe.current(E.Axis.Y, [12, 12, 12]))
        return e, pw, p1, p2
    (g, pg, a1, a2), (o, po, b1, b2) = build("gpu"), build("oracle")
    for n in (1, 99, 200):
        g.run(n)
        o.run(n)
    assert np.array_equal(bits(a1.values()), bits(o.probe_values(b1)))
    assert np.array_equal(bits(a2.values()), bits(o.probe_values(b2)))
    for x, y in zip(pg._aux(), po._aux()):
        assert np.array_equal(bits(x), bits(y))
    same_fields(g, o)


def test_planewave_graded_mesh_with_lumped_load():
    """Graded mesh, oblique incidence, a lumped resistor inside the box."""
    from ems_amd.lumped import LumpedBox, LumpedElement
    x, lines = 0.0, []
    for i in range(21):
        lines.append(x)
        x += 1e-3 * (1.0 + 0.2 * math.sin(i * 0.5))

    def build(kind):
        g = E.Grid([lines] * 3)
        e = K.make(kind, g, threads=3)
        spec = PW.PlaneWaveSpec.new([5] * 3, [15] * 3, [1.0, 0.5, 0.2], [0.0, -0.4, 1.0])
        e.add_operator(PW.PlaneWave.new(e, spec, K.sine_pulse(10e9, 5e9)))
        h = e.add_operator(LumpedElement.resistor(LumpedBox(E.Axis.Z, [10, 10, 8], [10, 10, 11]),
                                                  50.0).record_current())
        return e, h
    (g, hg), (o, ho) = build("gpu"), build("oracle")
    g.run(180)
    o.run(180)
    assert np.array_equal(bits(hg.current()), bits(ho.current()))
    same_fields(g, o)
