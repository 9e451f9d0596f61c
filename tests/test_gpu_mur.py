"""The Mur boundary on the device against the oracle, bit for bit: an
all-Mur graded box (face edges and domain corners) with a dipole pulse, on
every kernel schedule; openEMS's operation order and corner rule with a
held face; and the registration rule (before the operators that act on E
serially)."""
import math

import numpy as np
import pytest

import ems_amd as E
from oracle.bridge import OracleEngine

from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu

C0 = 299792458.0
TAU = 2.8e-11
SCHEDULES = [(0, 0, 0), (1, 7, 0), (2, 3, 2), (3, 7, 0), (4, 7, 0), (6, 8, 0)]


def dgaussian(tau):
    def f(t):
        x = (t - 4.0 * tau) / tau
        return -2.0 * x * math.exp(-x * x)
    return f


def build(kind, sched, openems=False, hold=0):
    cells = 12
    lines = lambda s: np.array([s * cells * (q / cells + 0.2 * math.sin(6.0 * q / cells) / 6.0)  # noqa
                                for q in range(cells + 1)])
    g = E.Grid([lines(1e-3), lines(1.1e-3), lines(0.9e-3)])
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    if kind == "gpu":
        e = E.Engine(g, dt, lanes)
        e.set_schedule(*sched)
    else:
        e = OracleEngine(g, dt, lanes, threads=2)
    m = E.MurBoundary.with_faces(e, [(f, C0 if f < 4 else 0.8 * C0) for f in E.Face])
    if openems:
        m.openems_order(True).openems_corners(True)
    if hold:
        m.hold_faces([E.Face.XMin], hold)
    e.add_operator(m)
    e.add_operator(E.EdgeSource(dgaussian(TAU), horizon=2000).edge(E.Axis.Z, [6, 6, 6], 1.0)
                   .edge(E.Axis.X, [4, 7, 6], 0.3))
    p = e.add_operator(E.Probe.voltage(E.Axis.Y, [1, 5, 5]))
    q = e.add_operator(E.Probe.voltage(E.Axis.Z, [0, 6, 6]))  # a face edge
    return e, p, q


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "mode%d_tj%d_xc%d" % s)
@pytest.mark.parametrize("variant", ["ems", "openems_hold"])
def test_mur_box_matches_oracle(sched, variant):
    kw = {} if variant == "ems" else {"openems": True, "hold": 37}
    g, gp, gq = build("gpu", sched, **kw)
    o, op, oq = build("oracle", sched, **kw)
    for n in (1, 36, 80, 200):
        g.run(n)
        o.run(n)
    for a, b in ((gp, op), (gq, oq)):
        assert np.array_equal(bits(a.values()), bits(o.probe_values(b)))
    assert np.max(np.abs(gq.values())) > 0.0  # the face edge moved
    same_fields(g, o)
    assert np.float64(g.energy()).tobytes() == np.float64(o.energy()).tobytes()
    g.close()
    o.close()


def test_mur_registered_after_a_voltage_probe_is_refused():
    g = E.Grid.uniform([8, 6, 6], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(g)
    e = E.Engine(g, dt, E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt))
    e.add_operator(E.Probe.voltage(E.Axis.Z, [3, 3, 3]))
    with pytest.raises(E.EmsError, match="before"):
        e.add_operator(E.MurBoundary.new(e, E.Face.XMin, C0))
    e.close()
