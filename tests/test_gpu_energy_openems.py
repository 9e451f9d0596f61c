"""openEMS's energy measure and end criterion on the device: the measure
bit for bit against the oracle (layers with their corrections, H sources,
both curl orders, a box without layers), and the openEMS fixture
``pml_end`` run by the device's run loop on openEMS's schedule (checks
after every Nyquist interval and once after half the Gaussian) stopping
where openEMS stopped, at the oracle's level, with the files' probes."""
import numpy as np
import pytest

import ems_amd as E
from oracle.bridge import OracleEngine

import openems_pml_case as PC
import openems_pml_end_case as PE
from test_gpu_parity import bits, same_fields
from test_gpu_upml import graded_box

pytestmark = pytest.mark.gpu


def same_energy(g, o):
    a, b = g.energy_openems(), o.energy_openems()
    assert np.float64(a).tobytes() == np.float64(b).tobytes(), (a, b)
    return a


@pytest.mark.parametrize("curl", [0, 1], ids=["paired", "openems"])
def test_measure_with_layers_and_h_source_matches_oracle(curl):
    g, gop, _ = graded_box("gpu", curl, True, h_in_layer=True)
    o, oop, _ = graded_box("oracle", curl, True, h_in_layer=True)
    for n in (0, 1, 37, 120, 200):
        g.run(n)
        o.run(n)
        w = same_energy(g, o)
    assert w > 0.0
    same_fields(g, o)
    g.close()
    o.close()


def test_measure_without_layers_matches_oracle():
    gr = E.Grid([np.array([0.0, 1.0, 2.5, 3.1, 4.0, 5.2, 6.0, 7.3]) * 1e-3,
                 np.array([0.0, 0.9, 2.0, 3.0, 4.4, 5.0]) * 1e-3,
                 np.array([0.0, 1.1, 2.0, 3.3, 4.0, 5.0, 6.1]) * 1e-3])
    dt = 0.9 * E.courant_limit(gr)
    lanes = E.coefficients(gr, None, [E.FaceCondition.Pec] * 6, dt)
    g = E.Engine(gr, dt, lanes)
    o = OracleEngine(gr, dt, lanes, threads=2)
    for e in (g, o):
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(20e9, 15e9)).edge(E.Axis.Z, [3, 2, 3], 1e-3))
    for n in (5, 40, 80):
        g.run(n)
        o.run(n)
        same_energy(g, o)
    g.close()
    o.close()


def test_pml_end_stops_where_openems_stops():
    gr, dt = PE.grid(), PE.dt()
    g = E.Engine(gr, dt, PE.lanes(gr, dt), curl_order=1)
    o = OracleEngine(gr, dt, PE.lanes(gr, dt), curl_order=1, threads=4)
    gop, gp = PE.build(g, gr)
    oop, opr = PE.build(o, gr)
    k, extra, pint = PE.limits(dt)
    lim = E.RunLimits(PE.MAX_STEPS, E.EndCriteria(PE.END_RATIO, k), None,
                      check_after_interval=True, extra_check=extra, energy_measure=1)
    out = g.run_until(lim)
    steps, db = PE.drive(o, o.energy_openems, dt)
    assert out.reason == E.StopReason.EnergyDecay
    assert out.steps == steps == PE.STOP_STEPS
    assert abs(out.ratio_db - db) < 1e-12
    assert abs(out.ratio_db - PE.STOP_DB) < 5e-5
    same_fields(g, o)
    g.run(1)
    o.run(1)
    for name in PE.FILES:
        rows = PE.read_rows(name)
        got = np.stack([PE.sampled(name, p.values(), pint, rows.shape[0]) for p in gp[name]], axis=1)
        assert PC.rel_l2_worst(rows[:, 1:], got) < 1e-10, name
        for a, b in zip(gp[name], opr[name]):
            assert np.array_equal(bits(a.values()), bits(o.probe_values(b)))
    g.close()
    o.close()


def test_batch_refuses_openems_schedule():
    gr = E.Grid.uniform([8, 6, 6], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(gr)
    lanes = E.coefficients(gr, None, [E.FaceCondition.Pec] * 6, dt)
    e = E.Engine(gr, dt, lanes)
    b = E.EngineBatch([e])
    lim = E.RunLimits(100, E.EndCriteria(1e-5, 10), None, check_after_interval=True,
                      energy_measure=1)
    with pytest.raises(E.EmsError):
        b.run_until(lim)
    b.close()
    e.close()
