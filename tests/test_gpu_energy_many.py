"""ems_gpu_energy_many (the batch run loop's energy step) against each
engine's own evaluation, bit for bit: engines of different sizes, with and
without CPML, stepped to different times, every one active or a subset,
both curl orders, one launch per kernel over all of them; and the engines
it evaluates one by one (mixed row widths; a dispersive medium)."""
import numpy as np
import pytest

import ems_amd as E
from ems_amd.dispersive import Dispersive, Field, Pole

pytestmark = pytest.mark.gpu

SHAPES = [((20, 16, 12), True, 37), ((24, 18, 30), False, 41), ((12, 30, 20), True, 58),
          ((33, 9, 40), True, 7), ((16, 16, 64), False, 80)]


def engine(n, cpml, steps, seed, curl=0, medium=False):
    rng = np.random.default_rng(seed)
    g = E.Grid([np.concatenate([[0.0], np.cumsum(0.5e-3 + 1e-3 * rng.random(k))]) for k in n])
    dt = 0.95 * E.courant_limit(g)
    e = E.Engine(g, dt, E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt), curl_order=curl)
    if cpml:
        e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(4)) for f in E.Face]))
    if medium:
        d = Dispersive.new()
        d.fill_cells(g, [(3, 7), (3, 7), (3, 7)], Field.Electric, [Pole.debye(2.0, 1e-11)])
        e.add_operator(d)
    e.add_operator(E.EdgeSource(E.Waveform.gaussian(20e9, 15e9))
                   .edge(E.Axis.Z, [k // 2 for k in n], 1e-3))
    e.run(steps)
    return e


def bits(v):
    return np.float64(v).tobytes()


@pytest.mark.parametrize("curl", [0, 1], ids=["paired", "openems"])
def test_energy_many_matches_single_evaluations(curl):
    es = [engine(n, cp, st, q, curl) for q, (n, cp, st) in enumerate(SHAPES)]
    want = [e.energy() for e in es]
    assert all(w > 0.0 for w in want)
    assert [bits(v) for v in E.energy_many(es)] == [bits(v) for v in want]
    act = [q % 2 == 0 for q in range(len(es))]
    got = E.energy_many(es, act)
    for q, (a, g, w) in enumerate(zip(act, got, want)):
        assert bits(g) == (bits(w) if a else bits(0.0)), q
    for e in es:  # the evaluations left every state as it was
        e.run(5)
    assert [bits(v) for v in E.energy_many(es)] == [bits(e.energy()) for e in es]
    for e in es:
        e.close()


def test_energy_many_one_by_one_where_it_cannot_batch():
    es = [engine((14, 12, 20), True, 30, 1), engine((14, 12, 100), False, 30, 2),
          engine((16, 12, 18), False, 25, 3, medium=True)]
    want = [e.energy() for e in es]
    assert [bits(v) for v in E.energy_many(es)] == [bits(v) for v in want]
    assert [bits(v) for v in E.energy_many(es[:2])] == [bits(v) for v in want[:2]]
    assert [bits(v) for v in E.energy_many([es[0], es[2]])] == [bits(want[0]), bits(want[2])]
    for e in es:
        e.close()
