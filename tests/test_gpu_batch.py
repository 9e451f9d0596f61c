"""Batched stepping of uneven models: each model of a batch gives exactly
what its own engine gives alone (port traces, fields, run-until outcome),
for the lockstep run loop and for runs of a subset."""
import numpy as np
import pytest

import ems_amd as E
from ems_amd import workloads as W

from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu

SIZES = [(8.0, 20.0, 1.0, 8), (10.0, 24.0, 2.5, 10), (9.0, 22.0, 0.0, 9), (12.0, 30.0, 3.0, 8)]


def models():
    return [W.patch_model(*s) for s in SIZES]


def engine(m):
    e = E.Engine(m.grid, m.dt, m.lanes)
    p = m.build_ops(e)
    return e, p


def limits():
    # -80 dB checked every 20 steps: these four models stop at three
    # different checks (3781, 3801 and 4501 steps on the oracle), so the
    # staggered stops of the batch run loop are exercised.
    return E.RunLimits(6000, E.EndCriteria(1e-8, check_interval=20),
                       excitation_end=E.Waveform.gaussian(2e9, 1e9).duration())


@pytest.mark.parametrize("prefetch", [False, True], ids=["noprefetch", "prefetch"])
@pytest.mark.parametrize("tile_rows", [3, 7, "packed"])
def test_batch_run_until_matches_single_runs(tile_rows, prefetch):
    ms = models()
    single = [engine(m) for m in ms]
    outs = []
    for e, p in single:
        e.set_schedule(1, 7 if tile_rows == "packed" else tile_rows, 0)
        outs.append(e.run_until(limits()))
    batch = [engine(m) for m in ms]
    b = E.EngineBatch([e for e, _ in batch])
    if tile_rows == "packed":
        b.set_schedule(7, 0, prefetch=prefetch, packed=True)
    else:
        b.set_schedule(tile_rows, 0, prefetch=prefetch)
    bouts = b.run_until(limits())
    for (e, p), (f, q), o, bo in zip(single, batch, outs, bouts):
        assert (o.reason, o.steps, o.timestep) == (bo.reason, bo.steps, bo.timestep)
        assert np.float64(o.energy).tobytes() == np.float64(bo.energy).tobytes()
        assert np.array_equal(bits(p.voltage()), bits(q.voltage()))
        assert np.array_equal(bits(p.current()), bits(q.current()))
        same_fields(e, f)
    assert len({o.steps for o in bouts}) > 1 and all(o.reason == E.StopReason.EnergyDecay
                                                     for o in bouts)
    b.close()


@pytest.mark.parametrize("packed", [False, True])
def test_batch_without_graphs_matches(packed):
    ms = models()
    runs = []
    for graphs in (True, False):
        batch = [engine(m) for m in ms]
        b = E.EngineBatch([e for e, _ in batch])
        b.set_schedule(7, 2 if packed else 0, packed=packed)
        b.set_graphs(graphs)
        b.run(37)  # K = 16 graphs twice, then single steps
        b.run(5, [True, False, True, False])
        runs.append((batch, b))
    for (e, p), (f, q) in zip(runs[0][0], runs[1][0]):
        same_fields(e, f)
        assert np.array_equal(bits(p.voltage()), bits(q.voltage()))
    for batch, b in runs:
        b.close()
        for e, _ in batch:
            e.close()


@pytest.mark.parametrize("packed", [False, True])
def test_batch_subset_runs(packed):
    ms = models()
    single = [engine(m) for m in ms]
    for e, _ in single:
        e.run(77)
    batch = [engine(m) for m in ms]
    b = E.EngineBatch([e for e, _ in batch])
    b.set_schedule(7, 2 if packed else 0, packed=packed)
    b.run(40)
    b.run(37)
    for (e, p), (f, q) in zip(single, batch):
        same_fields(e, f)
        assert np.array_equal(bits(p.voltage()), bits(q.voltage()))
    # Only models 1 and 3 advance; the others stay put.
    act = [False, True, False, True]
    b.run(25, act)
    for q, ((e, p), (f, _)) in enumerate(zip(single, batch)):
        if act[q]:
            e.run(25)
        same_fields(e, f)
    b.close()


def test_batch_refuses_models_out_of_step_and_duplicates():
    """Models of a batch step in lockstep: two models whose step parities
    differ (one has taken an odd number of steps more, as after a run loop
    with an odd check interval that stopped them at different checks) are
    refused, and so is one engine listed twice."""
    ms = models()[:2]
    (a, _), (b, _) = engine(ms[0]), engine(ms[1])
    a.run(1)
    batch = E.EngineBatch([a, b])
    with pytest.raises(E.EmsError, match="lockstep"):
        batch.run(2)
    batch.close()
    b.run(1)  # back in step: the same batch object works
    batch = E.EngineBatch([a, b])
    batch.run(2)
    batch.close()
    with pytest.raises(E.EmsError):
        E.EngineBatch([a, a])
    a.close()
    b.close()
