"""The patch-antenna sweep path against the oracle: seeded patch models
(2 mm FR-4 on a PEC ground, 8-cell CPML on five faces, a 50-ohm lumped port
excited by a 1-3 GHz Gaussian) run to -40 dB energy decay by the batched
device kernels (packed short rows and tiled, graph-captured, the batch run
loop) and by the oracle with the reference's run loop (src/engine/run.rs:
424-540, restated below): port voltage and current bit for bit, the same
stop step, the same f64 energy at the deciding check, S11 at 201 points
over 1-3 GHz identical (src/port/mod.rs:501-544, src/post/port.rs:166-210).
Then a 32-model sweep through properties that hold at any size: every model
stops by energy decay one step after a check below -40 dB, past the end of
the excitation; |S11| near 1 (these patches barely radiate below 3 GHz);
and two of its models, run inside the larger batch, equal their oracle
runs."""
import os

import numpy as np
import pytest

import ems_amd as E
from ems_amd import workloads as W

from test_gpu_parity import bits

pytestmark = pytest.mark.gpu

FREQS = [1e9 + 2e9 * q / 200.0 for q in range(201)]
MODELS = [(8.0, 20.0, 1.0, 8), (10.0, 24.0, 2.5, 10)]


def limits():
    return E.RunLimits(20000, E.EndCriteria.energy_decay_db(-40.0),
                       excitation_end=E.Waveform.gaussian(2e9, 1e9).duration())


def oracle_threads():
    return max(1, min(8, len(os.sched_getaffinity(0))))


def oracle_run_until(m, lim):
    """run::drive (run.rs:424-540) on the oracle: an energy step every
    check_interval steps (W^n of the state before it), stop one step after
    the deciding check once t > excitation_end and W/W_max < ratio."""
    from oracle.bridge import OracleEngine
    o = OracleEngine(m.grid, m.dt, m.lanes, threads=oracle_threads())
    port = m.build_ops(o)
    ec = lim.end_criteria
    w_max, steps, w = 0.0, 0, 0.0
    while steps < lim.max_timesteps:
        t = o.timestep() * o.dt
        w = o.step_with_energy()
        steps += 1
        w_max = max(w_max, w)
        assert np.isfinite(w)
        if t > lim.excitation_end and w_max > 0.0 and w / w_max < ec.energy_ratio:
            break
        n = min(ec.check_interval - 1, lim.max_timesteps - steps)
        o.run(n)
        steps += n
    return o, port, steps, w


@pytest.fixture(scope="module")
def oracle_runs():
    lim = limits()
    out = []
    for s in MODELS:
        m = W.patch_model(*s)
        o, port, steps, w = oracle_run_until(m, lim)
        # (the port reads the oracle's records: copied out before it closes)
        rec = (port.voltage().copy(), port.current().copy())
        out.append((m, rec, steps, w, port.analyze(FREQS).s11()))
        o.close()
    return out


def batch_of(models):
    engines, ports = [], []
    for m in models:
        e = E.Engine(m.grid, m.dt, m.lanes)
        ports.append(m.build_ops(e))
        engines.append(e)
    return engines, ports, E.EngineBatch(engines)


@pytest.mark.parametrize("sched", ["packed", "tiled7_x2", "tiled3"])
def test_batched_sweep_matches_oracle(oracle_runs, sched):
    models = [r[0] for r in oracle_runs]
    engines, ports, b = batch_of(models)
    if sched == "packed":
        b.set_schedule(7, 0, packed=True)
    elif sched == "tiled7_x2":
        b.set_schedule(7, 2)
    else:
        b.set_schedule(3, 0)
    outs = b.run_until(limits())
    for (m, po, steps, w, s11), q, o in zip(oracle_runs, ports, outs):
        assert o.reason == E.StopReason.EnergyDecay
        assert o.steps == steps, (m.name, o.steps, steps)
        assert np.float64(o.energy).tobytes() == np.float64(w).tobytes()
        assert np.array_equal(bits(q.voltage()), bits(po[0])), m.name
        assert np.array_equal(bits(q.current()), bits(po[1])), m.name
        assert q.analyze(FREQS).s11() == s11
    b.close()
    for e in engines:
        e.close()


def test_sweep_of_32_models_properties(oracle_runs):
    lim = limits()
    models = W.patch_sweep(30, seed=99) + [r[0] for r in oracle_runs]
    engines, ports, b = batch_of(models)
    b.set_schedule(7, 0, packed=True)
    outs = b.run_until(lim)
    ec = lim.end_criteria
    for m, p, o in zip(models, ports, outs):
        assert o.reason == E.StopReason.EnergyDecay, m.name
        assert o.ratio_db < -40.0
        # the deciding check is an energy step (every check_interval steps),
        # the run stops one step after it, past the end of the excitation
        assert (o.steps - 1) % ec.check_interval == 0
        assert (o.steps - 1) * m.dt > lim.excitation_end
        v, i = p.voltage(), p.current()
        assert v.size == i.size == o.steps and np.all(np.isfinite(v)) and np.all(np.isfinite(i))
        # The lumped port's power waves of a trace cut at -40 dB are not
        # exactly passive (the oracle's runs of these models reach 1.0085);
        # a bound that a wrong trace would break.
        s = np.abs(np.asarray(p.analyze(FREQS).s11()))
        assert np.all(s <= 1.05) and np.all(s > 0.5), (m.name, float(s.min()), float(s.max()))
    # the two oracle-run models, inside the larger batch
    for (m, po, steps, w, s11), p, o in zip(oracle_runs, ports[-2:], outs[-2:]):
        assert o.steps == steps
        assert np.array_equal(bits(p.voltage()), bits(po[0]))
        assert np.array_equal(bits(p.current()), bits(po[1]))
        assert p.analyze(FREQS).s11() == s11
    b.close()
    for e in engines:
        e.close()
