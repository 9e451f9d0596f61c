"""The end-criterion loop (run::drive, src/engine/run.rs:424-540) on
scripted energies, with the exact step counts of the reference's own tests
(run.rs:588-673), for the product's loop and the oracle's restatement."""
import math

import pytest

import ems_amd as E
from oracle.bridge import oracle_drive_script


def outcome(w):
    lim = E.RunLimits(100, E.EndCriteria(10.0 ** (-30.0 / 10.0), check_interval=1))
    return E.drive_script(w, lim)


def test_negative_energy_is_divergence_not_decay():
    o = outcome([1.0, 4.0, 4.5, -0.2, 3.0])
    assert o.reason == E.StopReason.Diverged and o.steps == 4
    for last in (0.0, 1e-20, -1e-20):
        assert outcome([1.0, 4.0, 2.0, last]).reason == E.StopReason.EnergyDecay
    assert outcome([1.0, float("nan")]).reason == E.StopReason.Diverged
    assert outcome([1.0, 2.0, 1.5]).reason == E.StopReason.MaxTimesteps
    # The oracle's loop agrees.
    ratio = 10.0 ** (-3.0)
    assert oracle_drive_script([1.0, 4.0, 4.5, -0.2, 3.0], 100, ratio, 1) == (3, 4)
    assert oracle_drive_script([1.0, 2.0, 1.5], 100, ratio, 1) == (1, 100)
    assert oracle_drive_script([1.0, 4.0, 2.0, 0.0], 100, ratio, 1)[0] == 0


HALF = 60.0


def tail(k):
    return 0.01 * math.exp(-k * math.log(2.0) / HALF)


def x(k):
    return tail(k) + (1.0 if k < 20 else 0.0)


def run(settle, signal, steady):
    w = [1.0 if k < 40 else (1e-4 if steady else tail(k) * tail(k)) for k in range(5000)]
    sig = [signal(k) for k in range(100_000)] if signal else None
    c = E.EndCriteria(10.0 ** (-30.0 / 10.0), check_interval=1, settle=settle)
    return E.drive_script(w, E.RunLimits(100_000, c), sig)


def test_a_slow_tail_keeps_the_run_going():
    """run.rs:588-642."""
    plain = run(None, x, False)
    assert (plain.steps, plain.criterion_steps) == (41, 41)
    assert run(E.Settle(0.0), None, False).steps == 41
    o = run(E.Settle(0.0, max_extension=50.0), x, False)
    assert o.reason == E.StopReason.EnergyDecay
    assert (o.steps, o.criterion_steps) == (560, 41)
    assert run(E.Settle(0.0, tolerance=1e-5, max_extension=50.0), x, False).steps == 775
    assert run(E.Settle(0.0, max_extension=1.0), x, False).steps == 82
    assert run(E.Settle(0.0, max_extension=3.0), x, True).steps == 164
    f = run(E.Settle(0.05, max_extension=50.0), x, False)
    assert f.steps < o.steps


def test_excitation_end_and_zero_energy():
    lim = E.RunLimits(100, E.EndCriteria(1.0, check_interval=1), excitation_end=50.0)
    o = E.drive_script([0.0] * 10 + [1.0] * 200, lim)
    assert o.reason == E.StopReason.MaxTimesteps
    # Energy below threshold only counts after the excitation end.
    w = [1.0] * 30 + [1e-6] * 200
    lim = E.RunLimits(1000, E.EndCriteria(1e-4, check_interval=5), excitation_end=40.0)
    o = E.drive_script(w, lim)
    assert o.reason == E.StopReason.EnergyDecay and o.steps == 46
    assert oracle_drive_script(w, 1000, 1e-4, 5, 40.0) == (0, 46)
