"""The running DFT on the device (ems_gpu_add_dft / _update / _result,
RunningDft of src/post/dft.rs:75-190 with re-anchoring every 1024
samples) against the host's ems_dft of the copied-back series: bit for
bit, whether updated once at the end or many times during the run, and
S11 from the device spectra equal to S11 from the host's."""
import numpy as np
import pytest

import ems_amd as E
from ems_amd import post
from ems_amd import workloads as W

pytestmark = pytest.mark.gpu

FREQS = [1e9 + 2e9 * q / 200.0 for q in range(201)]


def test_device_dft_equals_host_dft():
    m = W.patch_model(10.0, 24.0, 2.5, 10)
    e = E.Engine(m.grid, m.dt, m.lanes)
    p = m.build_ops(e)
    p.track(FREQS)  # from the first step: updated as the run goes
    other = FREQS[::7]
    for chunk in (300, 724, 1, 1999, 1000):  # crosses anchors at 1024, 2048, 3072
        e.run(chunk)
        r = p.analyze_device(FREQS)
    late = p.analyze_device(other)  # started after the run: catches up at once
    assert len(p.voltage()) == 4024
    host = p.analyze(FREQS)
    assert r.voltage() == host.voltage() and r.current() == host.current()
    assert r.s11() == host.s11()
    hl = p.analyze(other)
    assert late.voltage() == hl.voltage() and late.current() == hl.current()
    # and the host transform is the plain sum to rounding
    v = p.voltage()
    t = p.voltage_t0() + p.dt * np.arange(v.size)
    naive = [complex(np.sum(v * np.exp(-2j * np.pi * f * t)) * p.dt) for f in FREQS[:5]]
    for a, b in zip(host.voltage()[:5], naive):
        assert abs(a - b) <= 1e-9 * max(abs(b), 1e-30)
    e.close()


def test_batch_sweep_spectra_from_the_device():
    ms = [W.patch_model(*s) for s in [(8.0, 20.0, 1.0, 8), (12.0, 30.0, 3.0, 8)]]
    engines, ports = [], []
    for m in ms:
        e = E.Engine(m.grid, m.dt, m.lanes)
        ports.append(m.build_ops(e))
        engines.append(e)
    for p in ports:
        p.track(FREQS)
    b = E.EngineBatch(engines)
    b.set_schedule(7, 0, packed=True)
    lim = E.RunLimits(20000, E.EndCriteria.energy_decay_db(-40.0),
                      excitation_end=E.Waveform.gaussian(2e9, 1e9).duration())
    b.run_until(lim)
    for p in ports:
        assert p.analyze_device(FREQS).s11() == p.analyze(FREQS).s11()
    b.close()
    for e in engines:
        e.close()
