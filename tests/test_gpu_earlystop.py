"""The trained early-stop network (ems_amd/earlystop_patch.npz) against the
method without it, on fresh solver traces: twelve patch antennas of a
seed the network never saw, run on the device to the horizon. The
early-stopped spectra (1024 observed steps plus the predicted tail) must
be far closer to the whole-trace S11 than the spectra of the observed
steps alone. (Held-out figures of the committed network on 512 traces:
median |dS11| 0.006, 90th percentile 0.026, largest 0.44; truncation:
median 1.01.)"""
import numpy as np
import pytest

from ems_amd import learn as L
from ems_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_network_gains_over_truncation():
    es = L.EarlyStop.load()
    models = W.patch_sweep(12, seed=99)
    v, i, dts = L.solver_traces(models, es.fmt.horizon)
    rep = L.held_out_report(es, v, i, dts)
    net = np.array(rep["net_err"])
    trunc = np.array(rep["trunc_err"])
    assert np.median(trunc) > 0.5, rep["trunc_dS11_median"]
    assert np.median(net) < 0.05 * np.median(trunc), (np.median(net), np.median(trunc))
    assert np.median(net) < 0.05 and net.max() < 0.6, net
    # the prediction decays within the horizon (no fallback on this family)
    assert rep["residual_db_max"] < -40.0


def test_early_stop_sweep_uses_the_network_and_falls_back():
    """A small early-stopped sweep: no fallback at the trained threshold;
    with a threshold no prediction can meet, every model is run in full
    and its S11 is the full run's."""
    import ems_amd as E
    es = L.EarlyStop.load()
    models = W.patch_sweep(4, seed=5)
    lim = E.RunLimits(20000, E.EndCriteria.energy_decay_db(-40.0),
                      excitation_end=E.Waveform.gaussian(2e9, 1e9).duration())
    res = L.early_stop_sweep(es, models, lim)
    assert not any(res.fallback) and all(s == es.fmt.observed for s in res.steps)
    es.threshold_db = -400.0
    fb = L.early_stop_sweep(es, models, lim)
    assert all(fb.fallback) and all(s > es.fmt.observed for s in fb.steps)
    engines, ports = L._batch_engines(models)
    b = E.EngineBatch(engines)
    L._schedule(b, models, 7)
    b.run_until(lim)
    for q, p in enumerate(ports):
        full = p.analyze(L.FREQS_1_3_GHZ).s11()
        assert fb.s11[q] == full
    b.close()
