"""Host side of the learned early stop (ems_amd/learn.py) on the CPU: the
trace format, the spectra assembled from observed samples plus a
decimated tail, and the residual-power rule. Pinned by mathematics: with
the true tail in place of a prediction, the assembled spectra equal the
spectra of the whole trace up to the decimated sum's error, which is tiny
for a band-limited signal sampled far above its band."""
import math

import numpy as np

from ems_amd import learn as L
from ems_amd import post

DT = 1e-12


def traces(n=4096, f0=2e9, tau=0.483e-9, t0=1.8e-9, r=50.0):
    """A port pair whose voltage is a modulated Gaussian with a damped
    ringing tail and whose current follows a 40-ohm load plus ringing."""
    t = (np.arange(n) + 1.0) * DT
    ring = 0.3 * np.exp(-(t - t0).clip(0) / 0.4e-9) * np.sin(2 * np.pi * 2.3e9 * (t - t0)) * (t > t0)
    v = np.cos(2 * np.pi * f0 * (t - t0)) * np.exp(-((t - t0) / tau) ** 2) + ring
    i = v / 40.0 - ring / r
    return v, i


def test_inputs_and_targets():
    fmt = L.TraceFormat(4096, 8, 1024, gain=3.0)
    fmt.check()
    v, i = traces()
    src = np.linspace(0, 1, fmt.T, dtype=np.float32)
    x = fmt.inputs(v, i, src)
    y = fmt.targets(v, i)
    assert x.shape == (1, 4, 512) and y.shape == (1, 2, 512)
    assert np.allclose(x[0, 0, :128], 3.0 * v[:1024:8].astype(np.float32))
    assert np.allclose(x[0, 1, :128], 3.0 * 50.0 * i[:1024:8], rtol=1e-6)
    assert not x[0, :2, 128:].any() and (x[0, 2, :128] == 1).all() and not x[0, 2, 128:].any()
    assert np.array_equal(x[0, 3], src)
    assert np.allclose(y[0, 0], v[::8].astype(np.float32))
    assert np.allclose(y[0, 1], (50.0 * i[::8]).astype(np.float32))


def test_true_tail_reproduces_the_full_spectra():
    fmt = L.TraceFormat(4096, 8, 1024)
    v, i = traces()
    freqs = L.FREQS_1_3_GHZ
    full = post.PortResult(freqs, post.dft(DT, DT, v, freqs), post.dft(0.5 * DT, DT, i, freqs),
                           50.0).s11()
    pred = fmt.targets(v, i)[0]
    early = L.early_port_result(fmt, freqs, DT, v, i, pred).s11()
    err = max(abs(a - b) for a, b in zip(early, full))
    assert err < 2e-3, err
    trunc = L.truncated_port_result(fmt, freqs, DT, v, i).s11()
    terr = max(abs(a - b) for a, b in zip(trunc, full))
    assert terr > 0.1, terr  # the observed part alone misses the pulse


def test_residual_rule():
    T = 512
    t = np.arange(T)
    fast = np.exp(-t / 20.0) * np.sin(t / 3.0)
    slow = np.exp(-t / 400.0) * np.sin(t / 3.0)
    pred = np.stack([np.stack([fast, fast]), np.stack([slow, slow])]).astype(np.float32)
    r = L.residual_db(pred)
    assert r[0] < -40.0 < r[1]


def test_spectral_matrix_is_the_weighted_spectrum():
    """d^T K d equals the weighted power of d's DFT at the chosen
    frequencies, with weights 1/|S(f)|^2 normalised to mean 1."""
    fmt = L.TraceFormat(512, 8, 128)
    w = 0.5 * np.cos(2 * np.pi * 2e9 * (np.arange(512) * DT - 0.2e-9)) * \
        np.exp(-((np.arange(512) * DT - 0.2e-9) / 0.05e-9) ** 2)
    freqs = [1e9, 2e9, 3e9]
    K = L.spectral_matrix(fmt, w, DT, freqs).astype(np.float64)
    n = fmt.T - fmt.N
    assert K.shape == (n, n) and np.allclose(K, K.T)
    ev = np.linalg.eigvalsh(K)
    assert ev.min() > -1e-6 * ev.max()  # positive semidefinite (to f32 rounding)
    rng = np.random.default_rng(1)
    d = rng.standard_normal(n)
    S = np.abs(np.array(post.dft(0.0, DT, w, freqs)))
    wt = 1.0 / S ** 2
    wt /= wt.mean()
    t = np.arange(n) * 8 * DT
    want = sum(wf * abs(np.sum(d * np.exp(-2j * np.pi * f * t))) ** 2 for f, wf in zip(freqs, wt))
    assert np.isclose(d @ K @ d, want / (len(freqs) * n), rtol=1e-5)


def test_batched_early_spectra_match_the_per_model_form():
    """early_s11 (many models, DFT matrices shared by models of one time
    step) against early_port_result model by model: the same sums up to the
    transform's rounding."""
    fmt = L.TraceFormat(4096, 8, 1024)
    rng = np.random.default_rng(3)
    v0, i0 = traces()
    v = np.stack([v0 * (1 + 0.1 * q) for q in range(4)])
    i = np.stack([i0 * (1 - 0.05 * q) for q in range(4)])
    pred = np.stack([fmt.targets(v[q], i[q])[0] for q in range(4)])
    pred += 1e-3 * rng.standard_normal(pred.shape).astype(np.float32)
    dts = [DT, DT, 2 * DT, DT]
    got = L.early_s11(fmt, L.FREQS_1_3_GHZ, dts, v, i, pred)
    for q in range(4):
        ref = L.early_port_result(fmt, L.FREQS_1_3_GHZ, dts[q], v[q], i[q], pred[q]).s11()
        assert max(abs(a - b) for a, b in zip(got[q], ref)) < 1e-9
