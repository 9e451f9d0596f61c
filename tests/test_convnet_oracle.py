"""Pins the early-stop network's CPU oracle (oracle/convnet.py) by
mathematics: its convolution against numpy's 1-D convolution, its
gradient against central finite differences in float64, its bf16
rounding against torch's, and its Adam step against torch.optim.Adam."""
import numpy as np
import pytest

from oracle import convnet as O


def test_causal_dilated_conv_matches_numpy_convolve():
    rng = np.random.default_rng(3)
    B, Ci, Co, T, d = 2, 3, 4, 50, 3
    x = rng.standard_normal((B, Ci, T))
    w = rng.standard_normal((Co, Ci, O.TAPS))
    b = rng.standard_normal(Co)
    z = O.conv(x, w, b, d)
    for bb in range(B):
        for o in range(Co):
            ref = np.full(T, b[o])
            for c in range(Ci):
                # the dilated kernel: np.convolve(x, h)[t] = sum_m h[m] x[t-m],
                # tap k at lag m = d*(K-1-k)
                h = np.zeros(d * (O.TAPS - 1) + 1)
                for k in range(O.TAPS):
                    h[d * (O.TAPS - 1 - k)] = w[o, c, k]
                ref += np.convolve(x[bb, c], h)[:T]
            assert np.allclose(z[bb, o], ref, rtol=1e-12, atol=1e-12)


def test_causality():
    """An output never depends on later inputs."""
    rng = np.random.default_rng(5)
    s = O.Shape(2, 2, 5, (1, 2, 4))
    p = rng.standard_normal(s.count) * 0.3
    x = rng.standard_normal((1, 2, 40))
    y0 = O.forward(s, p, x, round_w=False)
    x2 = x.copy()
    x2[:, :, 25:] += 1.0
    y1 = O.forward(s, p, x2, round_w=False)
    assert np.array_equal(y0[:, :, :25], y1[:, :, :25])
    assert not np.allclose(y0[:, :, 25:], y1[:, :, 25:])


@pytest.mark.parametrize("t0,spec", [(0, False), (7, False), (7, True)])
def test_gradient_matches_finite_differences(t0, spec):
    rng = np.random.default_rng(11)
    s = O.Shape(3, 2, 4, (1, 2, 3))
    p = rng.standard_normal(s.count) * 0.4
    x = rng.standard_normal((2, 3, 24))
    y = rng.standard_normal((2, 2, 24))
    spectral = None
    if spec:
        M = rng.standard_normal((10, 10))
        spectral = (M @ M.T / 10.0, 12, 0.7)   # symmetric, over samples [12, 22)
    loss, g = O.loss_and_grad(s, p, x, y, t0, round_w=False, spectral=spectral)
    h = 1e-6
    idx = rng.choice(s.count, 60, replace=False)
    # every bias and a few weights of each layer as well
    idx = np.unique(np.concatenate([idx, [s.boff[l] for l in range(s.L)], [s.woff[-1]]]))
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        lp, _ = O.loss_and_grad(s, pp, x, y, t0, round_w=False, spectral=spectral)
        lm, _ = O.loss_and_grad(s, pm, x, y, t0, round_w=False, spectral=spectral)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
    # the loss is the mean over t >= t0 only (plus the quadratic term)
    out = O.forward(s, p, x, round_w=False)
    want = np.mean((out[:, :, t0:] - y[:, :, t0:]) ** 2)
    if spectral is not None:
        K, s0, w = spectral
        d = out[:, :, s0:s0 + 10] - y[:, :, s0:s0 + 10]
        want += w * np.mean([dd @ K @ dd for dd in d.reshape(-1, 10)])
    assert np.isclose(loss, want, rtol=1e-12)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 4000),
                        np.float32([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, -0.0])])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(O.bf16(x).view(np.uint32), ref.view(np.uint32))
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert O.bf16(np.float32([1.0 + 2.0 ** -8]))[0] == 1.0


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    p0 = rng.standard_normal(50).astype(np.float32)
    grads = [rng.standard_normal(50).astype(np.float32) for _ in range(5)]
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    a = O.Adam(50, 1e-2)
    p = p0.copy()
    for g in grads:
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        p = a.step(p, g)
    assert np.allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-6)
