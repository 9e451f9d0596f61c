"""The early-stop network's device kernels against the CPU oracle
(oracle/convnet.py): forward pass, loss gradient and Adam steps (one with
a clipped gradient), on
the same parameters and inputs. The device accumulates in f32 in its own
order (multiply-adds contracted), so values agree to a tolerance: 2e-5 of
the largest magnitude for outputs, 2e-4 for gradients and 1e-5 for
parameters after a step."""
import numpy as np
import pytest

from oracle import convnet as O

pytestmark = pytest.mark.gpu


def _net(cin, cout, ch, dil, seed):
    from ems_amd.learn import ConvNet
    return ConvNet(cin, cout, ch, dil, seed=seed)


def _close(a, b, tol):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    assert err <= tol, f"max error {err:.3e} of the largest magnitude {scale:.3e}"


@pytest.mark.parametrize("shape", [(4, 2, 64, (1, 2, 4, 8, 16, 32), 2, 512),
                                   (3, 1, 16, (1, 3), 3, 37),
                                   (2, 3, 8, (2, 1, 5), 1, 1)])
def test_forward_matches_oracle(shape):
    cin, cout, ch, dil, B, T = shape
    n = _net(cin, cout, ch, dil, 1)
    p = n.params()
    s = O.Shape(cin, cout, ch, dil)
    assert p.size == s.count
    rng = np.random.default_rng(4)
    x = rng.standard_normal((B, cin, T)).astype(np.float32)
    got = n.forward(x)
    want = O.forward(s, p.astype(np.float64), x.astype(np.float64))
    _close(got, want, 2e-5)
    n.close()


@pytest.mark.parametrize("t0", [0, 100])
def test_gradient_matches_oracle(t0):
    dil = (1, 2, 4, 8, 16, 32)
    n = _net(4, 2, 64, dil, 7)
    s = O.Shape(4, 2, 64, dil)
    rng = np.random.default_rng(t0 + 1)
    # non-trivial parameters: perturb the initial ones, biases too
    p = n.params() + rng.standard_normal(s.count).astype(np.float32) * 0.02
    n.set_params(p)
    x = rng.standard_normal((3, 4, 256)).astype(np.float32)
    y = rng.standard_normal((3, 2, 256)).astype(np.float32)
    loss, g = n.loss_grad(x, y, t0)
    lo, go = O.loss_and_grad(s, p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), t0)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    for l in range(s.L):
        _close(g[s.woff[l]:s.boff[l]], go[s.woff[l]:s.boff[l]], 2e-4)
        _close(g[s.boff[l]:s.boff[l] + s.lo[l]], go[s.boff[l]:s.boff[l] + s.lo[l]], 2e-4)
    n.close()


def test_spectral_term_gradient_matches_oracle():
    dil = (1, 2, 4)
    n = _net(4, 2, 16, dil, 2)
    s = O.Shape(4, 2, 16, dil)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((3, 4, 96)).astype(np.float32)
    y = rng.standard_normal((3, 2, 96)).astype(np.float32)
    M = rng.standard_normal((64, 64))
    K = (M @ M.T / 64.0).astype(np.float32)
    n.set_spectral(K, 32, 0.8)
    p = n.params()
    loss, g = n.loss_grad(x, y, 20)
    lo, go = O.loss_and_grad(s, p.astype(np.float64), x.astype(np.float64), y.astype(np.float64),
                             20, spectral=(K.astype(np.float64), 32, 0.8))
    assert abs(loss - lo) <= 1e-4 * abs(lo), (loss, lo)
    _close(g, go, 2e-4)
    n.set_spectral(None, 0, 0.0)
    l2, _ = n.loss_grad(x, y, 20)
    l2o, _ = O.loss_and_grad(s, p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), 20)
    assert abs(l2 - l2o) <= 1e-5 * abs(l2o)
    n.close()


def test_adam_step_matches_oracle():
    dil = (1, 2, 4)
    n = _net(2, 2, 16, dil, 3)
    s = O.Shape(2, 2, 16, dil)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((6, 2, 64)).astype(np.float32)
    y = rng.standard_normal((6, 2, 64)).astype(np.float32)
    n.set_data(x, y)
    p = n.params()
    adam = O.Adam(s.count, 1e-3)
    # the third step clips the gradient to a tenth of its norm
    for step, (idx, clip) in enumerate((([0, 2, 4], 0.0), ([5, 1, 3], 0.0), ([1, 1, 0], -1.0))):
        lo, go = O.loss_and_grad(s, p.astype(np.float64), x[idx].astype(np.float64),
                                 y[idx].astype(np.float64), 10)
        gn = float(np.sqrt(np.sum(go ** 2)))
        if clip < 0:
            clip = 0.1 * gn
        loss = n.train_batch(idx, 10, 1e-3, clip)
        assert abs(loss - lo) <= 1e-5 * abs(lo), step
        assert abs(n.last_grad_norm - gn) <= 1e-4 * gn, (n.last_grad_norm, gn)
        p = adam.step(p, go.astype(np.float32), clip)
        _close(n.params(), p, 1e-5)
    n.close()


def test_training_reduces_the_loss():
    """A small net learns to continue damped sinusoids from their first
    quarter: the held-out loss falls by 4x in 1500 steps (the CPU oracle,
    trained the same way, reaches 6.7x)."""
    from ems_amd.learn import ConvNet
    rng = np.random.default_rng(21)
    T, N = 128, 32
    t = np.arange(T)

    def batch(n):
        f = rng.uniform(0.03, 0.08, n)[:, None]
        a = rng.uniform(0.01, 0.04, n)[:, None]
        ph = rng.uniform(0, 2 * np.pi, n)[:, None]
        y = np.exp(-a * t) * np.sin(2 * np.pi * f * t + ph)
        x = np.zeros((n, 2, T), np.float32)
        x[:, 0, :N] = y[:, :N]
        x[:, 1, :N] = 1.0
        return x, y[:, None, :].astype(np.float32)

    xs, ys = batch(512)
    xh, yh = batch(64)
    net = ConvNet(2, 1, 32, (1, 2, 4, 8, 16), seed=5)
    net.set_data(xs, ys)
    before = float(np.mean((net.forward(xh)[:, :, N:] - yh[:, :, N:]) ** 2))
    for it in range(1500):
        net.train_batch(rng.integers(0, 512, 32), N, 2e-3)
    after = float(np.mean((net.forward(xh)[:, :, N:] - yh[:, :, N:]) ** 2))
    assert after < before / 4.0, (before, after)
    net.close()


def test_invalid_arguments():
    from ems_amd.learn import ConvNet
    from ems_amd._native import EmsError
    n = ConvNet(2, 1, 8, (1, 2), seed=0)
    with pytest.raises(EmsError):
        n.forward(np.zeros((1, 3, 10), np.float32))
    with pytest.raises(EmsError):
        n.set_params(np.zeros(3, np.float32))
    n.set_data(np.zeros((4, 2, 10), np.float32), np.zeros((4, 1, 10), np.float32))
    with pytest.raises(EmsError):
        n.train_batch([0, 4], 0, 1e-3)
    n.close()
