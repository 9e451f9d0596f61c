"""The learned early stop (include/ems_learn.h).

A causal dilated 1-D convolution network, trained and run on the device by
the library's own kernels, that maps the first part of a lumped port's
voltage/current traces (and the known excitation) to the whole trace. An
early-stopped solve runs the solver for `observed` steps, predicts the
rest of the horizon decimated by `decimation`, and assembles the port
spectra from the observed full-rate samples plus the predicted tail. When
the predicted port signal has not decayed by the end criterion within the
horizon, the caller falls back to the full run.

No counterpart in the reference: its run loop (src/engine/run.rs:424-540)
always steps until the end criterion, and its port analysis
(src/engine/port.rs:209-219) transforms the recorded traces; the tail
assembly below is that same DFT (src/post/dft.rs:213-222) applied to the
observed samples and to the predicted decimated samples at their times.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from . import post
from ._native import EmsError

FREQS_1_3_GHZ = [1e9 + 2e9 * q / 200.0 for q in range(201)]


class ConvNet:
    """An ems_cn network: `channels` wide, one layer per dilation (9 taps
    each), f32 master weights and their bf16 copy on the device."""

    def __init__(self, cin: int, cout: int, channels: int = 64,
                 dilation: Sequence[int] = (1, 2, 4, 8, 16, 32), seed: int = 0):
        self.cin, self.cout, self.channels = int(cin), int(cout), int(channels)
        self.dilation = [int(d) for d in dilation]
        dil = np.ascontiguousarray(np.array(self.dilation, dtype=np.int32))
        h = C.c_void_p()
        rc = nat.ems_cn_create(self.cin, self.cout, self.channels, len(self.dilation),
                               nat.ptr(dil, C.c_int32), C.c_uint64(seed), C.byref(h))
        if rc != 0:
            raise EmsError(rc, "ems_cn_create failed")
        self._h = h
        self.param_count = int(nat.ems_cn_param_count(h))
        self._data_T = 0

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise EmsError(rc, nat.ems_cn_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            nat.ems_cn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.empty(self.param_count, np.float32)
        self._check(nat.ems_cn_get_params(self._h, nat.ptr(p, C.c_float), p.size))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(-1))
        if p.size != self.param_count:
            raise EmsError(nat.EMS_E_INVALID,
                           f"{p.size} parameters given, the network has {self.param_count}")
        self._check(nat.ems_cn_set_params(self._h, nat.ptr(p, C.c_float), p.size))

    def _seq(self, x, ch: int, what: str) -> np.ndarray:
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
        if x.ndim != 3 or x.shape[1] != ch:
            raise EmsError(nat.EMS_E_INVALID, f"{what}: expected [n][{ch}][T], got {x.shape}")
        return x

    def set_data(self, x, y) -> None:
        x = self._seq(x, self.cin, "inputs")
        y = self._seq(y, self.cout, "targets")
        if x.shape[0] != y.shape[0] or x.shape[2] != y.shape[2]:
            raise EmsError(nat.EMS_E_INVALID, f"inputs {x.shape} and targets {y.shape} differ")
        self._check(nat.ems_cn_set_data(self._h, nat.ptr(x, C.c_float), nat.ptr(y, C.c_float),
                                        x.shape[0], x.shape[2]))
        self._data_n, self._data_T = x.shape[0], x.shape[2]

    def train_batch(self, idx, t0: int, lr: float, clip: float = 0.0) -> float:
        """One Adam step; returns the batch loss (before the step). The
        gradient norm before clipping is kept in last_grad_norm."""
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int32))
        loss, gn = C.c_float(), C.c_float()
        self._check(nat.ems_cn_train_batch(self._h, nat.ptr(idx, C.c_int32), idx.size, int(t0),
                                           C.c_float(lr), C.c_float(clip), C.byref(loss),
                                           C.byref(gn)))
        self.last_grad_norm = gn.value
        return loss.value

    def set_spectral(self, K, t0: int, weight: float) -> None:
        """Adds weight * d^T K d per sample and output channel (averaged) to
        the loss, d the error over samples [t0, t0 + len(K)); weight 0
        removes it."""
        if weight == 0.0:
            self._check(nat.ems_cn_set_spectral(self._h, None, 0, 0, C.c_float(0.0)))
            return
        K = np.ascontiguousarray(np.asarray(K, dtype=np.float32))
        if K.ndim != 2 or K.shape[0] != K.shape[1]:
            raise EmsError(nat.EMS_E_INVALID, f"spectral matrix must be square, got {K.shape}")
        self._check(nat.ems_cn_set_spectral(self._h, nat.ptr(K, C.c_float), int(t0), K.shape[0],
                                            C.c_float(weight)))

    def forward(self, x) -> np.ndarray:
        x = self._seq(x, self.cin, "inputs")
        out = np.empty((x.shape[0], self.cout, x.shape[2]), np.float32)
        self._check(nat.ems_cn_forward(self._h, nat.ptr(x, C.c_float), x.shape[0], x.shape[2],
                                       nat.ptr(out, C.c_float)))
        return out

    def loss_grad(self, x, y, t0: int) -> Tuple[float, np.ndarray]:
        x = self._seq(x, self.cin, "inputs")
        y = self._seq(y, self.cout, "targets")
        g = np.empty(self.param_count, np.float32)
        loss = C.c_float()
        self._check(nat.ems_cn_loss_grad(self._h, nat.ptr(x, C.c_float), nat.ptr(y, C.c_float),
                                         x.shape[0], x.shape[2], int(t0), C.byref(loss),
                                         nat.ptr(g, C.c_float)))
        return loss.value, g


@dataclass
class TraceFormat:
    """How port traces become network sequences: `horizon` solver steps
    decimated by `decimation` (sample m is step m*decimation), of which the
    first `observed` steps are known. The reference resistance scales the
    current to volts. Inputs per sample: observed V, observed R*I (both
    times `gain`, zero after the observed part), the observed-part mask and
    the excitation; outputs: V and R*I."""
    horizon: int
    decimation: int
    observed: int
    resistance: float = 50.0
    gain: float = 1.0

    @property
    def T(self) -> int:
        return self.horizon // self.decimation

    @property
    def N(self) -> int:
        return self.observed // self.decimation

    def check(self) -> None:
        if self.decimation < 1 or self.horizon % self.decimation or self.observed % self.decimation:
            raise EmsError(nat.EMS_E_INVALID, "horizon and observed steps must be multiples of "
                                              "the decimation")
        if not 0 < self.observed < self.horizon:
            raise EmsError(nat.EMS_E_INVALID, "observed steps must lie inside the horizon")

    def source(self, waveform, dt: float, amplitude: float = 1.0) -> np.ndarray:
        """The excitation at the decimated steps (mid-step times)."""
        D = self.decimation
        return np.array([amplitude * waveform.value((m * D + 0.5) * dt) for m in range(self.T)],
                        np.float32)

    def inputs(self, v, i, src, decimated: bool = False) -> np.ndarray:
        """[n][4][T] network inputs from full-rate traces (at least
        `observed` samples each; already decimated with `decimated`) and
        the decimated excitation."""
        v = np.atleast_2d(np.asarray(v, np.float64))
        i = np.atleast_2d(np.asarray(i, np.float64))
        T, N = self.T, self.N
        D = 1 if decimated else self.decimation
        x = np.zeros((v.shape[0], 4, T), np.float32)
        x[:, 0, :N] = v[:, :N * D:D] * self.gain
        x[:, 1, :N] = i[:, :N * D:D] * self.resistance * self.gain
        x[:, 2, :N] = 1.0
        x[:, 3, :] = src[None, :]
        return x

    def targets(self, v, i, decimated: bool = False) -> np.ndarray:
        """[n][2][T] decimated V and R*I over the horizon (traces padded
        with zeros beyond their end)."""
        v = np.atleast_2d(np.asarray(v, np.float64))
        i = np.atleast_2d(np.asarray(i, np.float64))
        T = self.T
        D = 1 if decimated else self.decimation
        y = np.zeros((v.shape[0], 2, T), np.float32)
        nv = min(T, (v.shape[1] + D - 1) // D)
        ni = min(T, (i.shape[1] + D - 1) // D)
        y[:, 0, :nv] = v[:, :nv * D:D]
        y[:, 1, :ni] = i[:, :ni * D:D] * self.resistance
        return y


def spectral_matrix(fmt: TraceFormat, src_full, dt: float, freqs=FREQS_1_3_GHZ) -> np.ndarray:
    """K for ConvNet.set_spectral over the predicted samples [N, T): the
    error's power at `freqs`, each frequency weighted by 1 / |S(f)|^2 (S the
    excitation's spectrum, so an error counts as it would in S11, whose
    incident wave is S/2), weights normalised to mean 1 and K to the scale
    of the mean squared error (a white error gives about the same value):
    K[t][u] = sum_f w_f cos(2 pi f (t - u) D dt) / (F (T - N))."""
    D, n = fmt.decimation, fmt.T - fmt.N
    S = np.abs(np.array(post.dft(0.0, dt, np.asarray(src_full, np.float64), freqs)))
    w = 1.0 / np.maximum(S, 1e-30) ** 2
    w /= w.mean()
    tau = np.arange(n) * (D * dt)
    diff = tau[:, None] - tau[None, :]
    K = np.zeros((n, n))
    for f, wf in zip(freqs, w):
        K += wf * np.cos(2.0 * np.pi * f * diff)
    return (K / (len(freqs) * n)).astype(np.float32)


def residual_db(pred: np.ndarray, tail: int = 16) -> np.ndarray:
    """Port signal power (V^2 + (R I)^2) over the last `tail` samples of
    the predicted horizon relative to its peak, in dB, per sample."""
    p = pred[:, 0].astype(np.float64) ** 2 + pred[:, 1].astype(np.float64) ** 2
    peak = p.max(axis=1)
    end = p[:, -tail:].mean(axis=1)
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(np.where(peak > 0, end / np.where(peak > 0, peak, 1.0), 1.0))


def early_port_result(fmt: TraceFormat, freqs, dt: float, v_obs, i_obs, pred,
                      stop: Optional[int] = None) -> post.PortResult:
    """Port spectra from the observed full-rate samples (V at (n+1)dt, I at
    (n+1/2)dt, n < observed) plus the predicted decimated tail up to sample
    `stop` (default: the horizon), each decimated sample standing for
    `decimation` steps. The first tail sample is weighted (D+1)/2 steps
    instead of D: the Euler-Maclaurin end terms of the full-rate sum and
    the decimated one differ by (D-1)/2 steps of that sample, so the
    junction adds no first-order error."""
    n0, D = fmt.observed, fmt.decimation
    stop = fmt.T if stop is None else stop
    vh = pred[0, fmt.N:stop].astype(np.float64)
    ih = pred[1, fmt.N:stop].astype(np.float64) / fmt.resistance
    if vh.size:
        vh[0] *= (D + 1) / (2.0 * D)
        ih[0] *= (D + 1) / (2.0 * D)
    v = np.asarray(v_obs, np.float64)[:n0]
    i = np.asarray(i_obs, np.float64)[:n0]
    vf = np.array(post.dft(dt, dt, v, freqs)) + np.array(post.dft((n0 + 1) * dt, D * dt, vh, freqs))
    i_f = np.array(post.dft(0.5 * dt, dt, i, freqs)) + \
        np.array(post.dft((n0 + 0.5) * dt, D * dt, ih, freqs))
    return post.PortResult(freqs, list(vf), list(i_f), complex(fmt.resistance, 0.0))


def _dft_matrix(t0: float, step: float, n: int, freqs) -> np.ndarray:
    """E[f, k] = exp(-j 2 pi f (t0 + k step)) step: the DFT of post.dft as a
    matrix, so that many traces of one time step transform at once."""
    t = t0 + np.arange(n) * step
    return np.exp(-2j * np.pi * np.outer(np.asarray(freqs, np.float64), t)) * step


def early_s11(fmt: TraceFormat, freqs, dts, v_obs, i_obs, pred) -> List[List[complex]]:
    """S11 of early_port_result for many models at once (models of one time
    step share their DFT matrices); the same sums as the per-model form up to
    the rounding of the transform."""
    n0, D, R = fmt.observed, fmt.decimation, fmt.resistance
    vh = pred[:, 0, fmt.N:fmt.T].astype(np.float64)
    ih = pred[:, 1, fmt.N:fmt.T].astype(np.float64) / R
    vh[:, 0] *= (D + 1) / (2.0 * D)
    ih[:, 0] *= (D + 1) / (2.0 * D)
    v = np.asarray(v_obs, np.float64)[:, :n0]
    i = np.asarray(i_obs, np.float64)[:, :n0]
    out: List[List[complex]] = [None] * len(dts)
    for dt in sorted(set(float(d) for d in dts)):
        qs = [q for q, d in enumerate(dts) if float(d) == dt]
        ev = _dft_matrix(dt, dt, n0, freqs)
        eh = _dft_matrix((n0 + 1) * dt, D * dt, vh.shape[1], freqs)
        ei = _dft_matrix(0.5 * dt, dt, n0, freqs)
        eih = _dft_matrix((n0 + 0.5) * dt, D * dt, ih.shape[1], freqs)
        vf = v[qs] @ ev.T + vh[qs] @ eh.T
        i_f = i[qs] @ ei.T + ih[qs] @ eih.T
        for k, q in enumerate(qs):
            out[q] = post.PortResult(freqs, list(vf[k]), list(i_f[k]), complex(R, 0.0)).s11()
    return out


def truncated_port_result(fmt: TraceFormat, freqs, dt: float, v_obs, i_obs) -> post.PortResult:
    """The method without the network: the spectra of the observed samples
    alone."""
    n0 = fmt.observed
    v = np.asarray(v_obs, np.float64)[:n0]
    i = np.asarray(i_obs, np.float64)[:n0]
    return post.PortResult(freqs, post.dft(dt, dt, v, freqs), post.dft(0.5 * dt, dt, i, freqs),
                           complex(fmt.resistance, 0.0))


# -- solver traces, training and the early-stopped sweep ---------------------

# The network trained on the patch-antenna family (python -m ems_amd.learn
# --spectral 4 --out ...): 4096 solver traces of seeded patches, horizon
# 4096 steps, decimation 8, 1024 observed steps, 40 000 Adam steps.
PATCH_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "earlystop_patch.npz")


def _schedule(b, models, tile_rows: int, x_chunks: int = 16) -> None:
    """Packed short rows when every model allows them (nz <= 64), else the
    tiled kernel of height tile_rows; x_chunks per model (16 measured best
    for the patch family, 0 automatic)."""
    if all(m.grid.nodes()[2] <= 64 for m in models):
        b.set_schedule(tile_rows, x_chunks, packed=True)
    else:
        b.set_schedule(tile_rows, x_chunks)


def _batch_engines(models, workers: int = 8):
    """One engine per model with its operators. The models are independent
    and the native calls (coefficient pooling, CPML layers, uploads) release
    the interpreter lock, so they are built by a few threads at once; the
    order of the result is the models'."""
    from concurrent.futures import ThreadPoolExecutor
    from .engine import Engine

    def one(m):
        e = Engine(m.grid, m.dt, m.lanes)
        p = m.build_ops(e)
        e.release_dense_lanes()
        return e, p
    if workers <= 1 or len(models) < 2:
        built = [one(m) for m in models]
    else:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            built = list(ex.map(one, models))
    return [b[0] for b in built], [b[1] for b in built]


def solver_traces(models, steps: int, group: int = 256, tile_rows: int = 7, log=None):
    """Runs the models `steps` steps in batches of `group` (one fused launch
    per step over each batch) and returns the port voltage and current
    traces [n][steps] (f32) and each model's time step."""
    from .engine import EngineBatch
    n = len(models)
    v = np.zeros((n, steps), np.float32)
    i = np.zeros((n, steps), np.float32)
    for g0 in range(0, n, group):
        part = models[g0:g0 + group]
        engines, ports = _batch_engines(part)
        b = EngineBatch(engines)
        _schedule(b, part, tile_rows)
        b.run(steps)
        for q, p in enumerate(ports):
            v[g0 + q] = p.voltage()[:steps]
            i[g0 + q] = p.current()[:steps]
        b.close()
        for e in engines:
            e.close()
        if log:
            log(f"traces {g0 + len(part)}/{n}")
    return v, i, np.array([m.dt for m in models])


@dataclass
class EarlyStop:
    """A trained network with its trace format and excitation."""
    net: ConvNet
    fmt: TraceFormat
    src: np.ndarray
    threshold_db: float = -40.0

    def predict(self, v_obs, i_obs) -> np.ndarray:
        """[n][2][T] predicted V and R*I from [n][>=observed] traces."""
        return self.net.forward(self.fmt.inputs(v_obs, i_obs, self.src))

    def save(self, path: str) -> None:
        f = self.fmt
        np.savez(path, params=self.net.params(), src=self.src,
                 meta=np.array([f.horizon, f.decimation, f.observed, self.net.cin, self.net.cout,
                                self.net.channels], np.int64),
                 dilation=np.array(self.net.dilation, np.int64),
                 floats=np.array([f.resistance, f.gain, self.threshold_db], np.float64))

    @staticmethod
    def load(path: str = PATCH_MODEL) -> "EarlyStop":
        z = np.load(path, allow_pickle=False)
        H, D, obs, cin, cout, ch = (int(q) for q in z["meta"])
        r, gain, thr = (float(q) for q in z["floats"])
        net = ConvNet(cin, cout, ch, [int(d) for d in z["dilation"]])
        net.set_params(z["params"])
        return EarlyStop(net, TraceFormat(H, D, obs, r, gain), z["src"].astype(np.float32), thr)


def train(es: EarlyStop, v, i, steps: int, batch: int = 32, lr: float = 6e-4,
          lr_end: float = 1e-5, warmup: int = 1000, seed: int = 0, log=None, every: int = 500,
          decimated: bool = False, spectral=None):
    """Trains es.net on full traces [n][horizon]: linear warm-up, then a
    cosine-decayed learning rate; the loss counts the predicted part
    (t >= N) only. Gradients longer than three times the running mean norm
    are clipped to it (a guard against the rare large step that throws
    Adam off). Returns the loss history (one value per step)."""
    fmt = es.fmt
    x = fmt.inputs(v, i, es.src, decimated)
    y = fmt.targets(v, i, decimated)
    es.net.set_data(x, y)
    if spectral is not None:
        K, weight = spectral
        es.net.set_spectral(K, fmt.N, weight)
    rng = np.random.default_rng(seed)
    hist = []
    n = x.shape[0]
    mean_norm = 0.0
    for s in range(steps):
        if s < warmup:
            rate = lr * (s + 1) / warmup
        else:
            a = 0.5 * (1.0 + math.cos(math.pi * (s - warmup) / max(steps - warmup - 1, 1)))
            rate = lr_end + (lr - lr_end) * a
        clip = 3.0 * mean_norm if s >= 100 else 0.0
        hist.append(es.net.train_batch(rng.integers(0, n, batch), fmt.N, rate, clip))
        g = es.net.last_grad_norm
        g = min(g, clip) if clip > 0 else g
        mean_norm = g if s == 0 else 0.99 * mean_norm + 0.01 * g
        if log and (s % every == 0 or s == steps - 1):
            log(f"step {s} loss {np.mean(hist[-every:]):.3e} lr {rate:.2e} |g| {mean_norm:.2e}")
    return np.array(hist)


@dataclass
class SweepResult:
    s11: List[List[complex]]
    fallback: List[bool]
    steps: List[int]
    residual_db: np.ndarray


def early_stop_sweep(es: EarlyStop, models, limits, freqs=FREQS_1_3_GHZ, tile_rows: int = 7,
                     truncated: bool = False) -> SweepResult:
    """Every model runs `observed` steps in one batch; the network predicts
    the rest of each trace; models whose predicted port power over the last
    16 samples of the horizon stays above the threshold (relative to its
    peak) are run again in full to `limits` (the fallback). With
    `truncated`, the spectra come from the observed samples alone (the
    method without the network) and nothing falls back."""
    from .engine import EngineBatch
    fmt = es.fmt
    engines, ports = _batch_engines(models)
    b = EngineBatch(engines)
    _schedule(b, models, tile_rows)
    b.run(fmt.observed)
    v = np.stack([p.voltage()[:fmt.observed] for p in ports])
    i = np.stack([p.current()[:fmt.observed] for p in ports])
    b.close()
    for e in engines:
        e.close()
    if truncated:
        s11 = [truncated_port_result(fmt, freqs, models[q].dt, v[q], i[q]).s11()
               for q in range(len(models))]
        return SweepResult(s11, [False] * len(models), [fmt.observed] * len(models),
                           np.zeros(len(models)))
    pred = es.predict(v, i)
    rdb = residual_db(pred)
    fb = [bool(r > es.threshold_db) for r in rdb]
    s11 = [None] * len(models)
    steps = [fmt.observed] * len(models)
    keep = [q for q in range(len(models)) if not fb[q]]
    if keep:
        got = early_s11(fmt, freqs, [models[q].dt for q in keep], v[keep], i[keep], pred[keep])
        for q, x in zip(keep, got):
            s11[q] = x
    redo = [q for q in range(len(models)) if fb[q]]
    if redo:
        engines, ports = _batch_engines([models[q] for q in redo])
        b = EngineBatch(engines)
        _schedule(b, [models[q] for q in redo], tile_rows)
        outs = b.run_until(limits)
        for k, q in enumerate(redo):
            s11[q] = ports[k].analyze_device(freqs).s11()
            steps[q] = fmt.observed + outs[k].steps
        b.close()
        for e in engines:
            e.close()
    return SweepResult(s11, fb, steps, rdb)


def _s11_err(a, b) -> float:
    return max(abs(x - y) for x, y in zip(a, b))


def held_out_report(es: EarlyStop, v, i, dts, freqs=FREQS_1_3_GHZ) -> dict:
    """|dS11| over 1-3 GHz of the early-stopped spectra (observed part plus
    prediction) and of the truncated spectra (observed part alone) against
    the spectra of the whole traces, per held-out trace (time steps dts)."""
    fmt = es.fmt
    pred = es.predict(v, i)
    net_err, trunc_err = [], []
    for q in range(v.shape[0]):
        dt = float(dts[q])
        full = post.PortResult(freqs, post.dft(dt, dt, v[q].astype(np.float64), freqs),
                               post.dft(0.5 * dt, dt, i[q].astype(np.float64), freqs),
                               complex(fmt.resistance, 0.0)).s11()
        net_err.append(_s11_err(early_port_result(fmt, freqs, dt, v[q], i[q], pred[q]).s11(), full))
        trunc_err.append(_s11_err(truncated_port_result(fmt, freqs, dt, v[q], i[q]).s11(), full))
    y = fmt.targets(v, i)
    rel = np.sqrt(np.mean((pred[:, :, fmt.N:] - y[:, :, fmt.N:]) ** 2) /
                  np.mean(y[:, :, fmt.N:] ** 2))
    return {"n": int(v.shape[0]), "trace_rel_rms": float(rel),
            "net_dS11_max": float(np.max(net_err)), "net_dS11_median": float(np.median(net_err)),
            "net_dS11_p90": float(np.percentile(net_err, 90)),
            "trunc_dS11_max": float(np.max(trunc_err)),
            "trunc_dS11_median": float(np.median(trunc_err)),
            "residual_db_max": float(residual_db(pred).max()),
            "net_err": [float(e) for e in net_err], "trunc_err": [float(e) for e in trunc_err]}


def main(argv=None) -> int:
    """Generates solver traces of seeded patch antennas on the device,
    trains the network on them there, reports held-out accuracy and saves
    the trained network."""
    import argparse
    import json
    import sys
    import time
    from . import workloads as W
    from .engine import Waveform
    ap = argparse.ArgumentParser(prog="python -m ems_amd.learn")
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--held-out", type=int, default=512)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--horizon", type=int, default=4096)
    ap.add_argument("--decimation", type=int, default=8)
    ap.add_argument("--observed", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=6e-4)
    ap.add_argument("--channels", type=int, default=64)
    ap.add_argument("--spectral", type=float, default=0.0,
                    help="weight of the 1-3 GHz spectral error term in the loss (0: off)")
    ap.add_argument("--out", default="")
    ap.add_argument("--traces", default="", help="save the traces (npz) here")
    ap.add_argument("--traces-in", default="",
                    help="train on traces saved with --traces instead of running the solver")
    ap.add_argument("--report", default="", help="write the held-out report (json) here")
    a = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    tr, ho = slice(0, a.train), slice(a.train, a.train + a.held_out)
    if a.traces_in:
        z = np.load(a.traces_in, allow_pickle=False)
        vt, it = z["v_train"][:a.train], z["i_train"][:a.train]
        vh, ih = z["v_held"][:a.held_out], z["i_held"][:a.held_out]
        dts = np.asarray(z["dt"], np.float64)
        dts = dts if dts.ndim else np.full(a.train + a.held_out, float(dts))
        dth = dts[-len(z["v_held"]):][:a.held_out]
        t_gen = 0.0
    else:
        models = W.patch_sweep(a.train + a.held_out, seed=a.seed)
        log(f"{len(models)} models, {sum(m.cells for m in models) / len(models):.0f} cells on average")
        v, i, dts = solver_traces(models, a.horizon, log=log)
        t_gen = time.perf_counter() - t0
        cell_steps = sum(m.cells for m in models) * a.horizon
        log(f"traces in {t_gen:.1f} s ({cell_steps / t_gen / 1e9:.2f} Gcell/s)")
        vt, it = v[tr, ::a.decimation], i[tr, ::a.decimation]
        vh, ih, dth = v[ho], i[ho], dts[ho]
    fmt = TraceFormat(a.horizon, a.decimation, a.observed)
    fmt.check()
    # (the models' time steps agree to rounding: the 0.5 mm mesh and the
    # Courant factor fix them)
    src = fmt.source(Waveform.gaussian(2e9, 1e9), float(dts[0]))
    peak_obs = float(np.abs(src[:fmt.N]).max())
    fmt.gain = 1.0 / peak_obs if peak_obs > 0 else 1.0
    es = EarlyStop(ConvNet(4, 2, a.channels, (1, 2, 4, 8, 16, 32), seed=a.seed), fmt, src)
    if a.traces and not a.traces_in:
        np.savez(a.traces, v_train=vt, i_train=it, v_held=vh, i_held=ih, dt=dts)
    t1 = time.perf_counter()
    spec = None
    if a.spectral > 0.0:
        w = Waveform.gaussian(2e9, 1e9)
        full = [w.value((m + 0.5) * float(dts[0])) for m in range(a.horizon)]
        spec = (spectral_matrix(fmt, full, float(dts[0])), a.spectral)
    hist = train(es, vt, it, a.steps, a.batch, a.lr, log=log, decimated=True, spectral=spec)
    t_train = time.perf_counter() - t1
    log(f"trained {a.steps} steps in {t_train:.1f} s")
    rep = held_out_report(es, vh, ih, dth)
    rep.update({"generate_s": t_gen, "train_s": t_train, "steps": a.steps, "batch": a.batch,
                "final_loss": float(np.mean(hist[-200:])), "horizon": a.horizon,
                "observed": a.observed, "decimation": a.decimation, "train": a.train,
                "seed": a.seed})
    log(json.dumps({k: v for k, v in rep.items() if not isinstance(v, list)}))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f)
    if a.out:
        es.save(a.out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
