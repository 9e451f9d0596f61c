"""Frequency-domain post-processing of recorded series (src/post/).

``dft`` runs in the C++ host library (RunningDft, src/post/dft.rs:75-222).
Port quantities follow src/post/port.rs:50-240 with the reference's own
complex arithmetic (src/post/complex.rs): products and sums as written
there, division by Smith's algorithm, so results match it bit for bit.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError

Cx = Tuple[float, float]


def dft(t0: float, dt: float, samples, freqs) -> List[complex]:
    """X(f) = sum x(t0 + n dt) e^{-j2 pi f t_n} dt (dft.rs:213-222)."""
    x = np.ascontiguousarray(np.asarray(samples, dtype=np.float64))
    f = np.ascontiguousarray(np.asarray(freqs, dtype=np.float64))
    out = np.empty(2 * f.size, dtype=np.float64)
    nat.ems_dft(float(t0), float(dt), nat.ptr(x, C.c_double), x.size, nat.ptr(f, C.c_double),
                f.size, nat.ptr(out, C.c_double))
    return [complex(out[2 * k], out[2 * k + 1]) for k in range(f.size)]


# complex.rs arithmetic on (re, im) pairs
def _add(a: Cx, b: Cx) -> Cx:
    return (a[0] + b[0], a[1] + b[1])


def _sub(a: Cx, b: Cx) -> Cx:
    return (a[0] - b[0], a[1] - b[1])


def _mul(a: Cx, b: Cx) -> Cx:
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def _scale(a: Cx, k: float) -> Cx:
    return (a[0] * k, a[1] * k)


def _div(x: Cx, y: Cx) -> Cx:
    """Smith's algorithm (complex.rs:163-180)."""
    a, b = x
    c, d = y
    if abs(c) >= abs(d):
        r = d / c
        den = c + d * r
        return ((a + b * r) / den, (b - a * r) / den)
    r = c / d
    den = c * r + d
    return ((a * r + b) / den, (b * r - a) / den)


def _conj(a: Cx) -> Cx:
    return (a[0], -a[1])


def _c(z) -> Cx:
    z = complex(z)
    return (z.real, z.imag)


def _z(a: Cx) -> complex:
    return complex(a[0], a[1])


class PortResult:
    """PortResult (post/port.rs:50-210): V(f), I(f) and the reference
    impedance per frequency."""

    def __init__(self, freqs: Sequence[float], v: Sequence[complex], i: Sequence[complex],
                 z_ref):
        freqs = [float(f) for f in freqs]
        zs = list(z_ref) if isinstance(z_ref, (list, tuple, np.ndarray)) else \
            [z_ref] * max(len(freqs), 1)
        if len(v) != len(freqs) or len(i) != len(freqs) or (
                isinstance(z_ref, (list, tuple, np.ndarray)) and len(zs) != len(freqs)):
            raise EmsError(nat.EMS_E_INVALID,
                           f"port result: {len(freqs)} frequencies but {len(v)} voltages, "
                           f"{len(i)} currents and {len(zs)} reference impedances")
        for z in zs:
            z = complex(z)
            if not (math.isfinite(z.real) and z.real > 0.0):
                raise EmsError(nat.EMS_E_INVALID, f"reference impedance {z} needs Re > 0")
        self.freqs = freqs
        self.v = [_c(x) for x in v]
        self.i = [_c(x) for x in i]
        self.z_ref = [_c(z) for z in zs]

    def with_z_ref(self, z_ref) -> "PortResult":
        """The same spectra with another reference impedance (constant, or
        one per frequency)."""
        return PortResult(self.freqs, [_z(v) for v in self.v], [_z(i) for i in self.i], z_ref)

    def voltage(self) -> List[complex]:
        return [_z(v) for v in self.v]

    def current(self) -> List[complex]:
        return [_z(i) for i in self.i]

    def _viz(self):
        return zip(self.v, self.i, self.z_ref)

    def a(self) -> List[complex]:
        """(V + Z I) / (2 sqrt(Re Z))."""
        return [_z(_scale(_add(v, _mul(z, i)), 0.5 / math.sqrt(z[0]))) for v, i, z in self._viz()]

    def b(self) -> List[complex]:
        """(V - Z* I) / (2 sqrt(Re Z))."""
        return [_z(_scale(_sub(v, _mul(_conj(z), i)), 0.5 / math.sqrt(z[0])))
                for v, i, z in self._viz()]

    def v_inc(self) -> List[complex]:
        return [_z(_scale(_add(v, _mul(z, i)), 0.5)) for v, i, z in self._viz()]

    def v_ref(self) -> List[complex]:
        return [_z(_scale(_sub(v, _mul(z, i)), 0.5)) for v, i, z in self._viz()]

    def i_inc(self) -> List[complex]:
        return [_z(_scale(_add(_div(v, z), i), 0.5)) for v, i, z in self._viz()]

    def i_ref(self) -> List[complex]:
        return [_z(_scale(_sub(i, _div(v, z)), 0.5)) for v, i, z in self._viz()]

    def z_in(self) -> List[complex]:
        """V / I."""
        return [_z(_div(v, i)) for v, i in zip(self.v, self.i)]

    def s11(self) -> List[complex]:
        return s_param(self, self)

    def incident_power(self) -> List[float]:
        return [0.5 * (a.real * a.real + a.imag * a.imag) for a in self.a()]

    def reflected_power(self) -> List[float]:
        return [0.5 * (b.real * b.real + b.imag * b.imag) for b in self.b()]

    def accepted_power(self) -> List[float]:
        return [0.5 * (v[0] * i[0] + v[1] * i[1]) for v, i in zip(self.v, self.i)]


def s_param(out: PortResult, inc: PortResult) -> List[complex]:
    """S = b_out / a_inc (post/port.rs:232-247)."""
    if out.freqs != inc.freqs:
        raise ValueError("port results at different frequencies")
    return [_z(_div(_c(b), _c(a))) for b, a in zip(out.b(), inc.a())]
