"""Total-field / scattered-field plane wave (src/engine/planewave.rs).

The host side of `PlaneWave::new` (planewave.rs:684-1049) runs here for a
homogeneous (optionally lossy) background: dispersion-matched spacing of
the auxiliary 1-D grid, reference corner, the corrections of the E edges
on the box surface and of the H edges just outside it with their 4-point
Lagrange stencils, and the auxiliary grid's coefficients (matched-loss
layer, taper for lossy backgrounds). The device applies the corrections
after the H and E updates and advances the auxiliary grid
(planewave.rs:1513-1561). Layered and dispersive backgrounds (the
spectral incident field) are not held.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError
from .engine import C0, EPS0, MU0, Z0, Axis

AUX_PML = 64          # planewave.rs:433
AUX_PML_R = 1e-12     # planewave.rs:435
AUX_TAPER = 256       # planewave.rs:438


class Medium:
    """Medium (src/engine/medium.rs:1-110)."""

    def __init__(self, eps_r=1.0, mu_r=1.0, kappa=0.0, sigma_m=0.0):
        self.eps_r, self.mu_r = float(eps_r), float(mu_r)
        self.kappa, self.sigma_m = float(kappa), float(sigma_m)

    VACUUM = None  # set below

    def velocity(self) -> float:
        return C0 / math.sqrt(self.eps_r * self.mu_r)

    def impedance(self) -> float:
        return Z0 * math.sqrt(self.mu_r / self.eps_r)

    def is_lossless(self) -> bool:
        return self.kappa == 0.0 and self.sigma_m == 0.0

    def validate(self) -> None:
        pos = lambda x: math.isfinite(x) and x > 0.0  # noqa: E731
        nonneg = lambda x: math.isfinite(x) and x >= 0.0  # noqa: E731
        if not (pos(self.eps_r) and pos(self.mu_r) and nonneg(self.kappa) and nonneg(self.sigma_m)):
            raise EmsError(nat.EMS_E_INVALID, f"invalid medium {vars(self)}")


Medium.VACUUM = Medium()


class PlaneWaveSpec:
    """PlaneWaveSpec (planewave.rs:255-420)."""

    def __init__(self, start, stop, direction, polarization):
        self.start = tuple(int(v) for v in start)
        self.stop = tuple(int(v) for v in stop)
        self.direction = tuple(float(v) for v in direction)
        self.polarization = tuple(float(v) for v in polarization)
        self.eps_r, self.mu_r, self.kappa, self.sigma_m = 1.0, 1.0, 0.0, 0.0
        self.match_frequency = 0.0

    @staticmethod
    def new(start, stop, direction, polarization) -> "PlaneWaveSpec":
        return PlaneWaveSpec(start, stop, direction, polarization)

    def with_background(self, eps_r: float, mu_r: float) -> "PlaneWaveSpec":
        self.eps_r, self.mu_r = float(eps_r), float(mu_r)
        return self

    def with_losses(self, kappa: float, sigma_m: float) -> "PlaneWaveSpec":
        self.kappa, self.sigma_m = float(kappa), float(sigma_m)
        return self

    def with_medium(self, m: Medium) -> "PlaneWaveSpec":
        return self.with_background(m.eps_r, m.mu_r).with_losses(m.kappa, m.sigma_m)

    def with_match_frequency(self, f: float) -> "PlaneWaveSpec":
        self.match_frequency = float(f)
        return self

    def medium(self) -> Medium:
        return Medium(self.eps_r, self.mu_r, self.kappa, self.sigma_m)


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _bisect(lo, hi, f):
    """planewave.rs:526-543."""
    s_lo = f(lo) < 0.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            break
        if (f(mid) < 0.0) == s_lo:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def match_spacing(dt, v, f, d, cell):
    """The auxiliary spacing and matched wavenumber (planewave.rs:546-594)."""
    s = math.sin(math.pi * f * dt) / (v * dt)
    k_hi = math.inf
    for a in range(3):
        if abs(d[a]) > 0.0:
            k_hi = min(k_hi, math.pi / (abs(d[a]) * cell[a]))

    def rhs(k):
        acc = -0.0
        for a in range(3):
            x = math.sin(0.5 * k * d[a] * cell[a]) / cell[a]
            acc += x * x
        return acc - s * s
    if not (math.isfinite(s) and s > 0.0 and rhs(k_hi) > 0.0):
        raise EmsError(nat.EMS_E_INVALID,
                       f"plane wave: matching frequency {f} Hz is beyond the resolution of the mesh")
    k = _bisect(0.0, k_hi, rhs)
    axes = [a for a in range(3) if d[a] != 0.0]
    if len(axes) == 1:
        return cell[axes[0]], k
    g = lambda dd: math.sin(0.5 * k * dd) / dd - s  # noqa: E731
    lo = v * dt
    if g(lo) <= 0.0:
        return lo, k
    hi = math.pi / k
    if g(hi) > 0.0:
        raise EmsError(nat.EMS_E_INVALID,
                       f"plane wave: cannot match the auxiliary grid dispersion at {f} Hz")
    return _bisect(lo, hi, g), k


def _csqrt(z):
    """Complex64::sqrt via polar form (post/complex.rs:77-80)."""
    r, t = math.hypot(z.real, z.imag), math.atan2(z.imag, z.real)
    rs = math.sqrt(r)
    return complex(rs * math.cos(0.5 * t), rs * math.sin(0.5 * t))


def aux_gamma(wdt, se, sm, r):
    """planewave.rs:597-605."""
    sn, cs = math.sin(0.5 * wdt), math.cos(0.5 * wdt)
    a = lambda s: complex(2.0 * s * cs, 2.0 * sn)  # noqa: E731
    z = _csqrt(a(se) * a(sm)) / (2.0 * r)
    w = z + _csqrt(z * z + 1.0)
    g = complex(math.log(math.hypot(w.real, w.imag)), math.atan2(w.imag, w.real)) * 2.0
    return -g if g.real < 0.0 else g


def stencil(s):
    """4-point Lagrange stencil (planewave.rs:641-651)."""
    m = math.floor(s)
    t = s - m
    w = (-t * (t - 1.0) * (t - 2.0) / 6.0,
         (t + 1.0) * (t - 1.0) * (t - 2.0) / 2.0,
         -(t + 1.0) * t * (t - 2.0) / 2.0,
         (t + 1.0) * t * (t - 1.0) / 6.0)
    return int(m) - 1, w


def interp(v, m, w):
    return w[0] * v[m] + w[1] * v[m + 1] + w[2] * v[m + 2] + w[3] * v[m + 3]


class PlaneWave:
    """PlaneWave (planewave.rs:440-1470): build with :meth:`new`, then
    ``engine.add_operator``."""

    def __init__(self):
        self._installed = None

    @staticmethod
    def new(engine, spec: PlaneWaveSpec, amplitude: Callable[[float], float]) -> "PlaneWave":
        """PlaneWave::new (planewave.rs:684-1049), homogeneous background."""
        pw = PlaneWave()
        grid = engine.grid
        lanes = engine._dense_lanes
        if lanes is None:
            raise EmsError(nat.EMS_E_INVALID, "plane waves need the host lanes")
        n = grid.nodes()
        lo, hi = spec.start, spec.stop
        if any(lo[a] >= hi[a] or hi[a] >= n[a] for a in range(3)):
            raise EmsError(nat.EMS_E_INVALID,
                           f"plane wave: box {lo}..={hi} must have start < stop within the {n} nodes")
        ln = math.sqrt(_dot(spec.direction, spec.direction))
        if not (math.isfinite(ln) and ln > 0.0):
            raise EmsError(nat.EMS_E_INVALID, f"plane wave: invalid propagation direction {spec.direction}")
        k_hat = [x / ln for x in spec.direction]
        e0 = list(spec.polarization)
        e_len = math.sqrt(_dot(e0, e0))
        if not (math.isfinite(e_len) and e_len > 0.0) or abs(_dot(e0, k_hat)) > 1e-9 * e_len:
            raise EmsError(nat.EMS_E_INVALID,
                           f"plane wave: polarisation {e0} must be non-zero and perpendicular to "
                           f"the direction {k_hat}")
        medium = spec.medium()
        try:
            medium.validate()
        except EmsError:
            raise EmsError(nat.EMS_E_INVALID, "plane wave: invalid background") from None
        er, mr = spec.eps_r, spec.mu_r
        v = medium.velocity()
        eta = medium.impedance()
        h0 = [x / eta for x in _cross(k_hat, e0)]
        dt = engine.dt
        se = spec.kappa * dt / (2.0 * EPS0 * er)
        sm = spec.sigma_m * dt / (2.0 * MU0 * mr)

        def expect(geo, c, g):
            cap, cond = c * geo, g * geo
            den = cap + 0.5 * cond * dt
            return np.float32((cap - 0.5 * cond * dt) / den), np.float32(dt / den)

        def matches(got, want):
            (a, b), (ea, eb) = got, want
            return abs(np.float32(a - ea)) <= 1e-5 and abs(np.float32(b - eb)) <= 1e-5 * abs(eb)

        mismatch = None
        L = grid.lines
        cell = [(L[a][hi[a]] - L[a][lo[a]]) / float(hi[a] - lo[a]) for a in range(3)]
        mean = (cell[0] + cell[1] + cell[2]) / 3.0
        f = spec.match_frequency if spec.match_frequency > 0.0 else v / (20.0 * mean)
        if not math.isfinite(f):
            raise EmsError(nat.EMS_E_INVALID, f"plane wave: invalid matching frequency {f}")
        dxi, k = match_spacing(dt, v, f, k_hat, cell)
        r_ref = [float(L[a][lo[a]] if k_hat[a] >= 0.0 else L[a][hi[a]]) for a in range(3)]
        xi = lambda r: _dot(k_hat, [r[0] - r_ref[0], r[1] - r_ref[1], r[2] - r_ref[2]])  # noqa: E731
        dual_line = lambda a, i: 0.5 * (float(L[a][i]) + float(L[a][i + 1]))  # noqa: E731

        oth = ((1, 2), (2, 0), (0, 1))  # Axis.others() as ints

        def tf_h(d, q):
            e_, f_ = oth[d]
            return lo[d] <= q[d] <= hi[d] and lo[e_] <= q[e_] < hi[e_] and lo[f_] <= q[f_] < hi[f_]

        def h_valid(d, q):
            e_, f_ = oth[d]
            return all(0 <= q[a] < n[a] for a in range(3)) and q[e_] + 1 < n[e_] and q[f_] + 1 < n[f_]

        def e_pos(a, p):
            b, c = oth[a]
            r = [0.0] * 3
            r[a], r[b], r[c] = dual_line(a, p[a]), float(L[b][p[b]]), float(L[c][p[c]])
            return r

        def h_pos(d, q):
            e_, f_ = oth[d]
            r = [0.0] * 3
            r[d], r[e_], r[f_] = float(L[d][q[d]]), dual_line(e_, q[e_]), dual_line(f_, q[f_])
            return r

        flat = [lanes[l].reshape(-1) for l in range(12)]
        e_raw, h_raw = [], []
        for a in range(3):
            b, c = oth[a]
            for pa in range(lo[a], hi[a]):
                for pb in range(lo[b], hi[b] + 1):
                    # the surface: every pc on the two b faces, else the c faces
                    pcs = range(lo[c], hi[c] + 1) if pb in (lo[b], hi[b]) else (lo[c], hi[c])
                    for pc in pcs:
                        p = [0, 0, 0]
                        p[a], p[b], p[c] = pa, pb, pc
                        ip = grid.index(p)
                        cbv = float(flat[3 + a][ip])
                        e_checked = False

                        def minus(t):
                            q = list(p)
                            q[t] -= 1
                            return q
                        terms = [(c, list(p), 1.0), (c, minus(b), -1.0), (b, list(p), -1.0),
                                 (b, minus(c), 1.0)]
                        for d, q, sgn in terms:
                            if not h_valid(d, q) or tf_h(d, q):
                                continue
                            if mismatch is None and cbv != 0.0 and not e_checked:
                                e_checked = True
                                geo = grid.dual_delta(b, pb) * grid.dual_delta(c, pc) / grid.delta(a, pa)
                                got = (flat[a][ip], flat[3 + a][ip])
                                if not matches(got, expect(geo, EPS0 * er, spec.kappa)):
                                    mismatch = f"E edge along {Axis(a).name} at {p}"
                            iq = grid.index(q)
                            dbv = float(flat[9 + d][iq])
                            if mismatch is None and dbv != 0.0:
                                de, df = oth[d]
                                geo = grid.delta(de, q[de]) * grid.delta(df, q[df]) / \
                                    grid.dual_delta(d, q[d])
                                got = (flat[6 + d][iq], flat[9 + d][iq])
                                if not matches(got, expect(geo, MU0 * mr, spec.sigma_m)):
                                    mismatch = f"H edge along {Axis(d).name} at {q}"
                            if cbv != 0.0:
                                pos = h_pos(d, q)
                                e_raw.append([a, tuple(p), cbv * sgn * grid.dual_delta(d, q[d]), d,
                                              xi(pos)])
                            if dbv != 0.0:
                                pos = e_pos(a, p)
                                h_raw.append([d, tuple(q), dbv * sgn * grid.delta(a, pa), a,
                                              xi(pos)])
        e_raw = [r for r in e_raw if h0[r[3]] != 0.0]
        h_raw = [r for r in h_raw if e0[r[3]] != 0.0]
        for r in e_raw:
            r[2] *= h0[r[3]]
        for r in h_raw:
            r[2] *= e0[r[3]]
        if mismatch is not None:
            raise EmsError(nat.EMS_E_INVALID,
                           f"plane wave: the {mismatch} on the box surface does not have the "
                           "background medium of the source")
        xi_min = xi_max = 0.0
        for r in e_raw + h_raw:
            xi_min, xi_max = min(xi_min, r[4]), max(xi_max, r[4])
        xi_dom = xi_max
        for cc in range(8):
            rr = [float(L[a][-1] if (cc >> a) & 1 else L[a][0]) for a in range(3)]
            xi_dom = max(xi_dom, xi(rr))
        m0 = max(math.ceil(-xi_min / dxi), 0.0)
        m0 = int(m0) + 3
        main = int(math.ceil(xi_dom / dxi)) + m0 + 4
        taper = 0 if se == sm else AUX_TAPER
        s_top = max(se, sm)
        pml0 = main + taper
        nodes = pml0 + AUX_PML + 1
        r = v * dt / dxi
        smax = -math.log(AUX_PML_R) * r / float(AUX_PML)

        def loss(x, bg):
            t = min(max((x - float(main)) / float(max(taper, 1)), 0.0), 1.0)
            g = t * t * t * (10.0 - 15.0 * t + 6.0 * t * t)
            pml = x - float(pml0)
            if pml > 0.0:
                u = pml / float(AUX_PML)
                p = smax * (u * u * u)
            else:
                p = 0.0
            return bg + (s_top - bg) * (0.0 if taper == 0 else g) + p
        ca, cb = [1.0] * nodes, [0.0] * nodes
        for m in range(1, nodes - 1):
            s = loss(float(m), se)
            ca[m] = (1.0 - s) / (1.0 + s)
            cb[m] = r / (1.0 + s)
        da, db = [1.0] * (nodes - 1), [0.0] * (nodes - 1)
        for m in range(nodes - 1):
            s = loss(float(m) + 0.5, sm)
            da[m] = (1.0 - s) / (1.0 + s)
            db[m] = r / (1.0 + s)

        def finish(raw, half):
            out = []
            for ax, node, coef, _, x in raw:
                m, w = stencil(x / dxi + float(m0) - half)
                out.append((ax, node, coef, m, w))
            return out
        omega = 2.0 * math.pi * f
        if se == 0.0 and sm == 0.0:
            lead, gain = float(m0) * dxi * k / omega, 1.0
        else:
            g = aux_gamma(omega * dt, se, sm, r) / dxi
            x = float(m0) * dxi
            lead, gain = g.imag * x / omega, math.exp(g.real * x)
        far = 0.0
        for cc in range(8):
            rr = [float(L[a][hi[a] if (cc >> a) & 1 else lo[a]]) for a in range(3)]
            far = max(far, xi(rr))
        pw.spec, pw.amplitude = spec, amplitude
        pw.k_hat, pw.e0, pw.h0, pw.r_ref = k_hat, e0, h0, r_ref
        pw.dxi, pw.m0, pw.lead, pw.gain, pw.main = dxi, m0, lead, gain, main
        pw.transit = 1.1 * far / v
        pw.dt = dt
        pw.ca, pw.cb, pw.da, pw.db = ca, cb, da, db
        pw.e_corr, pw.h_corr = finish(e_raw, 0.5), finish(h_raw, 0.0)
        return pw

    # -- properties (planewave.rs:1356-1470)
    def direction(self):
        return list(self.k_hat)

    def spacing(self) -> float:
        return self.dxi

    def reference_point(self):
        return list(self.r_ref)

    def transit_time(self) -> float:
        return self.transit

    def corrections(self) -> Tuple[int, int]:
        return len(self.e_corr), len(self.h_corr)

    def source_value(self, n: int) -> float:
        """e[0] after the E update of step n: gain * a((n + 1)dt + lead)."""
        return self.gain * float(self.amplitude((n + 1.0) * self.dt + self.lead))

    def _aux(self):
        eng, oid = self._installed
        return eng._planewave_aux(oid, len(self.ca))

    def _aux_at(self, r, half):
        s = self.r_ref
        x = _dot(self.k_hat, [r[0] - s[0], r[1] - s[1], r[2] - s[2]]) / self.dxi + float(self.m0) - half
        if not (x >= 1.0 and x + 3.0 < len(self.ca) - 1):
            return None
        return stencil(x)

    def incident_e(self, r):
        """E_inc at r at the current time (planewave.rs:1446-1456)."""
        st = self._aux_at(r, 0.0)
        if st is None:
            return None
        e, _ = self._aux()
        v = interp(e, st[0], st[1])
        return [x * v for x in self.e0]

    def incident_h(self, r):
        st = self._aux_at(r, 0.5)
        if st is None:
            return None
        _, h = self._aux()
        v = interp(h, st[0], st[1])
        return [x * v for x in self.h0]

    def _bind(self, engine, oid: int) -> None:
        self._installed = (engine, oid)


class PwCorr(C.Structure):
    """ems_pw_corrections (include/ems_gpu.h)."""
    _fields_ = [("n", C.c_int32), ("axis", C.POINTER(C.c_int32)), ("nodes", C.POINTER(C.c_int32)),
                ("coef", C.POINTER(C.c_double)), ("m", C.POINTER(C.c_uint32)),
                ("w", C.POINTER(C.c_double))]


class PwDesc(C.Structure):
    """ems_pw_desc (include/ems_gpu.h)."""
    _fields_ = [("n_aux", C.c_int32), ("ca", C.POINTER(C.c_double)), ("cb", C.POINTER(C.c_double)),
                ("da", C.POINTER(C.c_double)), ("db", C.POINTER(C.c_double)),
                ("e_corr", PwCorr), ("h_corr", PwCorr)]


def make_desc(pw: PlaneWave):
    keep = []

    def arr(x, t):
        a = np.ascontiguousarray(np.asarray(x, dtype=t))
        if a.size == 0:
            a = np.zeros(1, dtype=t)
        keep.append(a)
        return a

    P = lambda a, ct: a.ctypes.data_as(C.POINTER(ct))  # noqa: E731

    def corr(lst):
        ax = arr([c[0] for c in lst], np.int32)
        nd = arr([v for c in lst for v in c[1]], np.int32)
        cf = arr([c[2] for c in lst], np.float64)
        m = arr([c[3] for c in lst], np.uint32)
        w = arr([x for c in lst for x in c[4]], np.float64)
        return PwCorr(len(lst), P(ax, C.c_int32), P(nd, C.c_int32), P(cf, C.c_double),
                      P(m, C.c_uint32), P(w, C.c_double))
    cols = [arr(x, np.float64) for x in (pw.ca, pw.cb, pw.da, pw.db)]
    d = PwDesc(len(pw.ca), *(P(c, C.c_double) for c in cols), corr(pw.e_corr), corr(pw.h_corr))
    return d, keep


class SourceTable:
    """e[0] of the auxiliary grid by step, kept ahead of the engine."""

    def __init__(self, pw: PlaneWave):
        self.pw = pw
        self.vals = np.empty(0)

    def ensure(self, upto: int, upload) -> None:
        if self.vals.size >= upto:
            return
        want = max(upto, 2 * self.vals.size, 1024)
        new = np.empty(want, dtype=np.float64)
        new[:self.vals.size] = self.vals
        for n in range(self.vals.size, want):
            new[n] = self.pw.source_value(n)
        self.vals = new
        upload(new)
