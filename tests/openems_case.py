"""The openEMS fixture ``gauss_soft_e`` (tests/fixtures/openems/gauss_soft_e of
the reference: openEMS v0.37 output of a 21^3-line vacuum PEC box, 1 mm
cells, a soft Gaussian Ez excitation over two edges, U/I/E/H probes, 400
steps; data copied under tests/golden/openems_gauss_soft_e), built as the
reference's file path builds it (src/sim/map.rs) for any engine (device or
oracle): openEMS's Rennings time step, CoeffRounding::OpenEms coefficients,
the openEMS curl order, the openEMS Gaussian with openEMS timing, the probes'
edges, weights and sums."""
from __future__ import annotations

import os

import numpy as np

import ems_amd as E
from ems_amd import engine as EE

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "openems_gauss_soft_e")

# src/constants.rs (OPENEMS_EPS0, OPENEMS_MU0)
OE_EPS0 = 8.85418781762e-12
OE_MU0 = 1.256637062e-6

STEPS = 400
INTERVAL = 21  # the probes' sampling interval (openEMS: Nyquist 86 / oversampling 4)


def rennings_dt_uniform(delta):
    """openEMS's Rennings time step (src/engine/timestep.rs:269-309) on a
    uniform vacuum grid with spacing delta[a] per axis, where every edge
    has the same circuit values: C_a = f32(EPS0 d_b d_c / d_a), L_a =
    f32(MU0 d_b d_c / d_a), couplings X(C, L) = f32(1/f32(C L)). Per edge
    axis a with transverse axes b0 = a+1, b1 = a+2 (c = 3-a-b):
    T_b = 1/4 (f64(X + X) + f64(X + X)) with X = X(C_a, L_c), and the
    own-axis term (its half-differences vanish) max over the two slots of
    X(C_b, L_c); omega^2 = T_b0 + T_b1 + own in f64, dt = 1/sqrt(max)."""
    f32 = np.float32
    d = [float(x) for x in delta]
    geo = [d[(a + 1) % 3] * d[(a + 2) % 3] / d[a] for a in range(3)]
    C = [f32(OE_EPS0 * 1.0 * geo[a]) for a in range(3)]
    L = [f32(OE_MU0 * 1.0 * geo[a]) for a in range(3)]

    def X(c, l):
        return f32(f32(1.0) / f32(c * l))

    w2max = 0.0
    for a in range(3):
        bs = [(a + 1) % 3, (a + 2) % 3]
        t, own = [], []
        for b in bs:
            c = 3 - a - b
            x = X(C[a], L[c])
            below = f32(x + x)
            above = f32(x + x)
            t.append(0.25 * (float(below) + float(above)))
        c0, c1 = 3 - a - bs[0], 3 - a - bs[1]
        x0, x1 = float(X(C[bs[0]], L[c0])), float(X(C[bs[1]], L[c1]))
        lc0, dc0 = 0.5 * (x0 + x0), 0.5 * (x0 - x0)
        lc1, dc1 = 0.5 * (x1 + x1), 0.5 * (x1 - x1)
        end = max(lc1 + abs(dc0), lc0 + abs(dc1))
        own = end + end
        w2max = max(w2max, t[0] + t[1] + 0.5 * own)
    return 1.0 / np.sqrt(w2max)


def grid():
    lines = [np.array([q * 1e-3 for q in range(21)]) for _ in range(3)]
    return E.Grid(lines)


def dt():
    return float(rennings_dt_uniform([1e-3, 1e-3, 1e-3]))


def lanes(g, step):
    return E.coefficients(g, None, [E.FaceCondition.Pec] * 6, step, rounding=1)


def read_rows(name):
    rows = []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            if line.startswith("%") or not line.strip():
                continue
            rows.append([float(v) for v in line.split()])
    return np.array(rows)


def current_loop_entries():
    """it1: NormDir z, box x, y 8..12 at z = 10: the loop on the dual plane
    k = 9 (snap_cell of z = 10, ties down; src/sim/map.rs:1858-1870) around
    x, y 8..12 (snap_upper / snap_lower of the corners, :1816-1840);
    loop_entries (src/engine/port.rs:231-260) negated for +z
    (current_loop, map.rs:1739-1771), in openEMS's order
    (openems_loop_order, map.rs:1776-1805): the x-side below along y, the
    y-side above along x, the x-side above, the y-side below, each by
    ascending index."""
    lo, hi, layer = [8, 8], [12, 12], 9
    X, Y = 0, 1
    ents = []
    for pc in range(lo[1], hi[1] + 1):  # dual edges along y on x = hi, x = lo - 1
        ents.append((Y, (hi[0], pc, layer), -1.0))
        ents.append((Y, (lo[0] - 1, pc, layer), 1.0))
    for pb in range(lo[0], hi[0] + 1):  # dual edges along x on y = hi, y = lo - 1
        ents.append((X, (pb, hi[1], layer), 1.0))
        ents.append((X, (pb, lo[1] - 1, layer), -1.0))
    ents = [(a, p, -w) for a, p, w in ents]

    def key(e):
        a, p, _ = e
        if a == X and p[1] < lo[1]:
            return (0, p[0])
        if a == Y and p[0] >= hi[0]:
            return (1, p[1])
        if a == X:
            return (2, p[0])
        return (3, p[1])

    return sorted(ents, key=key)


def build(engine, g):
    """Registers the excitation and the probes (map_sources order: the
    excitation, then the probes ut1, it1, et1 (x, y, z), ht1 (x, y, z))."""
    zl = g.lines[2]
    w = 1.0 * (zl[10] - zl[9])
    wave = E.Waveform.openems_gaussian(1.5e9, 1.5e9)
    engine.add_operator(E.EdgeSource(EE.OpenEmsSignal(wave)).kind(E.SourceKind.SoftE)
                        .edge(E.Axis.Z, [10, 10, 9], w).edge(E.Axis.Z, [10, 10, 10], w))
    probes = {}
    probes["ut1"] = [engine.add_operator(E.Probe(E.ProbeQuantity.Voltage)
                                         .entry(E.Axis.Z, [12, 10, 9], 1.0)
                                         .entry(E.Axis.Z, [12, 10, 10], 1.0))]
    it = E.Probe(E.ProbeQuantity.Current).f32_sum()
    for a, p, wt in current_loop_entries():
        it.entry(a, p, wt)
    probes["it1"] = [engine.add_operator(it)]
    node = [13, 10, 10]
    probes["et1"] = [engine.add_operator(E.Probe(E.ProbeQuantity.Voltage).entry(
        a, node, 1.0 / (g.lines[a][node[a] + 1] - g.lines[a][node[a]]))) for a in range(3)]
    dual = [12, 9, 9]
    probes["ht1"] = [engine.add_operator(E.Probe(E.ProbeQuantity.Current).entry(
        a, dual, 1.0 / g.dual_delta(a, dual[a]))) for a in range(3)]
    return probes


def sampled(kind, values):
    """The fixture's rows from per-step samples: E-time files at t = 21 m dt
    (row 0 the initial zero, row m > 0 the sample of step 21 m - 1), H-time
    files at (21 m + 1/2) dt (the sample of step 21 m)."""
    rows = STEPS // INTERVAL + 1
    if kind in ("ut1", "et1"):
        return np.array([0.0] + [values[INTERVAL * m - 1] for m in range(1, rows)])
    return np.array([values[INTERVAL * m] for m in range(rows)])


def rel_l2(a, b):
    """tests/openems_core.rs:52-60: |a - b| / |a| over every value."""
    num = float(np.sum((a - b) ** 2))
    den = float(np.sum(a * a))
    return np.sqrt(num) if den == 0.0 else np.sqrt(num / den)
