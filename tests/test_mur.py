"""The Mur boundary (src/engine/boundary/mur.rs) on the oracle, pinned by
the reference's own tests: the edge counts of its unit test (mur.rs:
886-905), normal-incidence reflection below -30 dB on every axis of a TEM
line against a long reference line (tests/mur.rs:96-126), energy leaving an
all-Mur box (tests/mur.rs:185-196) and staying gone for 3000 steps on a
graded mesh (tests/mur.rs:198-226), the same bits at any thread count and in
the fused and unfused schedules (tests/mur.rs:232-250, tests/fused.rs:110)."""
import math

import numpy as np
import pytest

import ems_amd as E
from oracle.bridge import OracleEngine

D = 1e-3
TAU = 2.8e-11
C0 = 299792458.0


def gaussian(tau):
    return lambda t: math.exp(-((t - 4.0 * tau) / tau) ** 2)


def dgaussian(tau):
    def f(t):
        x = (t - 4.0 * tau) / tau
        return -2.0 * x * math.exp(-x * x)
    return f


def make(g, faces, threads=1, fused=0):
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, None, faces, dt)
    return OracleEngine(g, dt, lanes, threads=threads, fused=fused)


def test_counts_and_validation():
    g = E.Grid.uniform([4, 3, 5], [1e-3] * 3)
    o = make(g, [E.FaceCondition.Pec] * 6)
    m = E.MurBoundary.new(o, E.Face.XMin, 3e8)
    assert m.edge_count() == 3 * 6 + 4 * 5 and len(m.corners) == 0
    m = E.MurBoundary.with_faces(o, [(E.Face.XMin, 3e8), (E.Face.ZMin, 3e8)])
    assert len(m.corners) == 3
    assert m.edge_count() == (3 * 6 + 4 * 5) + (4 * 4 + 5 * 3) - 3
    m = E.MurBoundary.with_faces(o, [(f, 3e8) for f in E.Face])
    assert len(m.corners) == 4 * (4 + 3 + 5)
    for bad in ([(E.Face.XMin, 0.0)], [(E.Face.XMin, float("nan"))],
                [(E.Face.YMax, 1.0), (E.Face.YMax, 1.0)]):
        with pytest.raises(E.EmsError):
            E.MurBoundary.with_faces(o, bad)
    faces = [E.FaceCondition.Pec] * 6
    faces[E.Face.ZMax] = E.FaceCondition.Pmc
    o2 = make(g, faces)
    with pytest.raises(E.EmsError):
        E.MurBoundary.new(o2, E.Face.ZMax, 3e8)
    E.MurBoundary.new(o2, E.Face.ZMin, 3e8)
    o.close()
    o2.close()


def line(a, cells, src, probes, mur, steps):
    """A quasi-1-D TEM line along a (tests/mur.rs:30-94): E along b = a+1,
    PEC on the b faces, PMC on the c = a+2 faces, Mur (or PEC) on the a
    faces, a soft source on plane src; the E voltage at each probe plane."""
    b, c = (a + 1) % 3, (a + 2) % 3
    n = [1, 1, 1]
    n[a] = cells
    g = E.Grid.uniform(n, [D] * 3)
    faces = [E.FaceCondition.Pec] * 6
    faces[2 * c] = faces[2 * c + 1] = E.FaceCondition.Pmc
    o = make(g, faces)
    if mur:
        o.add_operator(E.MurBoundary.with_faces(o, [(E.Face(2 * a), C0), (E.Face(2 * a + 1), C0)]))

    def node(pos, q):
        p = [0, 0, 0]
        p[a], p[c] = pos, q
        return p
    o.add_operator(E.EdgeSource(gaussian(TAU), horizon=steps + 1)
                   .edge(E.Axis(b), node(src, 0), 1.0).edge(E.Axis(b), node(src, 1), 1.0))
    ps = [o.add_operator(E.Probe.voltage(E.Axis(b), node(p, 0))) for p in probes]
    o.run(steps)
    ub = o.u_component(b)
    for p in probes:  # the wave stays uniform across the line (exact TEM)
        v0, v1 = ub[tuple(node(p, 0))], ub[tuple(node(p, 1))]
        assert abs(v0 - v1) <= 1e-5 * max(abs(v0), 1e-6)
    out = [o.probe_values(p) for p in ps]
    o.close()
    return out


@pytest.mark.parametrize("a", [0, 1, 2])
def test_normal_incidence_reflection_below_30db(a):
    cells, src, probes, ext = 120, 60, [30, 90], 150
    dt = 0.99 * D / (C0 * math.sqrt(3.0))
    steps = int(8.5e-10 / dt)
    ref = line(a, cells + 2 * ext, src + ext, [p + ext for p in probes], False, steps)
    mur = line(a, cells, src, probes, True, steps)
    for r, m in zip(ref, mur):
        peak = np.max(np.abs(r))
        db = 20.0 * math.log10(np.max(np.abs(r - m)) / peak)
        assert db < -30.0, db
    # without Mur the PEC face reflects the whole pulse (tests/mur.rs:127-143)
    pec = line(a, cells, src, probes, False, steps)
    db = 20.0 * math.log10(np.max(np.abs(ref[0] - pec[0])) / np.max(np.abs(ref[0])))
    assert db > -1.0


def mur_box(cells, threads, graded, fused=0):
    """tests/mur.rs:147-171: Mur on all six faces, a dipole pulse inside."""
    def lines(s):
        out = []
        for q in range(cells + 1):
            t = q / cells
            w = 0.2 * math.sin(6.0 * t) / 6.0 if graded else 0.0
            out.append(s * cells * (t + w))
        return np.array(out)
    g = E.Grid([lines(D), lines(1.1 * D), lines(0.9 * D)])
    o = make(g, [E.FaceCondition.Pec] * 6, threads=threads, fused=fused)
    o.add_operator(E.MurBoundary.with_faces(o, [(f, C0) for f in E.Face]))
    m = cells // 2
    src = E.EdgeSource(dgaussian(TAU), horizon=100000).edge(E.Axis.Z, [m, m, m], 1.0) \
        .edge(E.Axis.X, [m - 2, m + 1, m], 0.3)
    o.add_operator(src)
    return o


def drain(o, steps):
    peak = 0.0
    for _ in range(steps):
        o.run(1)
        peak = max(peak, o.energy())
    return peak, o.energy()


def test_energy_leaves_through_mur_faces():
    o = mur_box(16, 2, False)
    steps = int((8.0 * TAU + 3.0e-10) / o.dt)
    peak, end = drain(o, steps)
    assert end < 1e-2 * peak, (end, peak)
    o.close()


def test_long_run_stable_graded():
    o = mur_box(12, 2, True)
    settle = int((8.0 * TAU + 3.0e-10) / o.dt)
    peak, w0 = drain(o, settle)
    wmax = 0.0
    for q in range(3000 // 50):
        o.run(50)
        wmax = max(wmax, o.energy())
    assert np.isfinite(o.energy())
    assert w0 < 1e-9 * peak and wmax < 1e-9 * peak, (peak, w0, wmax)
    o.close()


def test_same_bits_any_threads_and_schedule():
    def run(threads, fused):
        o = mur_box(10, threads, True, fused=fused)
        o.run(250)
        out = [o.u_component(a).view(np.uint32).copy() for a in range(3)] + \
              [o.i_component(a).view(np.uint32).copy() for a in range(3)]
        o.close()
        return out
    base = run(1, 2)
    for threads, fused in ((3, 2), (1, 1), (4, 1)):
        for x, y in zip(base, run(threads, fused)):
            assert np.array_equal(x, y)
