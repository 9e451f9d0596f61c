"""Lumped elements, ports, DFT and port analysis on the CPU: the host side
(install, fold_parallel, dft, PortResult) and the oracle's restatement of
the lumped hooks, pinned by the closed forms of the reference's own tests
(src/engine/lumped.rs:1283-1500, tests/lumped.rs, src/post/dft.rs:224-330,
src/post/port.rs:243-330)."""
import cmath
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import post
from ems_amd.lumped import (EdgeValues, LumpedBox, LumpedElement, LumpedPort, Topology,
                            fold_parallel)
from ems_amd._native import EmsError

import cases as K

TAU = 2.0 * math.pi


def rel(a, b):
    return abs((a - b) / b)


# ---------------------------------------------------------------- host --

def test_box_counts_and_validation():
    """lumped.rs:1287-1331."""
    b = LumpedBox(E.Axis.Z, [1, 2, 3], [2, 4, 6])
    assert b.series_edges() == 3 and b.columns() == 6
    e = b.edges()
    assert len(e) == 18 and e[0] == (1, 2, 3) and e[2] == (1, 2, 5)
    for args in ((E.Axis.Z, [1, 2, 3], [1, 2, 3]), (E.Axis.X, [2, 0, 0], [1, 0, 0]),
                 (E.Axis.X, [0, 1, 0], [1, 0, 0])):
        with pytest.raises(EmsError):
            LumpedBox(*args)
    b.check_fits([3, 5, 7])
    with pytest.raises(EmsError):
        b.check_fits([3, 5, 6])
    r = LumpedBox(E.Axis.Y, [0, 0, 0], [0, 1, 0])
    for el in (LumpedElement.resistor(r, -1.0), LumpedElement.resistor(r, float("nan")),
               LumpedElement.new(r), LumpedElement.new(r).c(1e-12).source(lambda t: 1.0)):
        with pytest.raises(EmsError):
            el.validate()
    LumpedElement.new(r).c(1e-12).topology(Topology.Series).source(lambda t: 1.0).validate()
    assert LumpedElement.new(r).l(1e-9).has_branch()
    assert not LumpedElement.new(r).r(1.0).c(1e-12).has_branch()
    assert LumpedElement.new(r).r(1.0).c(1e-12).topology(Topology.Series).has_branch()


def test_fold_composes_with_conductive_material():
    """lumped.rs:1333-1362: folding g into a conductive edge equals a
    material with G + g from the start."""
    g = E.Grid.uniform([4, 4, 4], [1e-3] * 3)
    kappa = 0.7
    m = E.CellMaterials(g, E.Material.dielectric(2.0).with_kappa(kappa))
    dt = 1e-12
    lanes = E.coefficients(g, m, [E.FaceCondition.Pec] * 6, dt)
    i = g.index([1, 2, 1])
    ca0, cb0 = float(lanes[2].reshape(-1)[i]), float(lanes[5].reshape(-1)[i])
    c0, g0 = E.EPS0 * 2.0 * 1e-3, kappa * 1e-3  # interior edge: C = eps A/l, G = kappa A/l
    gl, cl = 0.02, 3e-14
    ca, cb = fold_parallel(ca0, cb0, dt, gl, cl)
    cc, gg = c0 + cl, g0 + gl
    want_ca = (cc - 0.5 * gg * dt) / (cc + 0.5 * gg * dt)
    want_cb = dt / (cc + 0.5 * gg * dt)
    assert rel(ca, want_ca) < 1e-6 and rel(cb, want_cb) < 1e-6
    a0, b0 = fold_parallel(ca0, cb0, dt, 0.0, 0.0)
    assert rel(a0, ca0) < 1e-6 and rel(b0, cb0) < 1e-6
    a1, b1 = fold_parallel(ca0, cb0, dt, 0.01, 1e-14)
    a2, b2 = fold_parallel(a1, b1, dt, 0.01, 2e-14)
    assert rel(a2, want_ca) < 1e-5 and rel(b2, want_cb) < 1e-5
    assert rel(dt * (1.0 + ca) / (2.0 * cb), cc) < 1e-6  # Coefficients::capacitance
    with pytest.raises(EmsError):
        fold_parallel(0.0, 0.0, dt, 1.0, 0.0)


def test_install_distributes_over_edges():
    """lumped.rs:1364-1401: g_e = M/(N R), C_e = M C/N folded into the box
    edges only; PEC and out-of-grid boxes are rejected."""
    g = E.Grid.uniform([6, 6, 6], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    bx = LumpedBox(E.Axis.X, [1, 2, 2], [4, 3, 3])
    el = LumpedElement.new(bx).r(30.0).c(2e-12).source(lambda t: 1.0)
    plan = el._plan(g, dt, lanes)
    ge, ce = 3.0 / (4.0 * 30.0), 3.0 * 2e-12 / 4.0
    assert rel(plan["g_src"][0], ge) < 1e-15
    ca, cb = plan["ca"].reshape(-1), plan["cb"].reshape(-1)
    ca0, cb0 = lanes[0].reshape(-1), lanes[3].reshape(-1)
    for p in bx.edges():
        i = g.index(p)
        assert (ca[i], cb[i]) == tuple(np.float32(x) for x in fold_parallel(ca0[i], cb0[i], dt, ge, ce))
    i = g.index([0, 2, 2])
    assert cb[i] == cb0[i]
    assert plan["terms"][0][1] == [1.0 / 3.0] * 12
    with pytest.raises(EmsError):
        LumpedElement.resistor(LumpedBox(E.Axis.Y, [0, 1, 1], [0, 2, 1]), 50.0)._plan(g, dt, lanes)
    with pytest.raises(EmsError):
        LumpedElement.resistor(LumpedBox(E.Axis.Y, [1, 1, 1], [1, 7, 1]), 50.0)._plan(g, dt, lanes)


# ------------------------------------------------------- oracle hooks --

def _small(kind, el_fn, steps):
    g = E.Grid.uniform([5, 5, 5], [1e-3] * 3)
    e = K.make(kind, g, threads=1)
    h = e.add_operator(el_fn())
    e.run(steps)
    return e, h


def test_branch_resistor_matches_fold():
    """lumped.rs:1403-1428: a series R with a huge C (trapezoidal branch)
    behaves like the folded resistor."""
    bx = LumpedBox(E.Axis.Z, [2, 2, 1], [2, 2, 3])

    def run(series_rc):
        el = LumpedElement.resistor(bx, 20.0).source(lambda t: math.sin(t * 3e10)) \
            .topology(Topology.Series)
        if series_rc:
            el = el.c(1.0)
        assert el.has_branch() == series_rc
        e, _ = _small("oracle", lambda: el, 300)
        return e.u_component(E.Axis.Z)

    a, b = run(False), run(True)
    scale = float(np.max(np.abs(a)))
    assert scale > 0.0
    assert np.all(np.abs(a - b) <= 1e-4 * scale)


@pytest.mark.parametrize("topo", [Topology.Parallel, Topology.Series])
def test_edge_list_matches_box(topo):
    """lumped.rs:1430-1470: the edge list with distributed values and 1/M
    weights reproduces the box element."""
    r, l, c = 40.0, 2e-9, 1e-12
    bx = LumpedBox(E.Axis.Z, [2, 2, 1], [3, 2, 3])
    m, n = 2.0, 2.0
    src = lambda t: math.sin(t * 2e10)  # noqa: E731
    ea, ha = _small("oracle", lambda: LumpedElement.new(bx).r(r).l(l).c(c).topology(topo)
                    .source(src), 200)
    edges = [EdgeValues(p, r=n * r / m, l=n * l / m, c=m * c / n) for p in bx.edges()]
    el = LumpedElement.from_edges(E.Axis.Z, topo, edges).edge_source([1.0 / m] * len(edges), src)
    assert el.region() is None and el.series_edges() == 2
    eb, hb = _small("oracle", lambda: el, 200)
    a, b = ea.u_component(E.Axis.Z), eb.u_component(E.Axis.Z)
    scale = float(np.max(np.abs(a)))
    assert scale > 0.0 and np.all(np.abs(a - b) <= 1e-5 * scale)
    assert abs(ha.branch_current() - hb.branch_current()) <= 1e-5 * max(abs(ha.branch_current()), 1e-12)
    ev = lambda node, r: EdgeValues(node, r=r)  # noqa: E731
    for bad in ([ev([1, 1, 1], 1.0), ev([1, 1, 2], None)], [ev([1, 1, 1], 1.0), ev([1, 1, 1], 1.0)], []):
        with pytest.raises(EmsError):
            LumpedElement.from_edges(E.Axis.Z, Topology.Parallel, bad)


def _tail(x, n=200):
    s = -0.0
    for v in x[-n:]:
        s += float(v)
    return s / n


def test_port_dc_limits():
    """tests/lumped.rs:161-215: open, short and 150-ohm-loaded ports settle
    at V_s Z_L/(R + Z_L) and I = V_s/(R + Z_L)."""
    tr, r = 0.5e-9, 50.0
    g = E.Grid.uniform([12, 12, 12], [1e-3] * 3)
    e = K.make("oracle", g, threads=2)
    p = e.add_operator(LumpedPort(LumpedBox(E.Axis.Z, [6, 6, 5], [6, 6, 7]), r).excite(K.ramp(tr)))
    e.run(3000)
    v, i = _tail(p.voltage()), _tail(p.current())
    assert abs(v - 1.0) < 1e-4 and abs(i) < 1e-6
    g = E.Grid.uniform([12, 12, 4], [1e-3] * 3)
    e = K.make("oracle", g, threads=2)
    p = e.add_operator(LumpedPort(LumpedBox(E.Axis.Z, [6, 6, 0], [6, 6, 4]), r).excite(K.ramp(tr)))
    e.run(3000)
    v, i = _tail(p.voltage()), _tail(p.current())
    assert abs(v) < 1e-4 and abs(i * r - 1.0) < 1e-4
    g = E.Grid.uniform([20, 20, 20], [1e-3] * 3)
    e = K.make("oracle", g, K.loop_materials(g), threads=2)
    p = e.add_operator(LumpedPort(K.loop_column(K.LOOP["port_x"]), r).excite(K.ramp(tr)))
    e.add_operator(LumpedElement.resistor(K.loop_column(K.LOOP["load_x"]), 150.0))
    e.run(3000)
    v, i = _tail(p.voltage()), _tail(p.current())
    assert abs(v - 0.75) < 1e-4 and abs(i - 5e-3) < 1e-6


def test_loop_resistor_impedance():
    """tests/lumped.rs:224-259: 100-ohm load behind the loop inductance;
    S11 is the reflection of Z_in; power balance."""
    rl = 100.0
    _, p = K.run_loop("oracle", lambda b: LumpedElement.resistor(b, rl))
    res = p.analyze(K.FREQS)
    z, s = res.z_in(), res.s11()
    for k, f in enumerate(K.FREQS):
        l_eff = z[k].imag / (TAU * f)
        assert 3e-9 < l_eff < 5e-9, l_eff
        if f <= 500e6:
            assert abs(z[k].real / rl - 1.0) < 0.01
        if f <= 1e9:
            assert abs(z[k].real / rl - 1.0) < 0.02
        assert abs(s[k] - (z[k] - K.Z0_PORT) / (z[k] + K.Z0_PORT)) < 1e-9
    assert abs(s[0] - (rl - 50.0) / (rl + 50.0)) < 0.01
    pi, pr, pa = res.incident_power(), res.reflected_power(), res.accepted_power()
    for k in range(len(K.FREQS)):
        assert abs(pi[k] - pr[k] - pa[k]) < 1e-9 * pi[k] and pa[k] > 0.0


def _calibration():
    """Short-open-load one-port calibration (tests/lumped.rs:262-302)."""
    zs = K.run_loop("oracle", "short")[1].analyze(K.FREQS).z_in()
    zo = K.run_loop("oracle", "open")[1].analyze(K.FREQS).z_in()
    zr = K.run_loop("oracle", lambda b: LumpedElement.resistor(b, 100.0))[1].analyze(K.FREQS).z_in()
    c = [(zs[k] - zr[k]) / ((zr[k] - zo[k]) * 100.0) for k in range(len(K.FREQS))]
    a = [zo[k] * c[k] for k in range(len(K.FREQS))]
    return a, zs, c


def test_lumped_elements_deembedded_quick():
    """tests/lumped.rs:336-407 (quick, tolerance 1e-3): R, C, L, series RL,
    series RLC and parallel RLC de-embedded through the calibration match
    their nominal impedances."""
    a, b, c = _calibration()

    def zser(r, l, cap):
        return lambda w: complex(r, w * l + (-1.0 / (w * cap) if cap > 0 else 0.0))

    def zpar(r, l, cap):
        def f(w):
            y = complex(0.0, w * cap)
            if r > 0:
                y += 1.0 / r
            if l > 0:
                y += complex(0.0, -1.0 / (w * l))
            return 1.0 / y
        return f

    cases = [
        (lambda bx: LumpedElement.resistor(bx, 30.0), zser(30.0, 0.0, 0.0)),
        (lambda bx: LumpedElement.new(bx).c(1e-12), zpar(0.0, 0.0, 1e-12)),
        (lambda bx: LumpedElement.new(bx).l(10e-9), zpar(0.0, 10e-9, 0.0)),
        (lambda bx: LumpedElement.new(bx).r(40.0).l(3e-9).topology(Topology.Series),
         zser(40.0, 3e-9, 0.0)),
        (lambda bx: LumpedElement.new(bx).r(20.0).l(5e-9).c(2e-12).topology(Topology.Series),
         zser(20.0, 5e-9, 2e-12)),
        (lambda bx: LumpedElement.new(bx).r(200.0).l(8e-9).c(0.5e-12), zpar(200.0, 8e-9, 0.5e-12)),
    ]
    for load, want in cases:
        zin = K.run_loop("oracle", load)[1].analyze(K.FREQS).z_in()
        for k, f in enumerate(K.FREQS):
            zl = (b[k] - zin[k]) / (c[k] * zin[k] - a[k])
            w = want(TAU * f)
            assert abs(zl - w) / abs(w) < 1e-3, (f, zl, w)


def test_loop_lc_resonance():
    """tests/lumped.rs:413-440: the loop inductance resonates with 1 pF a
    few percent below 1/(2 pi sqrt(L C))."""
    l = K.run_loop("oracle", "short", freqs=[100e6])[1].analyze([100e6]).z_in()[0].imag / (TAU * 100e6)
    cap = 1e-12
    f_lc = 1.0 / (TAU * math.sqrt(l * cap))
    freqs = [1.5e9 + q * 7.5e6 for q in range(201)]
    r = K.run_loop("oracle", lambda b: LumpedElement.new(b).c(cap))[1].analyze(freqs)
    x = [z.imag for z in r.z_in()]
    q = next(k for k in range(200) if x[k] < 0.0 <= x[k + 1])
    f_r = freqs[q] - x[q] * (freqs[q + 1] - freqs[q]) / (x[q + 1] - x[q])
    assert 0.95 * f_lc < f_r < f_lc
    assert abs(r.z_in()[q].real) < 0.5


def test_two_port_reciprocity_and_determinism():
    """tests/lumped.rs:500-555: S21 = S12 and passivity; bit-identical
    results for any thread count."""
    e1, a1, a2 = K.two_port("oracle", 1, threads=1)
    r1, r2 = a1.analyze(K.FREQS), a2.analyze(K.FREQS)
    s21, s11 = post.s_param(r2, r1), r1.s11()
    _, b1, b2 = K.two_port("oracle", 2, threads=2)
    q1, q2 = b1.analyze(K.FREQS), b2.analyze(K.FREQS)
    s12, s22 = post.s_param(q1, q2), q2.s11()
    for k in range(len(K.FREQS)):
        assert abs(s21[k]) > 0.05
        assert abs(s21[k] - s12[k]) < 2e-3 * abs(s21[k])
        assert abs(s11[k]) ** 2 + abs(s21[k]) ** 2 <= 1.0 + 1e-6
        assert abs(s22[k]) ** 2 + abs(s12[k]) ** 2 <= 1.0 + 1e-6
    _, c1, c2 = K.two_port("oracle", 1, threads=3)
    assert np.array_equal(c2.voltage().view(np.uint64), a2.voltage().view(np.uint64))
    assert np.array_equal(c1.current().view(np.uint64), a1.current().view(np.uint64))


def test_lc_loaded_cavity_conserves_total_energy():
    """tests/lumped.rs:585-640: W_fields + W_branch is conserved; the field
    part alone is not; the in-step energy equals the read-only one."""
    for threads in (1, 3):
        e, hp, hs = K.lc_cavity("oracle", threads)
        e.run(5)
        w0 = e.energy()
        b0 = -0.0 + hp.branch_energy() + hs.branch_energy()
        assert w0 > 0.0
        dev = fdev = bmax = 0.0
        for _ in range(400):
            w_ro = e.energy()
            w_ops = -0.0 + hp.branch_energy() + hs.branch_energy()
            w = e.step_with_energy()
            assert np.float64(w).view(np.uint64) == np.float64(w_ro).view(np.uint64)
            dev = max(dev, abs(w / w0 - 1.0))
            fdev = max(fdev, abs((w - w_ops) / (w0 - b0) - 1.0))
            bmax = max(bmax, w_ops / w)
            e.run(9)
        assert dev < 1e-4, dev
        assert bmax > 0.05, bmax
        assert fdev > 100.0 * dev


# ----------------------------------------------------- DFT and ports --

def _naive(t0, dt, xs, f):
    return sum(cmath.exp(complex(0.0, -TAU * f * (t0 + n * dt))) * x for n, x in enumerate(xs)) * dt


def test_dft_matches_naive_sum():
    """dft.rs:239-256."""
    dt = 1.7e-12
    t0 = 0.5 * dt
    xs = [((n * 7919) % 1000) / 1000.0 - 0.5 for n in range(5000)]
    freqs = [0.0, 1.3e9, 7.77e9, 41e9]
    got = post.dft(t0, dt, xs, freqs)
    for g, f in zip(got, freqs):
        assert abs(g - _naive(t0, dt, xs, f)) < 1e-12 * len(xs) * dt


def test_dft_dc_and_impulse():
    """dft.rs:258-266 and the doc examples (dft.rs:60-72, 213-217)."""
    assert post.dft(0.0, 2.0, [1.0, 1.0, 1.0], [0.0])[0] == complex(6.0, 0.0)
    t0 = 0.125
    d = post.dft(t0, 1.0, [1.0], [1.0])[0]
    assert abs(d - cmath.exp(complex(0.0, -TAU * t0))) < 1e-15
    x = post.dft(0.0, 0.5, [1.0, 0.0, 0.0], [0.1, 0.3])
    assert abs(x[1].real - 0.5) < 1e-15 and abs(x[1].imag) < 1e-15
    dt = 1e-12
    xs = [math.cos(TAU * 1e9 * (n * dt)) for n in range(1000)]
    assert abs(post.dft(0.0, dt, xs, [1e9, 2e9])[0].real - 0.5e-9) < 1e-18


def test_dft_no_drift_over_many_samples():
    """dft.rs:291-320: 2,001,000 samples of an exact tone, E and H instants."""
    dt = 1e-12
    f = 0.125 / dt
    n_samples = 2_001_000
    phi = 0.3
    q = np.arange(n_samples) % 8
    for t0 in (0.0, 0.5 * dt):
        c0 = f * t0
        xs = np.cos(TAU * (c0 + q / 8.0) + phi)
        got = post.dft(t0, dt, xs, [f])[0]
        expect = cmath.exp(complex(0.0, phi)) * (0.5 * n_samples * dt)
        assert abs(got - expect) / abs(expect) < 1e-11


def test_port_result_waves():
    """post/port.rs:251-320."""
    zl = complex(100.0, 30.0)
    i = complex(2.0, 0.0) / (zl + 50.0)
    v = zl * i
    p = post.PortResult([1e9], [v], [i], complex(50.0, 0.0))
    assert abs(p.s11()[0] - (zl - 50.0) / (zl + 50.0)) < 1e-14
    assert abs(p.z_in()[0] - zl) < 1e-12
    assert abs(p.a()[0] - complex(1.0 / math.sqrt(50.0), 0.0)) < 1e-14
    pi, pr, pa = p.incident_power()[0], p.reflected_power()[0], p.accepted_power()[0]
    assert abs(pi - pr - pa) < 1e-14 * pi
    q = p.with_z_ref(complex(100.0, -30.0))
    assert abs(q.s11()[0]) < 1e-14
    assert abs(q.incident_power()[0] - q.accepted_power()[0]) < 1e-14
    q = p.with_z_ref(complex(20.0, 70.0))
    assert abs(q.incident_power()[0] - q.reflected_power()[0] - q.accepted_power()[0]) < \
        1e-13 * q.incident_power()[0]
    z = [complex(50.0, 0.0), complex(60.0, -5.0)]
    ii = [complex(0.01, 0.002), complex(-0.003, 0.02)]
    vv = [zl * x for x in ii]
    p = post.PortResult([1e9, 2e9], vv, ii, z)
    vi, vr, ci, cr = p.v_inc(), p.v_ref(), p.i_inc(), p.i_ref()
    for k in range(2):
        assert abs(vi[k] + vr[k] - vv[k]) < 1e-14
        assert abs(ci[k] + cr[k] - ii[k]) < 1e-16
        assert abs(vi[k] / ci[k] - z[k]) < 1e-12
        assert abs(vr[k] / cr[k] + z[k]) < 1e-12
    assert abs(vr[0] / vi[0] - (zl - z[0]) / (zl + z[0])) < 1e-14
    with pytest.raises(EmsError):
        post.PortResult([1e9, 2e9], vv, ii, [complex(50.0, 0.0), complex(0.0, 1.0)])
    with pytest.raises(EmsError):
        post.PortResult([1e9], [0j], [0j], complex(0.0, 1.0))
