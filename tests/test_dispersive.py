"""Dispersive media on the CPU: the host install and the oracle's
restatement of the pole corrections, pinned by the reference's validation
(tests/validation_dispersive.rs): energy conservation of lossless poles,
dissipation of lossy ones, thread determinism, and the reflection of a
Drude half-space against the continuum and the discrete (bilinear-map)
response of the Yee line."""
import cmath
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd._native import EmsError
from ems_amd.dispersive import Dispersive, Field, Pole, Scheme

import cases as K

TAU = 2.0 * math.pi


def test_pole_forms():
    """dispersive.rs:218-425: Drude = Lorentz with w0 = 0; Debye static
    limit; the discrete forms at the bilinear / Yee frequencies."""
    wp, g = 3e10, 2e9
    for w in (1e9, 3e10, 1e11):
        assert abs(Pole.drude(wp, g).susceptibility(w) - Pole.lorentz(wp, 0.0, g).susceptibility(w)) \
            < 1e-12 * abs(Pole.drude(wp, g).susceptibility(w))
    assert Pole.debye(4.0, 20e-12).susceptibility(0.0) == 4.0
    dt, f = 1e-12, 5e9
    wt = 2.0 / dt * math.tan(math.pi * f * dt)
    p = Pole.lorentz(wp, 2e10, g)
    assert p.discrete_susceptibility(f, dt) == p.susceptibility(wt)
    om = 2.0 / dt * math.sin(math.pi * f * dt)
    c = math.cos(math.pi * f * dt)
    want = wp * wp / complex(2e10 * 2e10 - om * om, om * g * c)
    assert abs(p.leapfrog_susceptibility(f, dt) - want) < 1e-12 * abs(want)
    assert Pole.debye(0.0, 1e-12).is_null() and Pole.drude(0.0, 1.0).is_null()
    with pytest.raises(EmsError):
        Pole.lorentz(1.0, -1.0, 0.0).validate()
    s = Pole.lorentz(wp, 2e10, g).scaled(0.25)
    assert s.susceptibility(3e10) == pytest.approx(0.25 * p.susceptibility(3e10), rel=1e-12)


def test_fill_cells_shares_and_install():
    """fill_cells (dispersive.rs:1896-1986): interior edges carry the full
    pole, face edges half, edge-of-box edges a quarter; PEC edges are
    skipped at install."""
    g = E.Grid.uniform([6, 6, 6], [1e-3] * 3)
    d = Dispersive.new()
    d.fill_cells(g, [(2, 4), (2, 4), (2, 4)], Field.Electric, [Pole.debye(2.0, 1e-11)])
    shares = {(a, n): p.p[0] / 2.0 for _, a, n, p in d.entries}
    assert shares[(0, (2, 3, 3))] == 1.0
    assert shares[(0, (2, 2, 3))] == 0.5
    assert shares[(0, (2, 2, 2))] == 0.25
    assert all(n[0] in (2, 3) for _, a, n, _ in d.entries if a == 0)
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    d2 = Dispersive.new()
    d2.push(Field.Electric, 0, (1, 0, 3), Pole.debye(1.0, 1e-12))  # y = 0 face: PEC
    d2.push(Field.Electric, 0, (1, 2, 3), Pole.debye(1.0, 1e-12))
    d2.push(Field.Electric, 0, (1, 2, 3), Pole.drude(1e10, 1e9))
    plan = d2._plan(g, dt, lanes)
    assert d2.skipped() == 1
    s = plan["sides"][0]
    assert s["pole_end"] == [2] and len(s["cidx"]) == 2
    m = E.EPS0 * 1e-3 * 1e-3 / 1e-3
    a, inv_c, z = Pole.debye(1.0, 1e-12).branch(m, dt)
    assert plan["consts"][s["cidx"][0]][:3] == (a, inv_c, 1.0 / z)
    assert s["sum_inv_z"][0] == 0.0 + 1.0 / z + plan["consts"][s["cidx"][1]][2]
    with pytest.raises(EmsError):
        bad = Dispersive.new()
        bad.push(Field.Electric, 2, (1, 1, 6), Pole.debye(1.0, 1e-12))
        bad._plan(g, dt, lanes)


def _drift(steps, magnetic):
    """energy_drift_with (validation_dispersive.rs:431-444)."""
    e, h = K.disp_cavity("oracle", lossless=True, magnetic=magnetic)
    e.run(10)
    w0 = e.energy()
    lo = hi = w0
    for _ in range(steps // 50):
        e.run(50)
        w = e.energy()
        lo, hi = min(lo, w), max(hi, w)
    return (hi - lo) / w0, h.energy() / w0


def test_lossless_poles_conserve_energy():
    """validation_dispersive.rs:446-456: drift < 1e-5 over 4000 steps with
    a sizeable pole share."""
    drift, share = _drift(4000, False)
    assert drift < 1e-5, drift
    assert share > 0.01, share


def test_lossless_magnetic_poles_conserve_energy():
    """validation_dispersive.rs:458-475, and step_with_energy == energy."""
    drift, _ = _drift(4000, True)
    assert drift < 1e-5, drift
    e, h = K.disp_cavity("oracle", lossless=True, magnetic=True)
    e.run(37)
    assert h.magnetic_energy() > 0.0
    w = e.energy()
    assert np.float64(w).tobytes() == np.float64(e.step_with_energy()).tobytes()


def test_lossy_poles_dissipate():
    """validation_dispersive.rs:486-501."""
    e, _ = K.disp_cavity("oracle", lossless=False)
    e.run(2)
    w = w0 = e.energy()
    for _ in range(200):
        e.run(10)
        wn = e.energy()
        assert wn <= w * (1.0 + 1e-6)
        w = wn
    assert w < 0.5 * w0


def test_magnetic_poles_stay_bounded():
    """validation_dispersive.rs:503-528."""
    for lossless in (True, False):
        e, _ = K.disp_cavity("oracle", lossless=lossless, magnetic=True)
        e.run(2)
        w0 = w = e.energy()
        wmax = 0.0
        for _ in range(60):
            e.run(50)
            wn = e.energy()
            if not lossless:
                assert wn <= w * (1.0 + 1e-6)
            w = wn
            wmax = max(wmax, wn)
        assert wmax < 1.05 * w0


def test_deterministic_across_threads():
    """validation_dispersive.rs:530-555."""
    def run(t):
        e, h = K.disp_cavity("oracle", threads=t, lossless=False, magnetic=True)
        e.run(300)
        f = [e.u_component(a) for a in range(3)] + [e.i_component(a) for a in range(3)]
        return f, (h.electric_energy(), h.magnetic_energy(), e.energy())
    f1, w1 = run(1)
    f4, w4 = run(4)
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(f1, f4))
    assert [np.float64(x).tobytes() for x in w1] == [np.float64(x).tobytes() for x in w4]


# ---- reflection of a Drude half-space (validation_dispersive.rs:156-265)

DZ, IFACE, SRC, PRB = 0.5e-3, 148, 48, 88


def _csqrt(z):
    r, t = abs(z), cmath.phase(z)
    return cmath.rect(math.sqrt(r), 0.5 * t)


def _casin(z):
    """asin z = -j ln(j z + sqrt(1 - z^2)) (tests/common/mod.rs:89-93)."""
    return -1j * cmath.log(1j * z + _csqrt(1.0 - z * z))


def _yee_k_complex(om, dz, eps, mu):
    x = eps * mu * (om * om * E.MU0 * E.EPS0 * dz * dz)
    k = _casin(_csqrt(x) * 0.5) * 2.0
    if k.imag > 1e-15 or (abs(k.imag) <= 1e-15 and k.real < 0.0):
        k = -k
    return k


def _discrete_response(eps, mu, dz, dt, f):
    """discrete_response (tests/common/mod.rs:348-374): reflection of the
    discrete 1-D Yee line with per-cell eps, mu."""
    n = len(eps)
    om = 2.0 / dt * math.sin(math.pi * f * dt)
    x = om * om * E.MU0 * E.EPS0 * dz * dz
    kr = _yee_k_complex(om, dz, eps[n - 1], mu[n - 1])
    kl = _yee_k_complex(om, dz, eps[0], mu[0])
    e = [0j] * (n + 1)
    e[n] = 1.0 + 0j
    e[n - 1] = cmath.exp(1j * kr)
    for k in range(n - 1, 0, -1):
        eps_k = (eps[k - 1] + eps[k]) * 0.5
        e[k - 1] = e[k] - mu[k - 1] * ((e[k + 1] - e[k]) / mu[k] + eps_k * x * e[k])
    p, q = cmath.exp(-1j * kl), cmath.exp(1j * kl)
    b = (e[1] - e[0] * p) / (q - p)
    a = e[0] - b
    return b / a


@pytest.mark.parametrize("scheme", [Scheme.Trapezoidal, Scheme.Leapfrog])
def test_drude_half_space_reflection(scheme):
    """drude_half_space_reflection (validation_dispersive.rs:253-265):
    |Gamma| against the continuum eps(w) within 3e-3 and against the exact
    discrete response of the scheme's eps_d within 5e-4 (trapezoidal:
    chi at the bilinear frequency; leapfrog: chi_d of
    Pole::leapfrog_susceptibility)."""
    eps_inf = 2.0
    pole = Pole.drude(TAU * 4e9 * math.sqrt(eps_inf), 1.0 / 0.2e-9)
    cells = IFACE + 100 + 24
    steps = 2400
    med = (eps_inf, 1.0, [pole], [], IFACE)
    et, pt = K.line("oracle", cells, DZ, 24, SRC, [PRB], medium=med, steps=steps, scheme=scheme)
    er, pr = K.line("oracle", cells, DZ, 24, SRC, [PRB], steps=steps)
    et.run(steps)
    er.run(steps)
    t, r = et.probe_values(pt[0]), er.probe_values(pr[0])
    dt = et.dt
    assert np.max(np.abs(t[-50:])) / np.max(np.abs(r)) < 1e-4
    freqs = [1e9 + 7e9 * q / 28 for q in range(29)]
    from ems_amd import post
    a = post.dft(dt, dt, r, freqs)
    b = post.dft(dt, dt, t - r, freqs)
    ec = ed = 0.0
    for q, f in enumerate(freqs):
        g = abs(b[q] / a[q])
        eps_c = eps_inf + pole.susceptibility(TAU * f)
        z = _csqrt(1.0 / eps_c)
        gc = abs((z - 1.0) / (z + 1.0))
        chi = pole.discrete_susceptibility(f, dt) if scheme == Scheme.Trapezoidal else \
            pole.leapfrog_susceptibility(f, dt)
        pe = [1.0 + 0j] * 4 + [eps_inf + chi] * 4
        gd = abs(_discrete_response(pe, [1.0 + 0j] * 8, DZ, dt, f))
        ec, ed = max(ec, abs(g - gc)), max(ed, abs(g - gd))
    assert ec < 3e-3, ec
    assert ed < 5e-4, ed


def test_fill_mask_matches_fill_cells_on_a_box():
    """fill_mask over a box of cells gives fill_cells' entries bit for bit,
    in its order, on a graded mesh; a sphere's edges get shares in (0, 1]."""
    lines = []
    for nn in (9, 8, 7):
        x, l = 0.0, []
        for q in range(nn + 1):
            l.append(x)
            x += 1e-3 * (1.0 + 0.3 * math.sin(1.3 * q))
        lines.append(l)
    g = E.Grid(lines)
    blk = [(2, 7), (1, 5), (3, 6)]
    poles = [Pole.lorentz(3e10, 2e10, 1e9), Pole.drude(5e10, 2e9)]
    a = Dispersive.new()
    a.fill_cells(g, blk, Field.Electric, poles)
    m = np.zeros(g.cells(), dtype=bool)
    m[2:7, 1:5, 3:6] = True
    b = Dispersive.new()
    b.fill_mask(g, m, Field.Electric, poles)
    assert len(a.entries) == len(b.entries) > 0
    for (fa, xa, na, pa), (fb, xb, nb, pb) in zip(a.entries, b.entries):
        assert (fa, xa, na, pa.kind) == (fb, xb, nb, pb.kind)
        assert np.array_equal(np.float64(pa.p), np.float64(pb.p))
    c = np.indices(g.cells()).astype(float) + 0.5
    sphere = ((c[0] - 4.5) ** 2 + (c[1] - 4.0) ** 2 + (c[2] - 3.5) ** 2) < 2.6 ** 2
    d = Dispersive.new()
    d.fill_mask(g, sphere, Field.Electric, [Pole.drude(5e10, 2e9)])
    shares = [e[3].p[0] / 5e10 for e in d.entries]
    assert len(shares) > 0 and min(shares) > 0.0 and max(shares) <= 1.0 + 1e-12


def test_dispersive_box_rejects_impossible_sphere_placement():
    from ems_amd import workloads as W
    with pytest.raises(ValueError):
        W.dispersive_box((40, 40, 32), pml=6, lorentz=2, drude=1, cylinder_cells=5.0,
                         radii=(2.5, 3.5))
