"""TF/SF plane wave on the CPU: the host set-up (dispersion-matched
spacing, stencils, auxiliary grid) and the oracle's restatement of the
corrections, pinned by the reference's leakage levels
(tests/planewave.rs:145-190) and unit tests (planewave.rs:1563-1603)."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import planewave as PW
from ems_amd._native import EmsError

import cases as K


def test_stencil_reproduces_cubics():
    f = lambda x: 0.3 - x + 0.2 * x * x - 0.05 * x * x * x  # noqa: E731
    v = [f(float(m)) for m in range(8)]
    for s in (1.0, 1.25, 2.5, 3.999, 4.0):
        m, w = PW.stencil(s)
        assert abs(PW.interp(v, m, w) - f(s)) < 1e-12


def test_axis_aligned_match_is_the_mesh_spacing():
    dt = 0.99 * 1e-3 / (E.C0 * math.sqrt(3.0))
    d, _ = PW.match_spacing(dt, E.C0, 10e9, [1.0, 0.0, 0.0], [1e-3] * 3)
    assert d == 1e-3
    s = math.sqrt(0.5)
    d, _ = PW.match_spacing(dt, E.C0, 10e9, [s, s, 0.0], [1e-3] * 3)
    assert E.C0 * dt < d < 1e-3


def test_aux_gamma_solves_the_dispersion_relation():
    dt = 0.99e-3 / (E.C0 * math.sqrt(3.0))
    er, kappa, f = 2.0, 0.13, 8e9 / math.sqrt(2.0)
    se = kappa * dt / (2.0 * E.EPS0 * er)
    r = E.C0 / math.sqrt(er) * dt / 1e-3
    g = PW.aux_gamma(2.0 * math.pi * f * dt, se, 0.0, r)
    om = 2.0 / dt * math.sin(math.pi * f * dt)
    c = math.cos(math.pi * f * dt)
    x = complex(om * om * E.MU0 * E.EPS0 * er * 1e-6, -om * E.MU0 * kappa * c * 1e-6)
    import cmath
    sh = (cmath.exp(g * 0.5) - cmath.exp(g * -0.5)) * 0.5
    assert abs(sh * sh + x * 0.25) < 1e-12 * abs(x)
    assert g.real > 0.0 and g.imag > 0.0


def test_normal_incidence_does_not_leak():
    """planewave.rs tests: < -80 dB at normal incidence."""
    l, _, _ = K.leakage("oracle", 36, 13, 23, [1.0, 0.0, 0.0], [0.0, 0.0, 1.0])
    assert l < -80.0, l
    l, _, _ = K.leakage("oracle", 36, 13, 23, [0.0, 0.0, -1.0], [0.0, 2.0, 0.0])
    assert l < -80.0, l


def test_oblique_incidence_leakage():
    s, c = math.sin(math.radians(30.0)), math.cos(math.radians(30.0))
    l30, _, _ = K.leakage("oracle", 36, 13, 23, [c, s, 0.0], [0.0, 0.0, 1.0])
    l45, _, _ = K.leakage("oracle", 36, 13, 23, [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0])
    l3d, _, _ = K.leakage("oracle", 36, 13, 23, [1.0, 1.0, 1.0], [1.0, -1.0, 0.0])
    assert l45 < -100.0 and l3d < -100.0, (l45, l3d)
    assert l30 < -60.0, l30


def test_dielectric_and_lossy_background_leakage():
    """dielectric_background / lossy_background leakage levels."""
    l0, _, _ = K.leakage("oracle", 36, 13, 23, [1.0, 0.0, 0.0], [0.0, 0.0, 1.0],
                         medium=PW.Medium(4.0, 1.0), f0=4e9)
    assert l0 < -100.0, l0
    l1, _, _ = K.leakage("oracle", 36, 13, 23, [1.0, 0.0, 0.0], [0.0, 0.0, 1.0],
                         medium=PW.Medium(2.0, 1.0, 0.13, 0.0), f0=8e9 / math.sqrt(2.0))
    assert l1 < -100.0, l1


def test_incident_field_at_the_reference_corner():
    """dielectric_background: E_inc(r_ref) = a(t) to 1e-3."""
    f0 = 4e9
    a = K.sine_pulse(f0, 0.5 * f0)
    errs = []

    def rec(e, pw, n):
        r = pw.reference_point()
        c = 0.5 * (e.grid.lines[1][18] + e.grid.lines[1][19])
        v = pw.incident_e([r[0], c, c])[2]
        errs.append(abs(v - a((n + 1) * e.dt)))
    K.leakage("oracle", 36, 13, 23, [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], medium=PW.Medium(4.0, 1.0),
              f0=f0, record=rec)
    assert max(errs) < 1e-3, max(errs)


def test_invalid_and_mismatched_specs_are_rejected():
    g = E.Grid.uniform([10] * 3, [1e-3] * 3)
    e = K.make("oracle", g)
    base = lambda: PW.PlaneWaveSpec.new([2] * 3, [8] * 3, [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])  # noqa: E731
    PW.PlaneWave.new(e, base(), lambda t: 0.0)
    bad = []
    s = base(); s.polarization = (1.0, 1.0, 0.0); bad.append(s)  # noqa: E702
    s = base(); s.stop = (8, 2, 8); bad.append(s)  # noqa: E702
    s = base(); s.stop = (8, 8, 11); bad.append(s)  # noqa: E702
    s = base(); s.direction = (0.0, 0.0, 0.0); bad.append(s)  # noqa: E702
    bad.append(base().with_background(-1.0, 1.0))
    bad.append(base().with_match_frequency(1e14))
    for s in bad:
        with pytest.raises(EmsError):
            PW.PlaneWave.new(e, s, lambda t: 0.0)
    g = E.Grid.uniform([12] * 3, [1e-3] * 3)
    e = K.make("oracle", g, E.CellMaterials(g, E.Material.dielectric(4.0)))
    spec = PW.PlaneWaveSpec.new([3] * 3, [8] * 3, [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    with pytest.raises(EmsError):
        PW.PlaneWave.new(e, spec, lambda t: 0.0)
    PW.PlaneWave.new(e, spec.with_background(4.0, 1.0), lambda t: 0.0)


def test_deterministic_across_threads():
    """planewave.rs tests: oblique incidence on a graded mesh, 1 vs 4 threads."""
    x = 0.0
    lines = []
    for i in range(25):
        lines.append(x)
        x += 1e-3 * (1.0 + 0.2 * math.sin(i * 0.5))
    g = E.Grid([lines] * 3)

    def run(t):
        e = K.make("oracle", g, threads=t)
        spec = PW.PlaneWaveSpec.new([6] * 3, [18] * 3, [1.0, 0.5, 0.2], [0.0, -0.4, 1.0])
        e.add_operator(PW.PlaneWave.new(e, spec, K.sine_pulse(10e9, 5e9)))
        e.run(150)
        return [e.u_component(a) for a in range(3)] + [e.i_component(a) for a in range(3)]
    one = run(1)
    assert np.any(one[2] != 0.0)
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(one, run(4)))
