"""Pins the oracle's kernel restatement against the closed forms of the
reference's own engine tests (tests/engine.rs): PEC cavity resonances on
uniform, graded and anisotropic-dielectric meshes, conservation of the
leapfrog energy invariant, the lossy decay rate, PEC/padding invariants and
bitwise determinism over thread counts. A dropped curl term, a wrong sign or
a transposed neighbour moves the resonances or breaks the invariant."""
import math

import numpy as np
import pytest

import ems_amd as E
from cases import (DIMS, cavity, check_modes, det_model, graded, make, pulse, small_cavity,
                   staggered_energy)


def test_pec_cavity_uniform():
    g = E.Grid.uniform([20, 9, 14], [0.005] * 3)
    e, p, t_end = cavity("oracle", g, None, 3000, 1.0)
    v = e.probe_values(p)
    t = (np.arange(v.size) + 1.0) * e.dt
    check_modes(t, v, t_end, 1.0, 0.005)


def test_pec_cavity_graded():
    g = E.Grid([graded(DIMS[0], 22, 0.3), graded(DIMS[1], 10, 0.2), graded(DIMS[2], 16, 0.3)])
    e, p, t_end = cavity("oracle", g, None, 4000, 1.0)
    v = e.probe_values(p)
    t = (np.arange(v.size) + 1.0) * e.dt
    check_modes(t, v, t_end, 1.0, 0.005)


def test_dielectric_cavity_scales():
    g = E.Grid.uniform([20, 9, 14], [0.005] * 3)
    m = E.CellMaterials(g, E.Material(eps_r=(4.0, 2.25, 9.0)))
    e, p, t_end = cavity("oracle", g, m, 4500, 1.0 / 1.5)
    v = e.probe_values(p)
    t = (np.arange(v.size) + 1.0) * e.dt
    check_modes(t, v, t_end, 2.25, 0.005)


def _drift(e, t_end, steps, interval):
    e.run(math.ceil(t_end / e.dt) + 1)
    w0, s0 = e.energy(), staggered_energy(e, e.dt)
    assert w0 > 0.0
    dev = sdev = 0.0
    for _ in range(steps // interval):
        w_ro = e.energy()
        w = e.step_with_energy()
        assert np.float64(w).tobytes() == np.float64(w_ro).tobytes()
        dev = max(dev, abs(w / w0 - 1.0))
        sdev = max(sdev, abs(staggered_energy(e, e.dt) / s0 - 1.0))
        e.run(interval - 1)
    return dev, sdev


def test_lossless_energy_conserved_uniform():
    e, t_end = small_cavity("oracle", threads=2)
    dev, sdev = _drift(e, t_end, 10_000, 10)
    assert dev < 1e-4, dev
    assert sdev > 100.0 * dev


def test_lossless_energy_conserved_graded_dielectric():
    g = E.Grid([graded(0.1, 15, 0.3), graded(0.07, 11, 0.2), graded(0.05, 9, 0.25)])
    m = E.CellMaterials.vacuum(g)
    m.fill_box([(3, 9), (2, 8), (1, 6)], E.Material.dielectric(4.0))
    m.fill_box([(9, 14), (0, 11), (4, 9)], E.Material(eps_r=(2.0, 3.0, 1.5), mu_r=(1.5,) * 3))
    m.fill_pec_box([(11, 12), (3, 8), (2, 7)])
    faces = [E.FaceCondition.Pec] * 6
    faces[E.Face.YMax] = E.FaceCondition.Pmc
    e = make("oracle", g, m, faces=faces, threads=3)
    f, t_end = pulse(3e9, 2e9)
    e.add_operator(E.EdgeSource(f, horizon=20000).edge(E.Axis.Z, [5, 4, 3], 1.0)
                   .edge(E.Axis.Y, [10, 5, 6], 0.7))
    dev, _ = _drift(e, t_end, 8000, 50)
    assert dev < 1e-4, dev


def test_lossy_energy_decays():
    kappa = 0.002
    e, t_end = small_cavity("oracle", lambda g: E.CellMaterials(g, E.Material().with_kappa(kappa)))
    e.run(math.ceil(t_end / e.dt) + 1)
    prev = e.energy()
    w1, t1 = prev, e.timestep() * e.dt
    for _ in range(500):
        w = e.step_with_energy()
        assert w <= prev
        prev = w

    def avg():
        s = 0.0
        for _ in range(100):
            s += e.step_with_energy()
        return s / 100.0, e.timestep() * e.dt - 50.5 * e.dt

    wa, ta = avg()
    e.run(400)
    wb, tb = avg()
    assert wb < wa < w1 and t1 < ta
    rate = math.log(wa / wb) / (tb - ta)
    want = kappa / E.EPS0
    assert abs(rate / want - 1.0) < 0.02, (rate, want)


def test_pec_walls_and_padding_stay_zero():
    e, _ = small_cavity("oracle", threads=3)
    e.run(200)
    nx, ny, nz = e.grid.nodes()
    uy, uz, ux, ix = e.u_component(1), e.u_component(2), e.u_component(0), e.i_component(0)
    assert np.all(uy[0] == 0) and np.all(uz[nx - 1] == 0) and np.all(ux[nx - 1] == 0)
    assert np.all(ix[0] == 0)
    assert e.energy() > 0.0


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


def test_deterministic_across_threads():
    ref = None
    for t in (1, 2, 3, 4):
        e, p = det_model("oracle", threads=t)
        e.run(300)
        arrays = [e.u_component(a) for a in range(3)] + [e.i_component(a) for a in range(3)]
        out = ([_bits(a) for a in arrays], _bits(e.probe_values(p)), e.energy())
        assert out[2] > 0.0
        if ref is None:
            ref = out
            continue
        for x, y in zip(ref[0], out[0]):
            assert np.array_equal(x, y)
        assert np.array_equal(ref[1], out[1])
        assert np.float64(ref[2]).tobytes() == np.float64(out[2]).tobytes()


def test_energy_invariant_in_step_equals_read_only_across_threads():
    for t in (1, 4):
        e, _ = det_model("oracle", threads=t)
        e.run(37)
        w_ro = e.energy()
        w = e.step_with_energy()
        assert np.float64(w).tobytes() == np.float64(w_ro).tobytes()
