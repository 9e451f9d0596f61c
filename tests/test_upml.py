"""openEMS's PML_n layers (OpenEmsPml, src/engine/boundary/upml.rs) in the
oracle, pinned by openEMS itself: the grading against the coefficients of
openEMS's operator dump (the reference's test upml.rs:783-804), and the
whole layer update against openEMS's probe files of the fixture
``pml_face`` (tests/openems_pml.rs:154-158 accepts 2e-5; the oracle
reproduces the files to their 12 printed digits). A planted mistake in
each of the four update coefficients and a run without the layer miss the
files by orders of magnitude. The edge counts follow from the geometry of
the fixture (closed form below)."""
import numpy as np
import pytest

import ems_amd as E
from ems_amd.engine import EPS0
from ems_amd.upml import openems_pml_sigma

import openems_pml_case as PC


def test_profile_matches_openems_dump():
    """upml.rs:783-804: vv = (1 - a)/(1 + a) in f32 at depths 0.5..3.5 mm of
    a uniform 1 mm PML_4, from openEMS's --debug-operator dump (8 digits)."""
    dt = 1.925833270632893e-12
    for d, want in [(0.5, 0.8589056), (1.0, 0.7856962), (1.5, 0.68101895), (2.0, 0.53842795),
                    (2.5, 0.35649642), (3.0, 0.14281084), (3.5, -0.08507279)]:
        a = openems_pml_sigma(d * 1e-3, 4e-3, 4) * dt / (2.0 * EPS0)
        v = np.float32((1.0 - a) / (1.0 + a))
        assert abs(float(v) - want) < 1e-7, (d, v, want)
    assert openems_pml_sigma(0.0, 4e-3, 4) == 0.0
    assert openems_pml_sigma(-1e-3, 4e-3, 4) == 0.0


def test_sigma_integral_gives_round_trip_attenuation():
    """The profile's total integral: exp(-2 Z0 int sigma dx) = 1e-6 at
    normal incidence (upml.rs:58-67), by quadrature of the closed form."""
    from ems_amd.engine import Z0
    for cells in (4, 6, 8):
        w = cells * 1e-3
        x = np.linspace(0.0, w, 200001)
        s = np.array([openems_pml_sigma(v, w, cells) for v in x])
        integral = float(np.sum(0.5 * (s[1:] + s[:-1]) * np.diff(x)))
        assert abs(np.exp(-2.0 * Z0 * integral) / 1e-6 - 1.0) < 1e-4


def test_rejects_bad_faces():
    """upml.rs:806-813."""
    from oracle.bridge import OracleEngine
    g = E.Grid.uniform([5, 5, 5], [1e-3] * 3)
    dt = 0.5 * E.courant_limit(g)
    o = OracleEngine(g, dt, E.coefficients(g, None, [0] * 6, dt))
    for faces in ([(E.Face.XMin, 7)], [(E.Face.XMin, 0)], [(E.Face.XMin, 2), (E.Face.XMin, 2)]):
        with pytest.raises(E.EmsError):
            E.OpenEmsPml.new(o, faces)


def run_fixture(mutate=None, threads=2, pml=True, materials=True):
    from oracle.bridge import OracleEngine
    g, dt = PC.grid(), PC.dt()
    o = OracleEngine(g, dt, PC.lanes(g, dt), curl_order=1, threads=threads)
    if pml and not materials:
        op = E.OpenEmsPml.new(o, [(E.Face.XMax, PC.PML_CELLS)])
        o.add_operator(op)
        op_, probes = PC.build(o, g, pml=False)
    else:
        if mutate is not None:
            orig = E.OpenEmsPml._build

            def patched(engine, faces, eps_r, mu_r):
                p = orig(engine, faces, eps_r, mu_r)
                mutate(p)
                return p
            E.OpenEmsPml._build = staticmethod(patched)
        try:
            op, probes = PC.build(o, g, pml=pml)
        finally:
            if mutate is not None:
                E.OpenEmsPml._build = staticmethod(orig)
    o.run(PC.STEPS)
    cols = PC.columns(probes, o.probe_values)
    fields = [o.u_component(a).copy() for a in range(3)] + [o.i_component(a).copy() for a in range(3)]
    o.close()
    return op, cols, fields


@pytest.fixture(scope="module")
def fixture_run():
    return run_fixture()


@pytest.mark.parametrize("name", PC.FILES)
def test_probe_files_to_printed_digits(fixture_run, name):
    _, cols, _ = fixture_run
    ref = PC.read_rows(name)
    got = cols[name]
    assert got.shape == (ref.shape[0], ref.shape[1] - 1) == (167, got.shape[1])
    e = PC.rel_l2_worst(ref[:, 1:], got)
    assert e < 1e-10, f"{name}: relative L2 {e:.3e}"


def test_edge_counts_closed_form(fixture_run):
    """31 x 11 x 11 nodes, layer cells 22..29 along x (interface line 22):
    E_x in cells 22..29 away from the PEC z planes and the PMC y planes
    (8 x 9 x 9), E_y and E_z on node planes 23..29 (the face plane 30 is
    PEC): E_y 7 x 10 x 9 (PEC z planes), E_z 7 x 9 x 10 (PMC y planes).
    H_x on node planes 23..29 (the normal H on x_max is cleared) away from
    the cleared dual layers y = 0, 9: 7 x 8 x 10; H_y in cells 22..29 off
    the PMC planes y = 0, 10: 8 x 9 x 10; H_z in cells 22..29 off the dual
    layers y = 0, 9 and the cleared normal H on z_max: 8 x 8 x 10."""
    op, _, _ = fixture_run
    assert op.edge_counts() == (8 * 9 * 9 + 7 * 10 * 9 + 7 * 9 * 10,
                                7 * 8 * 10 + 8 * 9 * 10 + 8 * 8 * 10)
    assert op.state_counts() == op.edge_counts()
    assert op.layers.contains(True, 0, (22, 5, 5))      # E_x across cell 22
    assert not op.layers.contains(True, 1, (22, 5, 5))  # E_y on the interface
    assert op.layers.contains(True, 1, (23, 5, 5))


def test_thread_count_invariant(fixture_run):
    """The layer runs per x plane in the slab phases: bit-identical fields
    for any thread count (tests/openems_pml.rs:287-306)."""
    _, _, fields = fixture_run
    _, _, one = run_fixture(threads=1)
    _, _, five = run_fixture(threads=5)
    for a, b, c in zip(fields, one, five):
        assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("col", [0, 1, 2, 3])
def test_planted_coefficient_mistake_misses_the_files(col):
    """Each of the four update coefficients matters: scaling one of them by
    (1 + 1e-3) on every E and H edge moves some probe file by 0.13..0.31
    (measured), against 1e-12 for the right ones."""
    def mutate(p):
        for side in (p.e, p.h):
            side[2][:, col] *= np.float32(1.001)
    _, cols, _ = run_fixture(mutate=mutate)
    worst = max(PC.rel_l2_worst(PC.read_rows(n)[:, 1:], cols[n]) for n in PC.FILES)
    assert worst > 1e-2, worst


def test_without_the_layer_misses_the_files():
    _, cols, _ = run_fixture(pml=False)
    worst = max(PC.rel_l2_worst(PC.read_rows(n)[:, 1:], cols[n]) for n in PC.FILES)
    assert worst > 0.5


def test_capacitance_variant_differs_in_last_bits():
    """OpenEmsPml::new takes m from the engine coefficients instead of
    eps0' eps_r (upml.rs:578-590): the same layer up to rounding, which
    openEMS's files tell apart (3.7e-6 measured, against 1e-12 with the
    materials)."""
    _, cols, _ = run_fixture(materials=False)
    worst = max(PC.rel_l2_worst(PC.read_rows(n)[:, 1:], cols[n]) for n in PC.FILES)
    assert 1e-8 < worst < 1e-5, worst


def test_refusals():
    from oracle.bridge import OracleEngine
    g, dt = PC.grid(), PC.dt()
    o = OracleEngine(g, dt, PC.lanes(g, dt), curl_order=1)
    op = E.OpenEmsPml.with_materials(o, [(E.Face.XMax, 4)], 1.0, 1.0)
    o.add_operator(op)
    with pytest.raises(AssertionError):
        o.add_operator(E.OpenEmsPml.with_materials(o, [(E.Face.XMin, 4)], 1.0, 1.0))
    with pytest.raises(E.EmsError):
        E.OpenEmsPml.with_materials(o, [(E.Face.XMin, 4)], np.ones((3, 2, 2, 2)), 1.0)
