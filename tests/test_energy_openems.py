"""openEMS's energy measure (EnergyMeasure::OpenEms, src/engine/run.rs:66-75;
Engine::energy_openems, mod.rs:727-741) in the oracle, pinned by openEMS
itself: the fixture ``pml_end`` (a patch on a lossy substrate with PML_6
on five faces) stopped by openEMS's energy criterion after 1956 steps at
-50.5179 dB; the oracle, checking its measure (with the layers'
corrections, upml.rs:505-537) on openEMS's schedule (sim/mod.rs:915-929),
stops at the same step at the same level to the printed digits, and its
probes match the files to their printed digits (tests/openems_pml.rs:
228-249 accepts 0.02 dB and 2e-5). Without the layers' corrections the
level is off. The measure's masks and weights are pinned in closed form on
states where only E or only H is non-zero, and the H sources' share by
the look-ahead of a stepped copy."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd.engine import EPS0, MU0
from oracle.bridge import OracleEngine

import openems_pml_case as PC
import openems_pml_end_case as PE


@pytest.fixture(scope="module")
def pml_end_run():
    g, dt = PE.grid(), PE.dt()
    o = OracleEngine(g, dt, PE.lanes(g, dt), curl_order=1, threads=4)
    op, probes = PE.build(o, g)
    steps, db = PE.drive(o, o.energy_openems, dt)
    no_layers = o.energy_openems(layers=False)
    with_layers = o.energy_openems()
    o.run(1)  # the files' last H row is the look-ahead of the state reached
    _, _, pint = PE.limits(dt)
    cols = {}
    for name in PE.FILES:
        rows = PE.read_rows(name).shape[0]
        cols[name] = np.stack([PE.sampled(name, o.probe_values(p), pint, rows)
                               for p in probes[name]], axis=1)
    o.close()
    return steps, db, cols, no_layers, with_layers


def test_openems_limits_of_the_fixture():
    """Nyquist interval floor(1/(2 (f0 + fc) dt)) = 12, the extra check
    after half the Gaussian's floor(T/dt) + 1 = 149 samples, probes every
    12 // 4 = 3 steps; the files' rows are 3 dt apart."""
    dt = PE.dt()
    assert PE.limits(dt) == (12, 74, 3)
    t = PE.read_rows("ut")[:, 0]
    assert abs(t[1] / dt - 3.0) < 1e-6


def test_stops_where_openems_stops(pml_end_run):
    steps, db, _, _, _ = pml_end_run
    assert steps == PE.STOP_STEPS
    assert abs(db - PE.STOP_DB) < 5e-5, db


@pytest.mark.parametrize("name", PE.FILES)
def test_probe_files_to_printed_digits(pml_end_run, name):
    _, _, cols, _, _ = pml_end_run
    e = PC.rel_l2_worst(PE.read_rows(name)[:, 1:], cols[name])
    assert e < 1e-10, f"{name}: {e:.3e}"


def test_layer_corrections_matter(pml_end_run):
    """At the stop the layers hold a share of the measure that the plain
    sum over the fields misreads (the flux increment instead of i')."""
    _, _, _, no_layers, with_layers = pml_end_run
    assert abs(no_layers / with_layers - 1.0) > 1e-3


def small_engine(step_da_db=None, curl=0):
    g = E.Grid([np.array([0.0, 1.0, 2.5, 3.1, 4.0, 5.2]) * 1e-3,
                np.array([0.0, 0.9, 2.0, 3.0, 4.4]) * 1e-3,
                np.array([0.0, 1.1, 2.0, 3.3, 4.0, 5.0, 6.1]) * 1e-3])
    dt = 0.5 * E.courant_limit(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    if step_da_db is not None:
        lanes[6:9] = step_da_db[0]
        lanes[9:12] = step_da_db[1]
    return g, OracleEngine(g, dt, lanes, curl_order=curl, threads=2)


@pytest.mark.parametrize("curl", [0, 1])
def test_e_part_closed_form(curl):
    """i = 0 and db = 0 (the look-ahead keeps i = 0): eps0 sum u^2 over the
    E edges off the last node plane of every axis."""
    g, o = small_engine((1.0, 0.0), curl)
    n = g.nodes()
    rng = np.random.default_rng(3)
    want = 0.0
    for a in range(3):
        u = rng.standard_normal(n).astype(np.float32)
        o.set_component(0, a, u)
        want += float(np.sum(u[:-1, :-1, :-1].astype(np.float64) ** 2))
    assert abs(o.energy_openems() / (EPS0 * want) - 1.0) < 1e-13
    o.close()


@pytest.mark.parametrize("curl", [0, 1])
def test_h_part_closed_form(curl):
    """u = 0, da = 1, db = 0 (the look-ahead is i itself): mu0 sum w i^2
    over the dual edges the H update visits (H_x: j < ny-1, k < nz-1;
    H_y: i < nx-1, k < nz-1; H_z: i < nx-1, j < ny-1), the ones on the outer
    planes normal to their own axis weighted 4."""
    g, o = small_engine((1.0, 0.0), curl)
    n = g.nodes()
    rng = np.random.default_rng(5)
    want = 0.0
    for a in range(3):
        i = rng.standard_normal(n).astype(np.float32)
        o.set_component(1, a, i)
        sl = [slice(None)] * 3
        for b in range(3):
            if b != a:
                sl[b] = slice(0, n[b] - 1)
        w = np.ones(n)
        idx = [slice(None)] * 3
        idx[a] = 0
        w[tuple(idx)] = 4.0
        idx[a] = n[a] - 1
        w[tuple(idx)] = 4.0
        want += float(np.sum((w * i.astype(np.float64) ** 2)[tuple(sl)]))
    assert abs(o.energy_openems() / (MU0 * want) - 1.0) < 1e-13
    o.close()


@pytest.mark.parametrize("layers", [False, True], ids=["plain", "in_layer"])
def test_h_source_share_is_the_look_ahead(layers):
    """A soft and a hard H source: the measure with them minus the plain
    field part is mu0 sum w ((k + s)^2 - k^2), k the look-ahead (the current
    a stepped copy without the sources reaches; on a layer edge the layer's
    completed value), k + s what the sources make of it (the copy with
    them)."""
    def make(with_src):
        g, o = small_engine()
        if layers:  # two cells on x_max: H_x at i = 4 is a layer edge
            o.add_operator(E.OpenEmsPml.new(o, [(E.Face.XMax, 2)]))
        rng = np.random.default_rng(9)
        for a in range(3):
            o.set_component(0, a, (1e-3 * rng.standard_normal(g.nodes())).astype(np.float32))
            o.set_component(1, a, (1e-6 * rng.standard_normal(g.nodes())).astype(np.float32))
        o.run(3)  # some layer state
        if with_src:
            def amp(t):
                return 2e-6 * (1.0 + 1e11 * t)
            o.add_operator(E.EdgeSource(amp, horizon=50).kind(E.SourceKind.SoftH)
                           .edge(E.Axis.Y, [2, 0, 3], 1.5).edge(E.Axis.Z, [1, 2, 0], 0.5)
                           .edge(E.Axis.X, [4, 1, 2], 0.7))
            o.add_operator(E.EdgeSource(amp, horizon=50).kind(E.SourceKind.HardH)
                           .edge(E.Axis.X, [3, 1, 2], 2.0))
        return g, o
    g, a = make(True)
    _, b = make(False)
    with_src, plain = a.energy_openems(), b.energy_openems()
    edges = [(1, (2, 0, 3)), (2, (1, 2, 0)), (0, (3, 1, 2)), (0, (4, 1, 2))]
    a.run(1)
    b.run(1)
    n = g.nodes()
    share = 0.0
    for c, p in sorted(edges):
        w = 4.0 if p[c] in (0, n[c] - 1) else 1.0
        ks = float(a.i_component(c)[p])
        k = float(b.i_component(c)[p])
        share += w * (ks * ks - k * k)
    assert share != 0.0
    assert abs((with_src - plain) / (MU0 * share) - 1.0) < 1e-9
    a.close()
    b.close()
