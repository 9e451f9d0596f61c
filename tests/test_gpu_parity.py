"""Parity of the HIP engine with the oracle, through the C ABI: the same
grid, coefficient lanes and operators on both, compared bit for bit (fields,
probe traces, energies). Small sizes, so the oracle finishes in seconds."""
import math

import numpy as np
import pytest

import ems_amd as E
from cases import det_model, graded, make, pulse, small_cavity

import cases

pytestmark = pytest.mark.gpu

# Every parity test runs on the unfused schedule and on the fused one with
# each tile height and with many x chunks (chunk seams inside the box), the
# full-row and the packed-row kernels (packed: grids with nz <= 64; the
# others run the tiled kernel), the band kernel (bands of 4, 8, 16 rows) and
# the tiled kernel with its z-normal CPML slabs in passes of their own
# (flag 1) and with its tiles in two classes (flag 2). The unfused kernels
# load the next plane ahead and the E one corrects its CPML slabs inside by
# default; flag 4 has both load ahead with every slab in launches of its
# own, flag 8 neither load ahead, flag 16 both correct their slabs inside,
# flag 32 only H loads ahead and corrects inside, flag 64 only H loads ahead
# and every slab is outside.
SCHEDULES = [(0, 0, 0), (0, 0, 3), (0, 0, 0, 4), (0, 0, 3, 4), (0, 0, 0, 8), (0, 0, 0, 16),
             (0, 0, 5, 16), (0, 0, 0, 32), (0, 0, 0, 64), (0, 0, 4, 64),
             (1, 3, 0), (1, 7, 0), (1, 15, 0), (1, 7, 5), (2, 7, 3), (2, 15, 0),
             (3, 3, 0), (3, 7, 0), (3, 7, 4), (4, 7, 0), (4, 7, 3), (5, 7, 0),
             (6, 4, 0), (6, 8, 3), (6, 16, 0), (6, 16, 5), (1, 7, 0, 1), (2, 3, 5, 1),
             (1, 7, 0, 2), (2, 7, 3, 2), (2, 15, 4, 2), (2, 3, 5, 3)]
UNFUSED = {4: ("pf_cpml_out", 3), 8: ("plain", 0), 16: ("pf_cpml_in", 15),
           32: ("pf_h_cpml_in_h", 5), 64: ("pf_cpml_out_h_only_pf", 1)}


def schedule_id(s):
    flags = s[3] if len(s) > 3 else 0
    return ("mode%d_tj%d_xc%d" % s[:3] + ("_split" if flags & 1 else "")
            + ("_classes" if flags & 2 else "")
            + ("_" + UNFUSED[flags & 124][0] if flags & 124 else ""))


@pytest.fixture(autouse=True, params=SCHEDULES, ids=schedule_id)
def schedule(request):
    f, t, c, *flags = request.param
    fl = flags[0] if flags else 0
    old = dict(cases.SCHEDULE)
    cases.SCHEDULE.update(fused=f, tile_rows=t, x_chunks=c, split=fl & 1, classes=(fl >> 1) & 1,
                          unfused_pf=UNFUSED[fl & 124][1] if fl & 124 else 11)
    yield
    cases.SCHEDULE.update(old)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def same_fields(g, o):
    for a in range(3):
        for name, fg, fo in (("u", g.u_component(a), o.u_component(a)),
                             ("i", g.i_component(a), o.i_component(a))):
            bg, bo = bits(fg), bits(fo)
            if not np.array_equal(bg, bo):
                bad = np.argwhere(bg != bo)
                raise AssertionError(f"{name}[{a}] differs at {bad[:5].tolist()} "
                                     f"({bad.shape[0]} entries): {fg[tuple(bad[0])]} vs {fo[tuple(bad[0])]}")


def both(build):
    return build("gpu"), build("oracle")


def test_vacuum_box_soft_source_probe():
    """Config-1 shape at 21^3 cells: soft Gaussian Ez source, Ez probe."""
    def build(kind):
        g = E.Grid.uniform([20, 20, 20], [1e-3] * 3)
        e = make(kind, g, threads=2)
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(2e9, 1e9)).edge(E.Axis.Z, [10, 10, 10], 1.0))
        p = e.add_operator(E.Probe.voltage(E.Axis.Z, [4, 10, 10]))
        return e, p
    (g, pg), (o, po) = both(build)
    g.run(300)
    o.run(300)
    assert np.array_equal(bits(pg.values()), bits(o.probe_values(po)))
    same_fields(g, o)


def test_graded_lossy_pec_pmc_current_probe():
    """run_det of tests/engine.rs: graded mesh, lossy and magnetic media, PEC
    box, PMC faces, current probe every 3 steps (moved post_h hook)."""
    (g, pg), (o, po) = det_model("gpu"), det_model("oracle", threads=3)
    g.run(300)
    o.run(300)
    assert np.array_equal(bits(pg.values()), bits(o.probe_values(po)))
    same_fields(g, o)
    assert np.float64(g.energy()).tobytes() == np.float64(o.energy()).tobytes()


def test_energy_invariant_bitwise():
    (g, _), (o, _) = det_model("gpu"), det_model("oracle", threads=2)
    for n in (0, 1, 17, 50):
        g.run(n)
        o.run(n)
        wg, wo = g.step_with_energy(), o.step_with_energy()
        assert np.float64(wg).tobytes() == np.float64(wo).tobytes(), (n, wg, wo)
        assert np.float64(g.energy()).tobytes() == np.float64(o.energy()).tobytes()
    same_fields(g, o)


def test_openems_curl_order():
    def build(kind):
        g = E.Grid([graded(0.05, 11, 0.2), graded(0.04, 9, 0.1), graded(0.03, 8, 0.0)])
        e = make(kind, g, curl_order=1)
        f, _ = pulse(5e9, 4e9)
        e.add_operator(E.EdgeSource(f, horizon=400).edge(E.Axis.Y, [5, 4, 3], 1.0))
        p = e.add_operator(E.Probe.voltage(E.Axis.Y, [7, 4, 4]))
        return e, p
    (g, pg), (o, po) = both(build)
    g.run(250)
    o.run(250)
    assert np.array_equal(bits(pg.values()), bits(o.probe_values(po)))
    same_fields(g, o)


def test_cpml_all_faces():
    """8-cell CPML on all six faces of a graded box with a dielectric block."""
    def build(kind):
        g = E.Grid([graded(0.03, 30, 0.15), graded(0.028, 28, 0.1), graded(0.026, 26, 0.2)])
        m = E.CellMaterials.vacuum(g)
        m.fill_box([(11, 18), (10, 17), (9, 16)], E.Material.dielectric(4.3).with_kappa(0.02))
        e = make(kind, g, m, threads=3)
        e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(8)) for f in E.Face]))
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(5e9, 4e9)).edge(E.Axis.Z, [15, 14, 13], 1.0))
        ps = [e.add_operator(E.Probe.voltage(a, [x, 14, 13])) for a in range(3) for x in (3, 9, 21)]
        return e, ps
    (g, pg), (o, po) = both(build)
    g.run(400)
    o.run(400)
    for a, b in zip(pg, po):
        assert np.array_equal(bits(a.values()), bits(o.probe_values(b)))
    same_fields(g, o)
    wg, wo = g.step_with_energy(), o.step_with_energy()
    assert np.float64(wg).tobytes() == np.float64(wo).tobytes()


def test_cpml_constant_rule_and_cfs():
    def build(kind):
        g = E.Grid.uniform([24, 20, 22], [1e-3] * 3)
        e = make(kind, g)
        p = E.Cpml(cells=6, kappa_max=3.0, alpha_max=0.05, convolution=0)
        q = E.Cpml(cells=5, grading_order=2.0, sigma_factor=1.5)
        e.add_operator(E.CpmlBoundary.new(e, [(E.Face.XMin, p), (E.Face.ZMax, q), (E.Face.YMin, p),
                                              (E.Face.XMax, q)]))
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(8e9, 6e9)).edge(E.Axis.Y, [12, 10, 11], 1.0))
        pr = e.add_operator(E.Probe.voltage(E.Axis.Y, [3, 10, 11]))
        return e, pr
    (g, pg), (o, po) = both(build)
    g.run(300)
    o.run(300)
    assert np.array_equal(bits(pg.values()), bits(o.probe_values(po)))
    same_fields(g, o)


def test_sources_of_every_kind_and_probe_intervals():
    def build(kind):
        g = E.Grid.uniform([12, 10, 9], [2e-3] * 3)
        e = make(kind, g)
        f, _ = pulse(6e9, 4e9)
        e.add_operator(E.EdgeSource(f, horizon=120).kind(E.SourceKind.SoftH).edge(E.Axis.X, [5, 4, 4], 2e-3))
        e.add_operator(E.Probe.current(E.Axis.X, [6, 4, 4]).interval(2))
        e.add_operator(E.EdgeSource(f, horizon=90).kind(E.SourceKind.HardE).edge(E.Axis.Z, [3, 3, 3], 0.5))
        e.add_operator(E.EdgeSource(f, horizon=150).edge(E.Axis.Y, [8, 5, 4], -1.0).edge(E.Axis.Y, [8, 5, 4], 0.25))
        e.add_operator(E.EdgeSource(f, horizon=60).kind(E.SourceKind.HardH).edge(E.Axis.Z, [6, 6, 2], 1e-3))
        ps = [e.add_operator(E.Probe.voltage(E.Axis.Z, [6, 5, 4]).interval(5)),
              e.add_operator(E.Probe(E.ProbeQuantity.Voltage).entry(E.Axis.Y, [8, 5, 4], 1.0)
                             .entry(E.Axis.Y, [8, 5, 5], -0.5).interval(3)),
              e.add_operator(E.Probe.current(E.Axis.Y, [4, 4, 4]))]
        return e, ps
    (g, pg), (o, po) = both(build)
    for n in (7, 1, 60, 132):
        g.run(n)
        o.run(n)
    for a, b in zip(pg, po):
        assert np.array_equal(bits(a.values()), bits(o.probe_values(b)))
    assert np.array_equal(pg[0].times(), (np.arange(pg[0].values().size) * 5 + 1.0) * g.dt)
    same_fields(g, o)


def test_pooled_and_dense_lanes_agree():
    def build(kind, pool):
        gr = E.Grid.uniform([14, 11, 13], [1e-3] * 3)
        e = make(kind, gr, pool=pool)
        e.set_component(0, E.Axis.Z, np.random.default_rng(1).standard_normal(gr.nodes()).astype(np.float32))
        return e
    a, b, o = build("gpu", True), build("gpu", False), build("oracle", True)
    for e in (a, b, o):
        e.run(40)
    same_fields(a, o)
    same_fields(b, o)


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 2, 5), (2, 7, 3), (5, 1 + 256, 2), (4, 3, 300),
                               (2, 40, 70), (3, 2, 700)])
def test_odd_shapes(n):
    """Every kernel on odd shapes; the last three also run the energy's
    several term widths (rows of 71, 301, 701 nodes: 2, 8 and 8 terms per
    lane, the last in two rounds) and E chains over several rounds."""
    def build(kind):
        g = E.Grid.uniform([n[0], n[1], n[2]], [1e-3] * 3)
        faces = [E.FaceCondition.Pmc] * 6
        e = make(kind, g, faces=faces)
        rng = np.random.default_rng(7)
        for a in range(3):
            e.set_component(0, a, rng.standard_normal(g.nodes()).astype(np.float32))
            e.set_component(1, a, 1e-3 * rng.standard_normal(g.nodes()).astype(np.float32))
        return e
    g, o = build("gpu"), build("oracle")
    g.run(9)
    o.run(9)
    same_fields(g, o)
    assert np.float64(g.energy()).tobytes() == np.float64(o.energy()).tobytes()


def test_subnormal_fields_flush_like_the_reference():
    """Values near 2^-126 exercise the final-value flush (Paired) and the
    per-op flush (OpenEms), which hardware flush-to-zero would break."""
    for order in (0, 1):
        def build(kind):
            g = E.Grid.uniform([6, 6, 6], [1e-3] * 3)
            e = make(kind, g, faces=[E.FaceCondition.Pmc] * 6, curl_order=order)
            rng = np.random.default_rng(3)
            for a in range(3):
                v = (rng.standard_normal(g.nodes()) * 2.0 ** -120).astype(np.float32)
                e.set_component(0, a, v)
                e.set_component(1, a, (v * np.float32(2.0 ** -8)).astype(np.float32))
            return e
        g, o = build("gpu"), build("oracle")
        for _ in range(5):
            g.run(3)
            o.run(3)
            same_fields(g, o)


def test_run_until_energy_decay_matches_oracle_loop():
    """run_until on a lossy cavity (tests/engine.rs:319-376) stops where the
    restated loop stops on the oracle's energies."""
    kappa = 0.002
    g, t_end = small_cavity("gpu", lambda gr: E.CellMaterials(gr, E.Material().with_kappa(kappa)))
    lim = E.RunLimits(100_000, E.EndCriteria(10.0 ** (-30.0 / 10.0), check_interval=20), t_end)
    out = g.run_until(lim)
    assert out.reason == E.StopReason.EnergyDecay
    assert -31.0 < out.ratio_db < -30.0
    assert out.timestep == out.steps
    o, _ = small_cavity("oracle", lambda gr: E.CellMaterials(gr, E.Material().with_kappa(kappa)))
    w_max, steps = 0.0, 0
    while True:
        n = o.timestep()
        w = o.step_with_energy()
        steps += 1
        w_max = max(w_max, w)
        if n * o.dt > t_end and w / w_max < 10.0 ** (-3.0):
            break
        o.run(19)
        steps += 19
    assert steps == out.steps
    assert np.float64(w).tobytes() == np.float64(out.energy).tobytes()
