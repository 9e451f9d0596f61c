"""Models of the reference's own engine tests (tests/engine.rs), built for any
engine (device or oracle) from the same grid, coefficients and operators."""
from __future__ import annotations

import math

import numpy as np

import ems_amd as E
from ems_amd import engine as EE

DIMS = (0.10, 0.045, 0.07)


def pulse(f0: float, bw: float):
    """tests/engine.rs:15-23: Gaussian-modulated pulse and its end 2*t0."""
    tau = 1.0 / (math.pi * bw)
    t0 = 4.0 * tau

    def f(t):
        x = (t - t0) / tau
        return math.exp(-x * x) * math.cos(2.0 * math.pi * f0 * (t - t0))

    return f, 2.0 * t0


def graded(length: float, cells: int, alpha: float):
    """tests/engine.rs:72-81: sinusoidally graded lines."""
    return [length * (q / cells - alpha * math.sin(2.0 * math.pi * q / cells) / (2.0 * math.pi))
            for q in range(cells + 1)]


# Schedule of the device engines the tests build: (fused, tile_rows, x_chunks)
# and whether the tiled step corrects its z-normal CPML slabs outside itself,
# splits its tiles in classes, and which unfused kernels load ahead.
SCHEDULE = {"fused": 1, "tile_rows": 0, "x_chunks": 0, "split": 0, "classes": 0,
            "unfused_pf": 11}


def make(kind: str, grid, mats=None, faces=None, ts=None, threads=1, curl_order=0, pool=True):
    faces = faces or [E.FaceCondition.Pec] * 6
    ts = ts or E.TimeStep.courant(0.99)
    dt = ts.resolve(grid)
    lanes = E.coefficients(grid, mats, faces, dt)
    if kind == "gpu":
        e = E.Engine(grid, dt, lanes, curl_order=curl_order, pool=pool)
        e.set_schedule(SCHEDULE["fused"], SCHEDULE["tile_rows"], SCHEDULE["x_chunks"])
        e.set_cpml_split(bool(SCHEDULE.get("split", 0)))
        e.set_tile_classes(bool(SCHEDULE.get("classes", 0)))
        e.set_unfused_prefetch(SCHEDULE.get("unfused_pf", 11))
        return e
    from oracle.bridge import OracleEngine
    return OracleEngine(grid, dt, lanes, curl_order=curl_order, threads=threads, pool=pool)


def nearest(grid, axis, x):
    l = grid.lines[axis]
    return int(np.argmin(np.abs(l - x)))


def cavity(kind, grid, mats, steps, scale, threads=4, horizon=None):
    """run_cavity (tests/engine.rs:92-128)."""
    src = [nearest(grid, 0, 0.3 * DIMS[0]), nearest(grid, 1, 0.3 * DIMS[1]),
           nearest(grid, 2, 0.4 * DIMS[2])]
    prb = [nearest(grid, 0, 0.7 * DIMS[0]), nearest(grid, 1, 0.5 * DIMS[1]),
           nearest(grid, 2, 0.6 * DIMS[2])]
    e = make(kind, grid, mats, threads=threads)
    f, t_end = pulse(3.2e9 * scale, 2.5e9 * scale)
    e.add_operator(E.EdgeSource(f, horizon=horizon or steps).edge(E.Axis.Y, src, 1.0))
    p = e.add_operator(E.Probe.voltage(E.Axis.Y, prb))
    e.run(steps)
    return e, p, t_end


def small_cavity(kind, mats_fn=None, threads=2, horizon=20000):
    """small_cavity (tests/engine.rs:196-211)."""
    g = E.Grid.uniform([10, 8, 6], [0.01, 0.009, 0.011])
    m = mats_fn(g) if mats_fn else None
    e = make(kind, g, m, threads=threads)
    f, t_end = pulse(3e9, 1e9)
    e.add_operator(E.EdgeSource(f, horizon=horizon).edge(E.Axis.Z, [3, 2, 2], 1.0)
                   .edge(E.Axis.X, [6, 5, 3], -0.5))
    return e, t_end


def det_model(kind, threads=1, faces=None, steps=300):
    """run_det (tests/engine.rs:497-531): graded, lossy dielectric, magnetic
    loss, PEC box, PMC faces, soft source, current probe every 3 steps."""
    g = E.Grid([graded(0.1, 13, 0.2), graded(0.08, 9, 0.1), graded(0.06, 7, 0.0)])
    m = E.CellMaterials.vacuum(g)
    m.fill_box([(2, 6), (1, 5), (2, 6)], E.Material.dielectric(3.0).with_kappa(0.02))
    m.fill_box([(7, 12), (0, 9), (0, 3)], E.Material.vacuum().with_mu_r(2.0).with_sigma_m(50.0))
    m.fill_pec_box([(9, 10), (2, 7), (1, 6)])
    if faces is None:
        faces = [E.FaceCondition.Pec] * 6
        faces[E.Face.XMax] = E.FaceCondition.Pmc
        faces[E.Face.ZMin] = E.FaceCondition.Pmc
    e = make(kind, g, m, faces=faces, ts=E.TimeStep.courant(0.95), threads=threads)
    f, _ = pulse(5e9, 4e9)
    e.add_operator(E.EdgeSource(f, horizon=steps + 10).edge(E.Axis.Z, [4, 4, 3], 1.0))
    p = e.add_operator(E.Probe.current(E.Axis.X, [5, 3, 3]).interval(3))
    return e, p


def spectrum(times, values, t_start, f):
    """tests/engine.rs:26-39: |DFT| of the Hann-windowed tail."""
    first = int(np.argmax(times >= t_start))
    t, v = times[first:], values[first:]
    n = float(t.size)
    q = np.arange(t.size)
    w = 0.5 - 0.5 * np.cos(2.0 * math.pi * q / (n - 1.0))
    ph = -2.0 * math.pi * f * t
    return math.hypot(float(np.sum(w * v * np.cos(ph))), float(np.sum(w * v * np.sin(ph))))


def peak(times, values, t_start, lo, hi):
    """tests/engine.rs:43-65: coarse scan then golden-section search."""
    s = lambda f: spectrum(times, values, t_start, f)
    n = 200
    step = (hi - lo) / n
    vals = [s(lo + q * step) for q in range(n + 1)]
    best = int(np.argmax(vals))
    a, b = lo + (best - 1.0) * step, lo + (best + 1.0) * step
    g = (math.sqrt(5.0) - 1.0) / 2.0
    for _ in range(60):
        c = b - g * (b - a)
        d = a + g * (b - a)
        if s(c) > s(d):
            b = d
        else:
            a = c
    f = 0.5 * (a + b)
    return f, s(f)


def cavity_freq(m, n, p, dims, eps_r):
    return E.C0 / (2.0 * math.sqrt(eps_r)) * math.sqrt((m / dims[0]) ** 2 + (n / dims[1]) ** 2
                                                       + (p / dims[2]) ** 2)


def check_modes(times, values, t_start, eps_r, tol):
    """check_modes (tests/engine.rs:130-154). Returns the relative errors."""
    errs = []
    for (m, n, p) in [(1.0, 0.0, 1.0), (2.0, 0.0, 1.0)]:
        fa = cavity_freq(m, n, p, DIMS, eps_r)
        f, mag = peak(times, values, t_start, 0.95 * fa, 1.05 * fa)
        err = (f - fa) / fa
        assert abs(err) < tol, f"mode ({m},{n},{p}) off by {err*100:.3f}%"
        assert 0.96 * fa < f < 1.04 * fa
        off = spectrum(times, values, t_start, 0.9 * fa)
        assert mag > 20.0 * off, f"no clear peak for ({m},{n},{p})"
        errs.append(err)
    return errs


def staggered_energy(e, dt):
    """Engine::energy_staggered (mod.rs:786-806), in numpy (diagnostic)."""
    lanes = e._dense_lanes
    w = 0.0
    for field, base in ((0, 0), (1, 6)):
        for a in range(3):
            x = (e.u_component(a) if field == 0 else e.i_component(a)).astype(np.float64)
            ca = lanes[base + a].astype(np.float64)
            cb = lanes[base + 3 + a].astype(np.float64)
            with np.errstate(divide="ignore", invalid="ignore"):
                c = np.where(cb == 0.0, 0.0, dt * (1.0 + ca) / (2.0 * cb))
            w += float(np.sum(0.5 * c * x * x))
    return w


# ---------------------------------------------------------------- lumped --
# Models of the reference's lumped-element tests (tests/lumped.rs), for any
# engine kind.
from ems_amd.lumped import LumpedBox, LumpedElement, LumpedPort, Topology  # noqa: E402

LOOP = {"port_x": 8, "load_x": 12, "y": 10, "z0": 9, "z1": 11}
FREQS = [50e6, 100e6, 200e6, 500e6, 1e9, 2e9]
Z0_PORT = 50.0


def ramp(tr):
    """tests/lumped.rs:147-159: raised-cosine 0 -> 1 over tr."""
    def f(t):
        if t >= tr:
            return 1.0
        if t <= 0.0:
            return 0.0
        return 0.5 * (1.0 - math.cos(math.pi * t / tr))
    return f


def loop_column(x):
    return LumpedBox(E.Axis.Z, [x, LOOP["y"], LOOP["z0"]], [x, LOOP["y"], LOOP["z1"]])


def loop_materials(g):
    """The two PEC strips of the loop (tests/lumped.rs:59-65)."""
    m = E.CellMaterials.vacuum(g)
    for z in (LOOP["z0"], LOOP["z1"]):
        m.fill_pec_box([(LOOP["port_x"], LOOP["load_x"] + 1), (LOOP["y"], LOOP["y"] + 1), (z, z + 1)])
    return m


def run_loop(kind, load, steps=4000, threads=2, freqs=FREQS):
    """run_loop (tests/lumped.rs:105-132): load is "short", "open" or a
    function of the load column returning a LumpedElement."""
    g = E.Grid.uniform([20, 20, 20], [1e-3] * 3)
    m = loop_materials(g)
    if load == "short":
        m.fill_pec_box([(LOOP["load_x"], LOOP["load_x"] + 1), (LOOP["y"], LOOP["y"] + 1),
                        (LOOP["z0"], LOOP["z1"] + 1)])
    e = make(kind, g, m, threads=threads)
    p = e.add_operator(LumpedPort(loop_column(LOOP["port_x"]), Z0_PORT)
                       .excite_waveform(E.Waveform.gaussian(0.0, 3e9), 1.0))
    if callable(load):
        e.add_operator(load(loop_column(LOOP["load_x"])))
    e.run(steps)
    return e, p


def two_port(kind, excite, threads=2, steps=5000):
    """two_port (tests/lumped.rs:452-496)."""
    g = E.Grid.uniform([20, 20, 20], [1e-3] * 3)
    y, z0, z1 = 10, 9, 11
    m = E.CellMaterials.vacuum(g)
    for z in (z0, z1):
        m.fill_pec_box([(5, 16), (y, y + 1), (z, z + 1)])
    col = lambda x: LumpedBox(E.Axis.Z, [x, y, z0], [x, y, z1])  # noqa: E731
    e = make(kind, g, m, threads=threads)
    w = E.Waveform.gaussian(0.0, 3e9)
    p1 = LumpedPort(col(5), 50.0)
    p2 = LumpedPort(col(15), 75.0)
    if excite == 1:
        p1.excite_waveform(w, 1.0)
    else:
        p2.excite_waveform(w, 1.0)
    e.add_operator(p1)
    e.add_operator(p2)
    e.add_operator(LumpedElement.new(col(7)).c(0.5e-12))
    e.add_operator(LumpedElement.new(col(10)).r(30.0).c(2e-12).topology(Topology.Series))
    e.run(steps)
    return e, p1, p2


def lc_cavity(kind, threads=1):
    """lc_loaded_cavity_conserves_total_energy (tests/lumped.rs:585-640):
    parallel L||C and series L+C in a 12 mm PEC cube, the parallel element
    charged to 0.25 V per edge."""
    g = E.Grid.uniform([12, 12, 12], [1e-3] * 3)
    e = make(kind, g, threads=threads)
    par = LumpedBox(E.Axis.Z, [4, 6, 4], [4, 6, 8])
    ser = LumpedBox(E.Axis.X, [6, 3, 7], [9, 3, 7])
    hp = e.add_operator(LumpedElement.new(par).l(2e-9).c(0.2e-12))
    hs = e.add_operator(LumpedElement.new(ser).l(3e-9).c(0.1e-12).topology(Topology.Series))
    u = e.u_component(E.Axis.Z)
    for p in par.edges():
        u[p] = 0.25
    e.set_component(0, E.Axis.Z, u)
    return e, hp, hs


# ------------------------------------------------------------ dispersive --
from ems_amd.dispersive import Dispersive, Field, Pole, Scheme  # noqa: E402

TAU2PI = 2.0 * math.pi


def disp_cavity(kind, threads=1, lossless=True, magnetic=False, scheme=None):
    """cavity (tests/validation_dispersive.rs:377-425): 10x8x6 mm PEC box,
    an eps_inf = 1.5 block with Lorentz + Drude + Debye poles (and a
    magnetic Lorentz pole), four edges charged, no source."""
    g = E.Grid.uniform([10, 8, 6], [1e-3] * 3)
    m = E.CellMaterials.vacuum(g)
    blk = [(2, 8), (1, 6), (1, 5)]
    m.fill_box(blk, E.Material.dielectric(1.5))
    e = make(kind, g, m, threads=threads)
    gam = 0.0 if lossless else 3e9
    w = TAU2PI * 20e9
    d = Dispersive.new()
    if scheme is not None:
        d.with_scheme(scheme)
    d.fill_cells(g, blk, Field.Electric, [Pole.lorentz(w, w, gam), Pole.drude(0.7 * w, gam),
                                          Pole.debye(2.0, 0.0 if lossless else 5e-12)])
    if magnetic:
        d.fill_cells(g, blk, Field.Magnetic, [Pole.lorentz(0.5 * w, 1.5 * w, gam)])
    h = e.add_operator(d)
    for q, n in enumerate([[3, 3, 2], [5, 2, 3], [1, 4, 2], [6, 5, 3]]):
        u = e.u_component(q % 3)
        u[tuple(n)] = 1.0 + q
        e.set_component(0, q % 3, u)
    return e, h


def sine_pulse(f0, bw):
    """tests/common/mod.rs:24-31."""
    tau = 1.0 / (math.pi * bw)
    t0 = 4.0 * tau

    def f(t):
        x = (t - t0) / tau
        return math.exp(-x * x) * math.sin(2.0 * math.pi * f0 * (t - t0))
    return f


def line(kind, cells, dz, pml_hi, src, probes, medium=None, steps=2400, f0=4.5e9, bw=3.5e9,
         scheme=None):
    """Line::build (tests/common/mod.rs:190-300): a 1-cell-wide TEM line
    along z, x-polarised (PMC y faces, PEC x faces), CPML 8 cells at z_min
    and `pml_hi` at z_max, soft source on the two x edges at z node `src`,
    E probes; `medium` = (eps_inf, mu_inf, electric poles, magnetic poles,
    first cell) fills z cells first.. with eps_inf/mu_inf and the poles."""
    z = [q * dz for q in range(cells + 1)]
    g = E.Grid([[0.0, 5e-3], [0.0, 5e-3], z])
    faces = [E.FaceCondition.Pec] * 6
    faces[E.Face.YMin] = faces[E.Face.YMax] = E.FaceCondition.Pmc
    m = E.CellMaterials.vacuum(g)
    if medium is not None:
        eps_inf, mu_inf, ep, mp, c0 = medium
        m.fill_box([(0, 1), (0, 1), (c0, cells)], E.Material.dielectric(eps_inf).with_mu_r(mu_inf))
    e = make(kind, g, m, faces=faces, threads=1)
    e.add_operator(E.CpmlBoundary.new(e, [(E.Face.ZMin, E.Cpml.with_cells(8)),
                                          (E.Face.ZMax, E.Cpml.with_cells(pml_hi))]))
    if medium is not None:
        d = Dispersive.new()
        if scheme is not None:
            d.with_scheme(scheme)
        d.fill_cells(g, [(0, 1), (0, 1), (c0, cells)], Field.Electric, ep)
        d.fill_cells(g, [(0, 1), (0, 1), (c0, cells)], Field.Magnetic, mp)
        e.add_operator(d)
    s = E.EdgeSource(sine_pulse(f0, bw), horizon=steps)
    for p in ([0, 0, src], [0, 1, src]):
        s.edge(E.Axis.X, p, 1.0)
    e.add_operator(s)
    ps = [e.add_operator(E.Probe.voltage(E.Axis.X, [0, 0, k])) for k in probes]
    return e, ps


# ------------------------------------------------------------- plane wave --
from ems_amd import planewave as PWM  # noqa: E402


def leakage(kind, n, lo, hi, d, pol, medium=None, f0=8e9, grid=None, record=None):
    """leakage_in (tests/planewave.rs:48-143): cube of n cells with an
    8-cell CPML, TF box lo..=hi; the worst scattered-field E (between the
    box and the CPML, every second step) over the incident peak, in dB."""
    medium = medium or PWM.Medium()
    g = grid or E.Grid.uniform([n] * 3, [1e-3] * 3)
    mat = E.Material.dielectric(medium.eps_r).with_mu_r(medium.mu_r) \
        .with_kappa(medium.kappa).with_sigma_m(medium.sigma_m)
    e = make(kind, g, E.CellMaterials(g, mat), threads=1)
    e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(8)) for f in E.Face]))
    bw = 0.5 * f0
    spec = PWM.PlaneWaveSpec.new([lo] * 3, [hi] * 3, d, pol).with_medium(medium) \
        .with_match_frequency(f0)
    pw = e.add_operator(PWM.PlaneWave.new(e, spec, sine_pulse(f0, bw)))
    steps = int((8.0 / (math.pi * bw) + 3.0 * n * 1e-3 / medium.velocity()) / e.dt)
    nn = g.nodes()
    p0, p1 = 9, n - 9
    lens = []
    masks = []
    for a in range(3):
        m = np.zeros(nn, dtype=bool)
        m[p0:p1, p0:p1, p0:p1] = True
        sl = [slice(lo, hi + 1)] * 3
        sl[a] = slice(lo, hi)
        m[tuple(sl)] = False
        masks.append(m)
        ln = np.array([g.delta(a, i) if i + 1 < nn[a] else 1.0 for i in range(nn[a])])
        shape = [1, 1, 1]
        shape[a] = nn[a]
        lens.append(np.broadcast_to(ln.reshape(shape), nn))
    amp = math.sqrt(sum(x * x for x in pol))
    centre = [0.5 * (g.lines[a][(lo + hi) // 2] + g.lines[a][(lo + hi) // 2 + 1]) for a in range(3)]
    worst = inc_peak = 0.0
    for step in range(steps):
        e.run(1)
        if record is not None:
            record(e, pw, step)
        if step % 2 == 1:
            continue
        for a in range(3):
            u = e.u_component(a).astype(np.float64)
            worst = max(worst, float(np.max(np.abs(u[masks[a]] / lens[a][masks[a]]))))
        inc = pw.incident_e(centre)
        inc_peak = max(inc_peak, math.sqrt(sum(x * x for x in inc)))
    assert 0.5 < inc_peak / amp <= 1.0, inc_peak
    return 20.0 * math.log10(worst / inc_peak), e, pw
