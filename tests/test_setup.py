"""The host side above the boundary, pinned by the reference's own unit
tests: Courant limit and coefficient formulas (src/engine/coeffs.rs:1088-1203),
CPML grading (src/engine/boundary/cpml.rs:1260-1429), and the product's
coefficient builder against the oracle's restatement of
Coefficients::from_edges fed with per-edge materials computed independently
here (numpy restatement of primary_edge_cg / dual_edge_lr)."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import engine as EE
from oracle.bridge import oracle_coefficients, oracle_cpml_coefficients, orc_courant_limit, P
import ctypes as C


def rel(a, b):
    return abs((a - b) / b)


def cap_from(lanes, axis, idx, dt):
    """Coefficients::capacitance (coeffs.rs:1070-1076)."""
    ca, cb = float(lanes[axis][idx]), float(lanes[3 + axis][idx])
    return 0.0 if cb == 0.0 else dt * (1.0 + ca) / (2.0 * cb)


def ind_from(lanes, axis, idx, dt):
    da, db = float(lanes[6 + axis][idx]), float(lanes[9 + axis][idx])
    return 0.0 if db == 0.0 else dt * (1.0 + da) / (2.0 * db)


def test_courant():
    g = E.Grid.uniform([4, 4, 4], [1e-3, 2e-3, 1e-3])
    lim = E.courant_limit(g)
    want = 1.0 / (E.C0 * math.sqrt(1e6 + 0.25e6 + 1e6))
    assert rel(lim, want) < 1e-14
    assert rel(E.TimeStep.courant(0.99).resolve(g), 0.99 * want) < 1e-14
    assert E.TimeStep.fixed(0.5 * lim).resolve(g) == 0.5 * lim
    E.TimeStep.fixed(lim).resolve(g)
    for bad in (E.TimeStep.fixed(1.0001 * lim), E.TimeStep.fixed(-1.0), E.TimeStep.fixed(float("nan")),
                E.TimeStep.courant(1.01), E.TimeStep.courant(0.0)):
        with pytest.raises(E.EmsError):
            bad.resolve(g)
    g2 = E.Grid([[0.0, 1.0, 1.5], [0.0, 1.0], [0.0, 1.0]])
    assert rel(E.courant_limit(g2), 1.0 / (E.C0 * math.sqrt(4.0 + 1.0 + 1.0))) < 1e-14
    # The oracle's restatement agrees bit for bit.
    x, y, z = g2.lines
    o = orc_courant_limit(P(x, C.c_double), 3, P(y, C.c_double), 2, P(z, C.c_double), 2)
    assert o == E.courant_limit(g2)


def test_edge_between_dielectric_and_vacuum():
    """coeffs.rs:1108-1137: area-weighted parallel combination."""
    g = E.Grid([[0.0, 1.0, 2.0], [0.0, 1.0, 3.0], [0.0, 2.0, 3.0]])
    m = E.CellMaterials.vacuum(g)
    m.fill_box([(0, 2), (0, 1), (0, 2)], E.Material.dielectric(4.0).with_kappa(2.0))
    dt = 1e-9
    lanes = E.coefficients(g, m, [E.FaceCondition.Pmc] * 6, dt)
    idx = (0, 1, 1)
    a_diel, a_vac = 0.5 * (1.0 + 0.5), 1.0 * (1.0 + 0.5)
    want_c = E.EPS0 * (4.0 * a_diel + a_vac) / 1.0
    want_g = 2.0 * a_diel
    ca, cb = float(lanes[0][idx]), float(lanes[3][idx])
    assert rel(cb, dt / (want_c + 0.5 * want_g * dt)) < 1e-6
    assert rel(ca, (want_c - 0.5 * want_g * dt) / (want_c + 0.5 * want_g * dt)) < 1e-6
    # Corner edge at the domain edge: a single quarter cell.
    cb = float(lanes[3][(1, 0, 0)])
    c = E.EPS0 * 4.0 * 0.5 * 1.0
    assert rel(cb, dt / (c + 0.5 * 2.0 * 0.5 * dt)) < 1e-6


def test_dual_edge_series():
    """coeffs.rs:1139-1161: half cells in series."""
    g = E.Grid([[0.0, 1.0, 3.0], [0.0, 1.0], [0.0, 2.0]])
    m = E.CellMaterials.vacuum(g)
    m.set((1, 0, 0), E.Material.vacuum().with_mu_r(3.0).with_sigma_m(5.0))
    dt = 1e-9
    lanes = E.coefficients(g, m, [E.FaceCondition.Pmc] * 6, dt)
    area = 2.0
    l1, l2 = E.MU0 * area / 0.5, E.MU0 * 3.0 * area / 1.0
    L = 1.0 / (1.0 / l1 + 1.0 / l2)
    R = L * L * (5.0 * area / 1.0) / (l2 * l2)
    da, db = float(lanes[6][(1, 0, 0)]), float(lanes[9][(1, 0, 0)])
    assert rel(db, dt / (L + 0.5 * R * dt)) < 1e-6
    assert rel(da, (L - 0.5 * R * dt) / (L + 0.5 * R * dt)) < 1e-6
    # Boundary dual edge: single half cell (mu_r 3 cell only).
    db = float(lanes[9][(2, 0, 0)])
    assert rel(db, dt / (E.MU0 * 3.0 * area / 1.0 + 0.5 * (5.0 * area / 1.0) * dt)) < 1e-6


def test_faces_pec_and_loss():
    """coeffs.rs:1163-1202."""
    g = E.Grid.uniform([3, 3, 3], [1e-3] * 3)
    m = E.CellMaterials(g, E.Material().with_kappa(1.0))
    m.set_pec_edge(2, (1, 1, 1))
    dt = E.TimeStep.courant(0.99).resolve(g)
    faces = [E.FaceCondition.Pec] * 6
    co = E.coefficients(g, m, faces, dt)
    assert co[4][0, 1, 1] == 0.0 and co[4][1, 1, 1] > 0.0
    assert co[5][1, 1, 1] == 0.0 and co[2][1, 1, 1] == 0.0
    assert co[3][0, 1, 1] > 0.0
    assert co[3][3, 1, 1] == 0.0 and co[9][1, 3, 1] == 0.0
    c = cap_from(co, 0, (1, 1, 1), dt)
    ca = float(co[0][1, 1, 1])
    geo = 1e-3
    want_c, want_g = E.EPS0 * geo, 1.0 * geo
    assert rel(ca, (want_c - want_g * dt / 2) / (want_c + want_g * dt / 2)) < 1e-6
    assert ca < 1.0
    assert rel(c, want_c) < 1e-6
    faces[E.Face.XMin] = E.FaceCondition.Pmc
    co = E.coefficients(g, m, faces, dt)
    assert rel(cap_from(co, 1, (0, 1, 1), dt), 0.5 * c) < 1e-6
    assert co[4][0, 1, 1] > 0.0
    assert co[4][0, 1, 0] == 0.0


def edge_materials(grid, mats):
    """numpy restatement of EdgeMaterials::from_cells (edge_material.rs:376-417)."""
    n = grid.nodes()
    d = [np.diff(l) for l in grid.lines]
    dual = [0.5 * (np.concatenate([[0.0], dd]) + np.concatenate([dd, [0.0]])) for dd in d]
    tab = np.array([m.row() for m in mats.table])
    ids = mats.ids
    eps = np.zeros((3,) + n); kap = np.zeros((3,) + n); mu = np.ones((3,) + n); sig = np.zeros((3,) + n)
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for i in range(n[0]):
            for j in range(n[1]):
                for k in range(n[2]):
                    p = (i, j, k)
                    if p[a] + 1 < n[a]:
                        l = d[a][p[a]]
                        cap = 0.0; cond = 0.0
                        for cb in (p[b] - 1, p[b]):
                            if cb < 0 or cb >= n[b] - 1:
                                continue
                            for cc in (p[c] - 1, p[c]):
                                if cc < 0 or cc >= n[c] - 1:
                                    continue
                                ce = [0, 0, 0]; ce[a] = p[a]; ce[b] = cb; ce[c] = cc
                                area = 0.25 * d[b][cb] * d[c][cc]
                                mm = tab[ids[tuple(ce)]]
                                cap += E.EPS0 * mm[a] * area
                                cond += mm[6 + a] * area
                        cap, cond = cap / l, cond / l
                        kk = d[a][p[a]] / (dual[b][p[b]] * dual[c][p[c]])
                        eps[a][p] = cap * kk / E.EPS0
                        kap[a][p] = cond * kk
                    if p[b] + 1 < n[b] and p[c] + 1 < n[c]:
                        area = d[b][p[b]] * d[c][p[c]]
                        inv_l = 0.0; rl2 = 0.0
                        for ca in (p[a] - 1, p[a]):
                            if ca < 0 or ca >= n[a] - 1:
                                continue
                            ce = list(p); ce[a] = ca
                            half = 0.5 * d[a][ca]
                            mm = tab[ids[tuple(ce)]]
                            lc = E.MU0 * mm[3 + a] * area / half
                            rc = mm[9 + a] * area / half
                            inv_l += 1.0 / lc
                            rl2 += rc / (lc * lc)
                        L = 1.0 / inv_l
                        R = L * L * rl2
                        kk = dual[a][p[a]] / (d[b][p[b]] * d[c][p[c]])
                        mu[a][p] = L * kk / E.MU0
                        sig[a][p] = R * kk
    return eps, kap, mu, sig


def test_builder_matches_oracle_from_edges_bitwise():
    from cases import graded
    g = E.Grid([graded(0.02, 7, 0.2), graded(0.015, 6, 0.1), graded(0.01, 5, 0.25)])
    m = E.CellMaterials.vacuum(g)
    m.fill_box([(1, 4), (2, 5), (1, 3)], E.Material(eps_r=(4.3, 2.0, 3.0), kappa=(0.02, 0.5, 0.0)))
    m.fill_box([(4, 7), (0, 3), (2, 5)], E.Material(mu_r=(1.5, 2.0, 1.2), sigma_m=(50.0, 0.0, 3.0)))
    m.fill_pec_box([(5, 6), (3, 5), (1, 4)])
    faces = [E.FaceCondition.Pec, E.FaceCondition.Pmc, E.FaceCondition.Pmc, E.FaceCondition.Pec,
             E.FaceCondition.Pec, E.FaceCondition.Pmc]
    dt = E.TimeStep.courant(0.95).resolve(g)
    mine = E.coefficients(g, m, faces, dt)
    eps, kap, mu, sig = edge_materials(g, m)
    pf = [1 if f == E.FaceCondition.Pec else 0 for f in faces]
    ref = oracle_coefficients(g, eps, kap, mu, sig, m.pec, pf, dt)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


def test_pooling_roundtrip_and_ratio():
    g = E.Grid.uniform([9, 7, 5], [1e-3] * 3)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, 1e-12)
    for l in range(12):
        pooled = E.pool_lane(lanes[l])
        assert pooled is not None
        rows, index = pooled
        assert np.array_equal(rows[index].reshape(lanes[l].shape).view(np.uint32),
                              lanes[l].view(np.uint32))
    rng = np.random.default_rng(0)
    dense = rng.standard_normal((8, 8, 4)).astype(np.float32)
    assert E.pool_lane(dense) is None  # 64 distinct rows > 64/4


# ---- CPML (cpml.rs:1260-1429) ----

def prof(b, face, side):
    return b.e_profile(face) if side == "e" else b.h_profile(face)


def test_cpml_grading_positions():
    g = E.Grid.uniform([20, 4, 4], [1e-3] * 3)
    dt = 1e-12
    p = E.Cpml.with_cells(4)
    b = E.CpmlBoundary.with_grid(g, dt, [(E.Face.XMin, p), (E.Face.XMax, p)])
    for face in (E.Face.XMin, E.Face.XMax):
        e, h = b.e_profile(face), b.h_profile(face)
        assert (len(e), len(h)) == (4, 4)
        ew, hw = [1.0, 0.75, 0.5, 0.25], [0.875, 0.625, 0.375, 0.125]
        if face == E.Face.XMax:
            ew.reverse(), hw.reverse()
        for a, w in list(zip([c["rho"] for c in e], ew)) + list(zip([c["rho"] for c in h], hw)):
            assert abs(a - w) < 1e-12
    c = b.e_profile(E.Face.XMin)[0]
    smax = 0.8 * 4.0 / (E.Z0 * 1e-3)
    assert abs(c["sigma"] / smax - 1.0) < 1e-12
    assert abs(c["b"] - math.exp(-smax * dt / E.EPS0)) < 1e-15
    x = smax * dt / E.EPS0
    a1 = -(1.0 - c["b"] - x * c["b"]) / x
    assert abs(c["a"] - (c["b"] - 1.0 - a1 + a1 / c["b"])) < 1e-12
    assert abs(c["direct"] + a1 / c["b"]) < 1e-12
    old = E.cpml_coefficients(E.Cpml(cells=4, convolution=0), 1.0, 1e-3, dt)
    assert abs(old["a"] - (old["b"] - 1.0)) < 1e-15
    assert old["direct"] == 0.0
    q = E.Cpml(cells=4, kappa_max=5.0, alpha_max=0.05)
    c = E.cpml_coefficients(q, 0.0, 1e-3, dt)
    assert (c["sigma"], c["kappa"], c["a"]) == (0.0, 1.0, 0.0)
    assert abs(c["alpha"] - 0.05) < 1e-15
    c = E.cpml_coefficients(q, 1.0, 1e-3, dt)
    assert (c["kappa"], c["alpha"]) == (5.0, 0.0)


def test_cpml_linear_convolution_is_second_order():
    w = 2.0 * math.pi * 9e9

    def err(conv, dt):
        q = E.Cpml(cells=6, kappa_max=2.0, alpha_max=0.3, convolution=conv)
        c = E.cpml_coefficients(q, 0.5, 2e-4, dt)
        p = (c["sigma"] / c["kappa"] + c["alpha"]) / E.EPS0
        cc = c["sigma"] / (E.EPS0 * c["kappa"] ** 2)
        d = p * p + w * w
        want = (-cc * p / d, cc * w / d)
        re, im = 1.0 - c["b"] * math.cos(w * dt), c["b"] * math.sin(w * dt)
        n = re * re + im * im
        got = (c["direct"] + 1.0 - 1.0 / c["kappa"] + c["a"] * re / n, -c["a"] * im / n)
        return math.hypot(got[0] - want[0], got[1] - want[1]) / math.hypot(*want)

    c1, c2 = err(0, 2e-13), err(0, 1e-13)
    l1, l2 = err(1, 2e-13), err(1, 1e-13)
    assert abs(c1 / c2 - 2.0) < 0.05
    assert abs(c2 / (w * 1e-13 / 2.0) - 1.0) < 0.05
    assert abs(l1 / l2 - 4.0) < 0.1
    assert l1 < c1 / 100.0
    assert err(1, 1e-17) < 1e-6 and err(0, 1e-17) < 1e-5
    big = E.Cpml(cells=4, sigma_factor=20.0)
    lin = E.cpml_coefficients(big, 1.0, 1e-3, 1e-12)
    con = E.cpml_coefficients(E.Cpml(cells=4, sigma_factor=20.0, convolution=0), 1.0, 1e-3, 1e-12)
    assert lin["sigma"] * 1e-12 / E.EPS0 > 8.0
    assert (lin["a"], lin["direct"]) == (con["a"], con["direct"])
    # The oracle's restatement is the same function bit for bit.
    for conv in (0, 1):
        for rho in (0.0, 0.3, 0.875, 1.0):
            q = E.Cpml(cells=6, kappa_max=2.0, alpha_max=0.3, convolution=conv)
            assert E.cpml_coefficients(q, rho, 2e-4, 1e-13) == oracle_cpml_coefficients(q, rho, 2e-4, 1e-13)


def test_cpml_graded_local_delta():
    y = [0.0, 1e-3, 2e-3, 3e-3, 5e-3, 8e-3]
    g = E.Grid([[0.0, 1e-3], y, [0.0, 1e-3]])
    p = E.Cpml.with_cells(3)
    b = E.CpmlBoundary.with_grid(g, 1e-12, [(E.Face.YMax, p)])
    smax = lambda d: 0.8 * 4.0 / (E.Z0 * d)
    h = b.h_profile(E.Face.YMax)
    assert abs(h[2]["rho"] - 2.5 / 3.0) < 1e-12
    assert abs(h[2]["sigma"] / (smax(3e-3) * (2.5 / 3.0) ** 3) - 1.0) < 1e-12
    e = b.e_profile(E.Face.YMax)
    assert abs(e[1]["sigma"] / (smax(2.5e-3) * (2.0 / 3.0) ** 3) - 1.0) < 1e-12
    assert abs(e[2]["sigma"] / smax(3e-3) - 1.0) < 1e-12


def test_cpml_layer_refractive_index():
    g = E.Grid.uniform([12, 6, 6], [1e-3] * 3)

    class _E:  # the engine attributes CpmlBoundary.new reads
        def __init__(self, m):
            self.grid = g
            self.dt = E.TimeStep.courant(0.99).resolve(g)
            self._dense_lanes = E.coefficients(g, E.CellMaterials(g, m), [E.FaceCondition.Pec] * 6,
                                               self.dt)

    p = E.Cpml.with_cells(4)
    b = E.CpmlBoundary.new(_E(E.Material()), [(E.Face.XMin, p), (E.Face.ZMax, p)])
    assert b.refractive_index(E.Face.XMin) == 1.0
    assert b.refractive_index(E.Face.YMin) is None
    b = E.CpmlBoundary.new(_E(E.Material.dielectric(4.0)), [(E.Face.XMin, p), (E.Face.ZMax, p)])
    n = b.refractive_index(E.Face.ZMax)
    assert abs(n - 2.0) < 1e-6
    e = b.e_profile(E.Face.ZMax)[-1]
    assert abs(e["sigma"] / (0.8 * 4.0 / (E.Z0 * 1e-3) / n) - 1.0) < 1e-12
    n = E.CpmlBoundary.new(_E(E.Material(eps_r=(2.0,) * 3, mu_r=(3.0,) * 3)),
                           [(E.Face.XMin, p)]).refractive_index(E.Face.XMin)
    assert abs(n - math.sqrt(6.0)) < 1e-6


def test_cpml_validation():
    g = E.Grid.uniform([10, 4, 4], [1e-3] * 3)
    ok = lambda l: E.CpmlBoundary.with_grid(g, 1e-12, l)
    ok([(E.Face.XMin, E.Cpml.with_cells(5)), (E.Face.XMax, E.Cpml.with_cells(5))])
    for bad in ([(E.Face.XMin, E.Cpml.with_cells(6)), (E.Face.XMax, E.Cpml.with_cells(5))],
                [(E.Face.YMin, E.Cpml.with_cells(5))],
                [(E.Face.XMin, E.Cpml.with_cells(2)), (E.Face.XMin, E.Cpml.with_cells(2))],
                [(E.Face.XMin, E.Cpml.with_cells(0))]):
        with pytest.raises(E.EmsError):
            ok(bad)
    with pytest.raises(E.EmsError):
        E.CpmlBoundary.with_grid(g, 0.0, [])


def test_gaussian_waveform():
    """waveform.rs:240-248, 442-453, 506-513: zero outside [0, 2T], peak at T,
    ends below 1e-6 (waveform.rs:591-610)."""
    w = E.Waveform.gaussian(2e9, 1e9)
    T = 0.5 * w.duration()
    tau = math.sqrt(math.log(10.0)) / (math.pi * 1e9)
    assert abs(T - tau * math.sqrt(6.0 * math.log(10.0))) < 1e-24
    assert w.value(-1e-15) == 0.0 and w.value(2 * T + 1e-15) == 0.0
    assert w.value(T) == 1.0
    assert abs(w.value(0.0)) < 1e-6 and abs(w.value(2 * T)) < 1e-6
    t = np.linspace(0.0, 2 * T, 101)
    want = np.cos(2 * math.pi * 2e9 * (t - T)) * np.exp(-((t - T) / tau) ** 2)
    got = np.array([w.value(x) for x in t])
    assert np.max(np.abs(got - want)) < 1e-15
