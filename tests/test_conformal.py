"""Conformal PEC coefficient patches (ems_amd/conformal.py) on the CPU:
the circle-rectangle area against brute force, the cut fractions against
chord lengths and the disk area, the patch arithmetic of
ConformalPec::apply (conformal.rs:786-793, 1542-1566), and the physics:
the TM010 resonance of a circular cavity, whose frequency the Bessel zero
fixes, is found closer with the conformal wall than with a staircase,
and the energy invariant of the patched lossless cavity stays constant."""
import math

import numpy as np
import pytest

import ems_amd as E
from ems_amd import conformal as K
from oracle.bridge import OracleEngine

J01 = 2.404825557695773  # first zero of J0


@pytest.mark.parametrize("rect", [(0.3, 0.8, -0.1, 0.4), (0.5, 1.5, 0.5, 1.5), (-2, 2, -2, 2),
                                  (-0.2, 0.1, -1.2, -0.7), (0.9, 1.3, -0.2, 0.2),
                                  (1.1, 2.0, 0.0, 1.0)])
def test_disk_rectangle_area_matches_brute_force(rect):
    x0, x1, y0, y1 = rect
    m = 2000
    xs = x0 + (np.arange(m) + 0.5) * (x1 - x0) / m
    ys = y0 + (np.arange(m) + 0.5) * (y1 - y0) / m
    inside = (xs[:, None] ** 2 + ys[None, :] ** 2) < 1.0
    ref = inside.mean() * (x1 - x0) * (y1 - y0)
    got = K._disk_rect_area_exact(x0, x1, y0, y1, 0.0, 0.0, 1.0)
    assert abs(got - ref) < 2e-3 * (x1 - x0) * (y1 - y0)


def test_cut_fractions_add_up_to_the_disk():
    g = E.Grid.uniform([20, 20, 2], [1e-3] * 3)
    c, r = 10.3e-3, 7.1e-3
    m = E.CellMaterials.vacuum(g)
    pec = K.cylinder_z(g, (c, c), r, mats=m, metal="inside")
    xs = np.asarray(g.lines[0])
    # closed area of the z faces of one plane = pi r^2 (cell area 1 mm^2):
    # whole cells inside plus the cut ones
    cut = {f.node[:2]: f.fraction for f in pec.faces if f.axis == 2 and f.node[2] == 0}
    whole = 0
    for i in range(20):
        for j in range(20):
            if (i, j) in cut:
                continue
            corners = [((xs[i + a] - c) ** 2 + (xs[j + b] - c) ** 2) < r * r
                       for a in (0, 1) for b in (0, 1)]
            whole += all(corners)
    closed = whole + sum(1.0 - f for f in cut.values())
    assert abs(closed * 1e-6 - math.pi * r * r) < 1e-9 * 1e-6 * 1e3
    # the closed length of the x edges on line y_j is the chord there
    for j in range(21):
        d = r * r - (xs[j] - c) ** 2
        chord = 2 * math.sqrt(d) if d > 0 else 0.0
        edges = [e for e in pec.edges if e.axis == 0 and e.node[1] == j and e.node[2] == 0]
        pec_x = int(m.pec[0, :, j, 0].sum()) if m.pec is not None else 0
        closed_len = (pec_x + sum(1.0 - e.fraction for e in edges)) * 1e-3
        assert abs(closed_len - chord) < 1e-12


def test_patch_arithmetic():
    """Edge scaling (cb f_e, PEC at 0), db / f_A for a face above the
    threshold, and a small face without a centroid merged with the
    neighbour across its most open edge: both frozen (db = 0), driven by
    one group current with a deviation on the small member."""
    g = E.Grid.uniform([6, 6, 4], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    pec = K.ConformalPec(K.ConformalOptions(0.5))
    pec.push_edge(K.ConformalEdge(0, (2, 3, 1), 0.37))
    pec.push_edge(K.ConformalEdge(2, (3, 3, 1), 0.0))
    pec.push_face(K.ConformalFace(2, (2, 2, 1), 0.8))
    pec.push_face(K.ConformalFace(1, (2, 3, 1), 0.2))      # below the threshold
    changed, rep, op = pec.build(lanes, g, dt)
    assert sorted(changed) == [3, 5, 10, 11]
    assert changed[3][2, 3, 1] == np.float32(float(lanes[3][2, 3, 1]) * 0.37)
    assert changed[5][3, 3, 1] == 0.0
    assert changed[11][2, 2, 1] == np.float32(float(lanes[11][2, 2, 1]) / 0.8)
    assert changed[10][2, 3, 1] == 0.0 and changed[10][1, 3, 1] == 0.0
    assert (rep.edges, rep.faces, rep.merged, rep.groups, rep.floored, rep.pec_edges) == \
        (2, 1, 2, 1, 0, 1)
    # the group current drives both faces, the small one plus its deviation
    assert op.faces == [(1, (1, 3, 1), [(0, 1.0)]), (1, (2, 3, 1), [(0, 1.0), (1, 1.0)])]
    assert op.scalar == 0 and len(op.blocks) == 1 and len(op.blocks[0].members) == 2
    # everything else is untouched
    for l, lane in changed.items():
        diff = np.argwhere(lane != lanes[l])
        assert len(diff) == (2 if l == 10 else 1)
    bad = K.ConformalPec()
    bad.push_edge(K.ConformalEdge(0, (6, 0, 0), 0.5))   # runs off the grid
    with pytest.raises(E.EmsError):
        bad.build(lanes, g, dt)
    bad = K.ConformalPec()
    bad.push_face(K.ConformalFace(2, (1, 1, 1), 1.5))
    with pytest.raises(E.EmsError):
        bad.build(lanes, g, dt)


def _cavity(mode, a_cells, steps, tau=0.5):
    h = 1e-3
    n = int(math.ceil(2 * a_cells)) + 2
    g = E.Grid.uniform([n, n, 2], [h] * 3)
    c, a = n * h / 2, a_cells * h
    m = E.CellMaterials.vacuum(g)
    xs = np.asarray(g.lines[0])
    if mode == "conformal":
        pec = K.cylinder_z(g, (c, c), a, mats=m, metal="outside", options=K.ConformalOptions(tau))
    else:
        N = g.nodes()
        for i in range(N[0]):
            for j in range(N[1]):
                for k in range(N[2]):
                    # staircase: an edge is metal where its midpoint is
                    if i + 1 < N[0] and ((xs[i] + xs[i + 1]) / 2 - c) ** 2 + (xs[j] - c) ** 2 >= a * a:
                        m.set_pec_edge(0, (i, j, k))
                    if j + 1 < N[1] and (xs[i] - c) ** 2 + ((xs[j] + xs[j + 1]) / 2 - c) ** 2 >= a * a:
                        m.set_pec_edge(1, (i, j, k))
                    if k + 1 < N[2] and (xs[i] - c) ** 2 + (xs[j] - c) ** 2 >= a * a:
                        m.set_pec_edge(2, (i, j, k))
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, m, [E.FaceCondition.Pec] * 6, dt)
    o = OracleEngine(g, dt, lanes, threads=4)
    if mode == "conformal":
        pec.apply(o)
    ic = n // 2
    w = E.Waveform.gaussian(15e9, 10e9)
    o.add_operator(E.EdgeSource(w).edge(E.Axis.Z, [ic - 1, ic, 0], 1.0))
    p = o.add_operator(E.Probe.voltage(E.Axis.Z, [ic, ic, 0]))
    return o, p, dt, a, int(w.duration() / dt) + 10


def _peak(v, dt, lo, hi):
    t = np.arange(v.size) * dt
    f = np.linspace(lo, hi, 4001)
    X = np.abs(np.exp(-2j * np.pi * f[:, None] * t[None, :]) @ v)
    k = int(np.argmax(X))
    y0, y1, y2 = X[k - 1], X[k], X[k + 1]
    return f[k] + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2) * (f[1] - f[0])


def test_tm010_resonance_closer_than_staircase():
    errs = {}
    for mode in ("staircase", "conformal"):
        o, p, dt, a, w_end = _cavity(mode, 7.3, 5000)
        o.run(5000)
        v = o.probe_values(p)[w_end:]
        fa = J01 * 299792458.0 / (2 * math.pi * a)
        errs[mode] = abs(_peak(v, dt, 13e9, 18e9) - fa) / fa
        o.close()
    assert errs["conformal"] < 0.015, errs
    assert errs["conformal"] < 0.4 * errs["staircase"], errs


def test_patched_cavity_conserves_energy():
    o, p, dt, a, w_end = _cavity("conformal", 6.3, 0)
    o.run(w_end)
    w0 = o.step_with_energy()
    for _ in range(6):
        o.run(2000)
        w = o.step_with_energy()
        assert abs(w - w0) <= 1e-5 * w0, (w, w0)
    o.close()
