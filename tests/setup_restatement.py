"""Python restatements of the setup producers the host library runs in C++
(ems_dispersive_fill_cells, ems_dispersive_plan, ems_lumped_plan): the
edges a dispersive cell box reaches and their pole shares
(Dispersive::fill_cells, src/engine/dispersive.rs:1896-1986), a medium's
install plan (Dispersive::install, :2063-2319, pole constants :300-393)
and a lumped element's install (LumpedElement::install,
src/engine/lumped.rs:829-1012). Test infrastructure: the tests compare the
library with these bit for bit."""
from __future__ import annotations

import math
from typing import Dict, List

import numpy as np

from ems_amd.engine import EPS0, MU0

KINDS = ("lorentz", "drude", "debye")


def scaled(kind, p, s):
    """Pole::scaled: omega_p by sqrt(s), a Debye delta_eps by s."""
    if kind == "debye":
        return (p[0] * s, p[1], p[2])
    return (p[0] * math.sqrt(s), p[1], p[2])


def fill_cells(grid, cells, field, poles):
    """[(field, axis, node, kind, p)] in the producer's order."""
    n = grid.nodes()
    out = []
    if any(r[0] == r[1] for r in cells):
        return out
    inside = lambda a, c: cells[a][0] <= c < cells[a][1]  # noqa: E731
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3

        def rng(t, along):
            return range(cells[t][0], cells[t][1]) if along else \
                range(cells[t][0], min(cells[t][1] + 1, n[t]))
        if field == 0:
            ra, rb, rc = rng(a, True), rng(b, False), rng(c, False)
        else:
            ra, rb, rc = rng(a, False), rng(b, True), rng(c, True)
        for pa in ra:
            for pb in rb:
                for pc in rc:
                    node = [0, 0, 0]
                    node[a], node[b], node[c] = pa, pb, pc
                    w_in = w = 0.0
                    if field == 0:
                        for cb in (pb - 1, pb):
                            if cb < 0 or cb >= n[b] - 1:
                                continue
                            for cc in (pc - 1, pc):
                                if cc < 0 or cc >= n[c] - 1:
                                    continue
                                area = grid.delta(b, cb) * grid.delta(c, cc)
                                w += area
                                if inside(b, cb) and inside(c, cc):
                                    w_in += area
                    else:
                        for ca in (pa - 1, pa):
                            if ca < 0 or ca >= n[a] - 1:
                                continue
                            ln = grid.delta(a, ca)
                            w += ln
                            if inside(a, ca):
                                w_in += ln
                    share = w_in / w if w > 0.0 else 0.0
                    if share > 0.0:
                        for kind, p in poles:
                            q = scaled(kind, p, share)
                            if q[0] != 0.0:
                                out.append((field, a, tuple(node), kind, q))
    return out


def lrc(kind, p, m):
    if kind == "lorentz":
        wp, w0, g = p
        l = 1.0 / (m * wp * wp)
        return l, g * l, w0 * w0 * l
    if kind == "drude":
        wp, g = p[:2]
        l = 1.0 / (m * wp * wp)
        return l, g * l, 0.0
    de, tau = p[:2]
    c = m * de
    return 0.0, tau / c, 1.0 / c


def dispersive_plan(entries, leapfrog: bool, grid, dt, lanes) -> dict:
    """entries [(field, axis, node, kind, p)] -> the plan."""
    key = lambda e: (e[2][0], e[0], e[1], e[2][1], e[2][2])  # noqa: E731
    live = [q for q in range(len(entries)) if entries[q][4][0] != 0.0]
    order = sorted(live, key=lambda q: key(entries[q]))
    flat = [lanes[l].reshape(-1) for l in range(12)]
    sides: Dict[int, dict] = {f: {"axis": [], "nodes": [], "sum_inv_z": [], "pole_end": [],
                                  "cidx": []} for f in (0, 1)}
    consts: List[tuple] = []
    cmap = {}
    skipped = 0
    last = None
    for q in order:
        field, axis, node, kind, p = entries[q]
        k = key(entries[q])
        ab, ac = (axis + 1) % 3, (axis + 2) % 3
        idx = grid.index(node)
        if field == 0:
            coef = float(flat[3 + axis][idx])
            geo = grid.dual_delta(ab, node[ab]) * grid.dual_delta(ac, node[ac]) / \
                grid.delta(axis, node[axis])
            m = EPS0 * geo
        else:
            coef = float(flat[9 + axis][idx])
            geo = grid.delta(ab, node[ab]) * grid.delta(ac, node[ac]) / \
                grid.dual_delta(axis, node[axis])
            m = MU0 * geo
        new_edge = last != k
        last = k
        if coef == 0.0:
            if new_edge:
                skipped += 1
            continue
        s = sides[field]
        if new_edge:
            s["axis"].append(axis)
            s["nodes"].append(tuple(node))
            s["sum_inv_z"].append(0.0)
            s["pole_end"].append(len(s["cidx"]))
        l, r, inv_c = lrc(kind, p, m)
        if leapfrog and kind != "debye":
            d = l / dt + 0.5 * r
            cst = ((l / dt - 0.5 * r) / d, inv_c, 1.0 / d, 0.5 * l, 1)
        else:
            a_ = 2.0 * l / dt
            z = a_ + r + 0.5 * dt * inv_c
            cst = (a_, inv_c, 1.0 / z, 0.25 * a_ * dt, 0)
            s["sum_inv_z"][-1] += cst[2]
        ck = tuple(np.float64(v).tobytes() for v in cst[:4]) + (cst[4],)
        if ck not in cmap:
            cmap[ck] = len(consts)
            consts.append(cst)
        s["cidx"].append(cmap[ck])
        s["pole_end"][-1] = len(s["cidx"])
    return {"sides": sides, "consts": consts, "skipped": skipped}


def lumped_parts(box, series, columns, r, l, c, edge_vals, parallel):
    """Per-edge (g, C, branch (R, L, C) or None); values None when absent."""
    inf = math.inf
    if box:
        pl, pc = l is not None, c is not None
    else:
        pl, pc = edge_vals[0][1] is not None, edge_vals[0][2] is not None
    has_branch = pl if parallel else (pl or pc)
    if box:
        m, n = float(series), float(columns)
        if parallel:
            p = (0.0 if r is None else m / (n * r), 0.0 if c is None else m * c / n,
                 None if l is None else (0.0, n * l / m, inf))
        elif not has_branch:
            p = (0.0 if r is None else m / (n * r), 0.0, None)
        else:
            p = (0.0, 0.0, (0.0 if r is None else n * r / m, 0.0 if l is None else n * l / m,
                            inf if c is None else m * c / n))
        return [p] * len(edge_vals), has_branch
    out = []
    for er, el, ec in edge_vals:
        if parallel:
            out.append((0.0 if er is None else 1.0 / er, 0.0 if ec is None else ec,
                        None if el is None else (0.0, el, inf)))
        elif not has_branch:
            out.append((0.0 if er is None else 1.0 / er, 0.0, None))
        else:
            out.append((0.0, 0.0, (0.0 if er is None else er, 0.0 if el is None else el,
                                   inf if ec is None else ec)))
    return out, has_branch


def fold_parallel(ca, cb, dt, g, c):
    """fold_parallel (lumped.rs:275-282)."""
    a, b = float(np.float32(ca)), float(np.float32(cb))
    cap = dt * (1.0 + a) / (2.0 * b) + c
    cond = (1.0 - a) / b + g
    den = cap + 0.5 * cond * dt
    return float(np.float32((cap - 0.5 * cond * dt) / den)), float(np.float32(dt / den))


def lumped_plan(nodes, axis, edges, parts, has_branch, parallel, explicit, dt, ca, cb) -> dict:
    """ca, cb: flat copies of the axis's lanes, folded in place."""
    n = nodes
    out = {"cb_e": [], "g_src": [], "rec_g": [], "rec_c": [], "two_l_dt": [], "inv_c": [], "z": []}
    for p, (g_e, c_e, br) in zip(edges, parts):
        i = (p[0] * n[1] + p[1]) * n[2] + p[2]
        if g_e != 0.0 or c_e != 0.0:
            ca[i], cb[i] = fold_parallel(ca[i], cb[i], dt, g_e, c_e)
        out["cb_e"].append(float(cb[i]))
        out["rec_g"].append(g_e)
        out["rec_c"].append(c_e)
        out["g_src"].append(0.0 if (has_branch and not parallel) else g_e)
        if has_branch:
            r, l, c = br if br is not None else (0.0, 0.0, math.inf)
            ic = 1.0 / c if math.isfinite(c) else 0.0
            out["two_l_dt"].append(2.0 * l / dt)
            out["inv_c"].append(ic)
            out["z"].append(2.0 * l / dt + r + 0.5 * dt * ic)
    out["branch"] = (2 if (explicit and parallel) else 1) if has_branch else 0
    out["src_in_branch"] = has_branch and not parallel
    return out
