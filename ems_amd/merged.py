"""Small cut faces without their own current: interpolation and merging
(src/engine/conformal.rs, ConformalPec::construct :1041-1589 and the
MergedFaces operator :2313-2700).

A cut face whose open area fraction f_A is below the merge threshold tau
loses its own current (conformal.rs:34-95). Where a stencil exists (the
face has an open-area centroid, a host face across its most open edge and
a neighbour of the host along each in-plane axis), its current is the
affine interpolation of those faces' currents at the centroid plus a
deviation current of its own; otherwise it is merged with the neighbour
across its most open edge into a group sharing one current (again with a
deviation per small member). The reduced currents g obey the Galerkin
restriction of the Dey-Mittra scheme,

    (P^T L P) dg/dt + (P^T R P) g = -P^T K u ,

solved each step after the H update: a merge group without deviations is
a scalar update, coupled currents form blocks solved with an envelope
Cholesky factor in reverse Cuthill-McKee order. The faces the reduced
currents drive are frozen in the kernel (da = 1, db = 0) and written by
the operator.

This module builds the operator on the host, in f64, in the reference's
order (the iteration orders of its maps are pinned by sorting there and
here): the stencils and merge groups, the mass rows, the blocks and their
factors, and the coefficient patch. The per-step update runs on the device
(ems_gpu_add_merged) and in the oracle.

Inside CPML layers (set_cpml) small faces are merged, and the layer's
H-side correction of every driven face there is added to the circulations
of the reduced currents driving it (PmlTerm, conformal.rs:2398-2412).
Scope: Cartesian grids, no thin sheets, no cylindrical PML. The stability
estimate (power iteration) runs in the host library
(ems_conformal_stability); `stability` and `grid_bound` feed the rounds of
ConformalPec.build.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Set, Tuple

import numpy as np

CA, CB, DA, DB = 0, 3, 6, 9
SIGNS = (1, 1, -1, -1)  # conformal.rs:2160: +u_b(p) + u_c(p+e_b) - u_b(p+e_c) - u_c(p)
MAX_WEIGHT = 2.5        # conformal.rs:1878

Key = Tuple[int, int]  # (normal axis, flat node index)


def others(a: int) -> Tuple[int, int]:
    """Axis::others (axis.rs:55-57): (next, prev) in cyclic order."""
    return (a + 1) % 3, (a + 2) % 3


def index(n, p) -> int:
    return (p[0] * n[1] + p[1]) * n[2] + p[2]


def face_exists(a: int, p, n) -> bool:
    """conformal.rs:2163-2169."""
    b, c = others(a)
    return all(p[d] < n[d] for d in range(3)) and p[b] + 1 < n[b] and p[c] + 1 < n[c]


def face_edges(a: int, p) -> List[Tuple[int, Tuple[int, int, int]]]:
    """conformal.rs:2173-2181: the four edges bounding the face, in
    perimeter order (b at p, c at p + e_b, b at p + e_c, c at p)."""
    b, c = others(a)
    pb = list(p)
    pb[b] += 1
    pc = list(p)
    pc[c] += 1
    p = tuple(p)
    return [(b, p), (c, tuple(pb)), (b, tuple(pc)), (c, p)]


def in_plane_neighbours(a: int, p, n):
    """conformal.rs:2184-2208: (shared edge axis, shared edge node,
    neighbour face node) of the existing in-plane neighbours."""
    b, c = others(a)
    out = []
    for edge, off in ((b, c), (c, b)):
        if p[off] > 0:
            q = list(p)
            q[off] -= 1
            out.append((edge, tuple(p), tuple(q)))
        q = list(p)
        q[off] += 1
        q = tuple(q)
        if face_exists(a, q, n):
            out.append((edge, q, q))
    return out


def touches_layer(absorbing: Sequence[int], a: int, p, n) -> bool:
    """conformal.rs:678-697: whether the face touches an absorbing layer of
    `absorbing` cells per face (x-, x+, y-, y+, z-, z+), its interface
    plane included."""
    for f in range(6):
        cells = absorbing[f]
        if cells == 0:
            continue
        d = f // 2
        last = n[d] - 1
        normal = d == a
        if f % 2 == 1:
            lo = max(last - cells, 0)
            if p[d] + (0 if normal else 1) > lo or (normal and p[d] >= lo):
                return True
        elif p[d] < cells or (normal and p[d] <= cells):
            return True
    return False


def interpolation_weights(d: Sequence[Tuple[float, float]]) -> Optional[List[float]]:
    """conformal.rs:1887-1917: weights with sum 1 and sum w*d = 0 (exact for
    affine fields) for 1, 2 or 3 stencil offsets; None if degenerate, off
    the point/line, or a weight above 2.5 in magnitude."""
    on = 1e-6
    if len(d) == 1:
        p = d[0]
        if math.hypot(p[0], p[1]) < on:
            w = [1.0]
        else:
            return None
    elif len(d) == 2:
        p, q = d
        e = (q[0] - p[0], q[1] - p[1])
        len2 = e[0] * e[0] + e[1] * e[1]
        if len2 < 1e-6:
            return None
        t = -(p[0] * e[0] + p[1] * e[1]) / len2
        if math.hypot(p[0] + t * e[0], p[1] + t * e[1]) > on:
            return None
        w = [1.0 - t, t]
    elif len(d) == 3:
        p, q, r = d
        det = (q[0] - p[0]) * (r[1] - p[1]) - (r[0] - p[0]) * (q[1] - p[1])
        if abs(det) < 1e-3:
            return None
        wq = ((-p[0]) * (r[1] - p[1]) - (r[0] - p[0]) * (-p[1])) / det
        wr = ((q[0] - p[0]) * (-p[1]) - (-p[0]) * (q[1] - p[1])) / det
        w = [1.0 - wq - wr, wq, wr]
    else:
        return None
    return w if all(abs(x) <= MAX_WEIGHT for x in w) else None


def add_terms(terms: Dict[Key, list], a: int, p, w: float, n) -> None:
    """conformal.rs:1942-1955: adds w times the curl of face (a, p)."""
    for (e, q), s in zip(face_edges(a, p), SIGNS):
        k = (e, index(n, q))
        t = terms.get(k)
        if t is None:
            terms[k] = [e, q, w * float(s)]
        else:
            t[2] += w * float(s)


def sorted_terms(terms: Dict[Key, list], n) -> List[Tuple[int, Tuple[int, int, int], float]]:
    """conformal.rs:1958-1967: the non-zero terms in (axis, index) order."""
    t = [tuple(v) for v in terms.values() if abs(v[2]) > 1e-12]
    t.sort(key=lambda v: (v[0], index(n, v[1])))
    return t


class Groups:
    """Union-find over keys with sorted member lists and areas per root
    (conformal.rs:2211-2272); the smaller key stays the root."""

    def __init__(self):
        self.parent: Dict = {}
        self.node: Dict = {}
        self.members: Dict = {}
        self.area: Dict = {}

    def add(self, k, node, area: float) -> None:
        self.parent[k] = k
        self.node[k] = node
        self.members[k] = [k]
        self.area[k] = area

    def contains(self, k) -> bool:
        return k in self.parent

    def find(self, k):
        r = k
        while self.parent[r] != r:
            r = self.parent[r]
        c = k
        while self.parent[c] != r:
            nxt = self.parent[c]
            self.parent[c] = r
            c = nxt
        return r

    def union(self, a, b) -> None:
        ra, rb = self.find(a), self.find(b)
        if ra == rb:
            return
        keep, gone = (ra, rb) if ra < rb else (rb, ra)
        self.parent[gone] = keep
        self.area[keep] += self.area.pop(gone, 0.0)
        mv = self.members.pop(gone, [])
        lst = self.members[keep]
        lst.extend(mv)
        lst.sort()

    def groups(self):
        g = [(r, list(m)) for r, m in self.members.items() if len(m) > 1]
        g.sort()
        return g


@dataclass
class Group:
    """One reduced current (conformal.rs:2275-2308): circulation terms
    (edge axis, edge node, weight), the update factors and the inductance."""
    terms: list
    da: float
    db: float
    l: float

    @staticmethod
    def new(terms, l: float, r: float, dt: float) -> "Group":
        den = l + 0.5 * r * dt
        return Group(terms, (l - 0.5 * r * dt) / den, dt / den, l)


def rcm_order(comp: Sequence[int], adj) -> List[int]:
    """conformal.rs:2128-2152: reverse Cuthill-McKee order of one connected
    component, from the first node of minimum degree, deterministic."""
    deg = {q: len(adj(q)) for q in comp}
    done: Set[int] = set()
    order: List[int] = []
    for s in sorted(comp, key=lambda q: (deg[q], q)):
        if s in done:
            continue
        done.add(s)
        queue = [s]
        head = 0
        while head < len(queue):
            q = queue[head]
            head += 1
            order.append(q)
            nb = sorted((k for k in adj(q) if k not in done), key=lambda k: (deg.get(k, 0), k))
            for k in nb:
                if k not in done:
                    done.add(k)
                    queue.append(k)
    order.reverse()
    return order


class Envelope:
    """Lower Cholesky factor in envelope storage (conformal.rs:2056-2124):
    row i holds columns first[i]..i."""

    def __init__(self, first: List[int], rows: List[List[float]]):
        self.first = first
        self.rows = rows

    @staticmethod
    def factor(a: Sequence[Sequence[Tuple[int, float]]]) -> Optional["Envelope"]:
        n = len(a)
        first = [min(min((e[0] for e in r), default=i), i) for i, r in enumerate(a)]
        rows = [[0.0] * (i - first[i] + 1) for i in range(n)]
        for i, r in enumerate(a):
            for j, v in r:
                if j <= i:
                    rows[i][j - first[i]] += v
        for i in range(n):
            for j in range(first[i], i + 1):
                lo = max(first[i], first[j])
                s = rows[i][j - first[i]]
                for k in range(lo, j):
                    s -= rows[i][k - first[i]] * rows[j][k - first[j]]
                if j == i:
                    if math.isnan(s) or s <= 0.0:
                        return None
                    rows[i][i - first[i]] = math.sqrt(s)
                else:
                    rows[i][j - first[i]] = s / rows[j][j - first[j]]
        return Envelope(first, rows)

    def solve(self, x: List[float]) -> None:
        n = len(x)
        for i in range(n):
            f, r = self.first[i], self.rows[i]
            s = x[i]
            for j in range(f, i):
                s -= r[j - f] * x[j]
            x[i] = s / r[i - f]
        for i in range(n - 1, -1, -1):
            f, r = self.first[i], self.rows[i]
            x[i] /= r[i - f]
            xi = x[i]
            for j in range(f, i):
                x[j] -= r[j - f] * xi


@dataclass
class Block:
    """Coupled reduced currents (conformal.rs:1969-2052): A g' = B g - dt c
    with A = M + (dt/2) R, B = M - (dt/2) R, M = P^T L P; rows in
    elimination order, columns local positions."""
    members: List[int]
    factor: Envelope
    m: List[List[Tuple[int, float]]]
    b: List[List[Tuple[int, float]]]
    dt: float
    mfactor: Optional[Envelope] = None  # the factor of M where it differs from A (lossy)

    @staticmethod
    def build(comp: Sequence[int], ml: Sequence[Dict[int, Tuple[float, float]]], first: int,
              dt: float) -> "Block":
        def adj(q):
            return sorted(k for k in ml[q] if k != q)
        order = rcm_order(comp, adj)
        pos = {q: i for i, q in enumerate(order)}

        def rows(f):
            out = []
            for q in order:
                row = [(pos[k], f(l, r)) for k, (l, r) in ml[q].items()]
                row.sort(key=lambda e: e[0])
                out.append(row)
            return out
        m = rows(lambda l, r: l)
        a = rows(lambda l, r: l + 0.5 * dt * r)
        b = rows(lambda l, r: l - 0.5 * dt * r)
        fac = Envelope.factor(a)
        if fac is None:
            raise ValueError("P^T L P is not positive definite")
        mfac = None if a == m else Envelope.factor(m)
        return Block([first + q for q in order], fac, m, b, dt, mfac)


@dataclass
class MergedFaces:
    """The reduced currents of a conformal geometry (conformal.rs:2313-2340):
    groups[:scalar] are merge groups with their own (da, db), the rest are
    solved in blocks; faces are the driven faces (axis, node, [(group,
    weight)])."""
    n: Tuple[int, int, int]
    dt: float
    groups: List[Group]
    scalar: int
    blocks: List[Block]
    faces: list
    pml: list = None

    def __len__(self):
        return len(self.groups)

    def desc_arrays(self) -> Dict[str, np.ndarray]:
        """Flat arrays of the device operator (ems_merged_desc)."""
        n = self.n
        g = self.groups
        ts = np.zeros(len(g) + 1, np.uint32)
        tax, tidx, tw = [], [], []
        for q, gr in enumerate(g):
            for (a, p, w) in gr.terms:
                tax.append(a)
                tidx.append(index(n, p))
                tw.append(w)
            ts[q + 1] = len(tax)
        bs = [0]
        mem, ef, eo, ev, brs, bc, bv = [], [], [0], [], [0], [], []
        mrs, mc, mv = [0], [], []
        cs, cr, cv = [0], [], []
        for blk in self.blocks:
            # column j of the factor: (k, L[k][j]) over the rows k > j whose
            # envelope holds j, in decreasing k (the order in which the
            # backward substitution applies them)
            nr = len(blk.members)
            cols = [[] for _ in range(nr)]
            for k in range(nr - 1, -1, -1):
                f = blk.factor.first[k]
                for j in range(f, k):
                    cols[j].append((k, blk.factor.rows[k][j - f]))
            for j in range(nr):
                for k, v in cols[j]:
                    cr.append(k)
                    cv.append(v)
                cs.append(len(cr))
            mem += blk.members
            for i in range(len(blk.members)):
                ef.append(blk.factor.first[i])
                ev += blk.factor.rows[i]
                eo.append(len(ev))
                for k, v in blk.b[i]:
                    bc.append(k)
                    bv.append(v)
                brs.append(len(bc))
                for k, v in blk.m[i]:
                    mc.append(k)
                    mv.append(v)
                mrs.append(len(mc))
            bs.append(len(mem))
        ds = [0]
        dax, didx, dg, dw = [], [], [], []
        for (a, p, w) in self.faces:
            dax.append(a)
            didx.append(index(n, p))
            for (q, x) in w:
                dg.append(q)
                dw.append(x)
            ds.append(len(dg))
        pml = self.pml or []
        per = [[] for _ in g]
        for q, t in enumerate(pml):
            for (gq, w) in t.targets:
                per[gq].append((q, w))
        gps, gpt, gpw = [0], [], []
        for lst in per:
            for q, w in lst:
                gpt.append(q)
                gpw.append(w)
            gps.append(len(gpt))
        return dict(
            pml_axis=np.array([t.other for t in pml], np.int32),
            pml_idx0=np.array([index(n, t.nodes[0]) for t in pml], np.uint64),
            pml_idx1=np.array([index(n, t.nodes[1]) for t in pml], np.uint64),
            pml_coef=np.array([[t.sign, t.b, t.a, t.k] for t in pml], np.float64).reshape(-1),
            pml_start=np.array(gps, np.uint32), pml_term=np.array(gpt, np.uint32),
            pml_w=np.array(gpw, np.float64),
            da=np.array([x.da for x in g], np.float64), db=np.array([x.db for x in g], np.float64),
            l=np.array([x.l for x in g], np.float64),
            mrow_start=np.array(mrs, np.uint32), mcol=np.array(mc, np.uint32),
            mval=np.array(mv, np.float64),
            col_start=np.array(cs, np.uint32), col_row=np.array(cr, np.uint32),
            col_val=np.array(cv, np.float64),
            term_start=ts, term_axis=np.array(tax, np.int32), term_idx=np.array(tidx, np.uint64),
            term_w=np.array(tw, np.float64),
            blk_start=np.array(bs, np.uint32), members=np.array(mem, np.uint32),
            env_first=np.array(ef, np.uint32), env_off=np.array(eo, np.uint64),
            env_val=np.array(ev, np.float64), brow_start=np.array(brs, np.uint32),
            bcol=np.array(bc, np.uint32), bval=np.array(bv, np.float64),
            face_axis=np.array(dax, np.int32), face_idx=np.array(didx, np.uint64),
            face_start=np.array(ds, np.uint32), face_group=np.array(dg, np.uint32),
            face_w=np.array(dw, np.float64))

    def _bind(self, engine, oid: int) -> None:
        self._engine, self._id = engine, oid

    def currents(self) -> Tuple[np.ndarray, np.ndarray]:
        """The reduced currents g and the last circulations, from the engine
        the operator runs in."""
        return self._engine._merged_state(self._id, len(self.groups))

    # -- host restatement of the per-step update (tests and energy) --

    def advance(self, cur: Sequence[float], c: Sequence[float]) -> List[float]:
        """MergedFaces::advance (conformal.rs:2569-2576) / Block::advance
        (:2037-2052): g^{n+1/2} of every reduced current."""
        out = [0.0] * len(self.groups)
        for q in range(self.scalar):
            g = self.groups[q]
            out[q] = float(_flush32(g.da * cur[q] - g.db * c[q]))
        for b in self.blocks:
            x = []
            for row, q in zip(b.b, b.members):
                s = -0.0
                for k, v in row:
                    s += v * cur[b.members[k]]
                x.append(s - b.dt * c[q])
            b.factor.solve(x)
            for q, v in zip(b.members, x):
                out[q] = float(_flush32(v))
        return out

    def energy(self, cur: Sequence[float], c: Sequence[float]) -> float:
        """Operator::energy of MergedFaces (conformal.rs:2684-2698):
        1/2 g^{n+1/2}^T M g^{n+3/2}."""
        nxt = self.advance(cur, c)
        w = 0.0
        for q in range(self.scalar):
            w += 0.5 * self.groups[q].l * cur[q] * nxt[q]
        for b in self.blocks:
            for row, q in zip(b.m, b.members):
                mn = -0.0
                for k, v in row:
                    mn += v * nxt[b.members[k]]
                w += 0.5 * cur[q] * mn
        return w


class Overlay:
    """Coefficient lanes (12, N) flat, with changed values held aside:
    co[lane, k] reads the change if there is one, else the lane."""

    def __init__(self, base: np.ndarray):
        self.base = base
        self.over: Dict[Tuple[int, int], np.float32] = {}

    def __getitem__(self, key):
        v = self.over.get(key)
        return self.base[key] if v is None else v

    def __setitem__(self, key, value):
        self.over[(int(key[0]), int(key[1]))] = np.float32(value)


def _flush32(x: float) -> np.float32:
    """flush(x as f32) (kernel.rs:30-33)."""
    v = np.float32(x)
    return np.float32(0.0) if abs(v) < np.finfo(np.float32).tiny else v


@dataclass
class Built:
    """What construct returns: the coefficient patch ((lane, flat index) ->
    f32 value), the operator and the counts."""
    patch: Dict[Tuple[int, int], np.float32]
    merged: MergedFaces
    interpolated: int = 0
    hosts: int = 0
    merged_faces: int = 0
    groups: int = 0
    floored: int = 0
    kept: int = 0
    faces: int = 0


def cell_stability(co: np.ndarray, grid, dt: float, q, fractions: Dict[Key, float]) -> float:
    """conformal.rs:1591-1700: the stability factor squared of one cell by
    itself (the element bound), the cell with lowest node q; co holds the
    edge-scaled lanes (12, N) flat, fractions the open areas of cut faces."""
    n = grid.nodes()
    if any(q[d] + 1 >= n[d] for d in range(3)):
        return 0.0
    delta = grid.delta
    faces = []
    for d in range(3):
        i = q[d]
        for o in range(2):
            p = list(q)
            p[d] += o
            k = index(n, p)
            da, db = float(co[DA + d, k]), float(co[DB + d, k])
            if db <= 0.0:
                continue
            l = dt * (1.0 + da) / (2.0 * db)
            f = fractions.get((d, k), 1.0)
            if not (l > 0.0 and f > 0.0):
                continue
            if o == 0:
                other = delta(d, i - 1) if i > 0 else None
            else:
                other = delta(d, i + 1) if i + 2 < n[d] else None
            share = 1.0 if other is None else delta(d, i) / (delta(d, i) + other)
            faces.append((d, tuple(p), share / (f * l)))
    inv_c: Dict[Key, float] = {}
    for d in range(3):
        b, c = others(d)
        for ob, oc in ((0, 0), (1, 0), (0, 1), (1, 1)):
            p = list(q)
            p[b] += ob
            p[c] += oc
            k = index(n, p)
            ca, cb = float(co[CA + d, k]), float(co[CB + d, k])
            if cb <= 0.0:
                continue
            cap = dt * (1.0 + ca) / (2.0 * cb)

            def side(a, i):
                v = []
                if i > 0:
                    v.append(delta(a, i - 1))
                if i + 1 < n[a]:
                    v.append(delta(a, i))
                return v
            total = -0.0
            for x in side(b, p[b]):
                for y in side(c, p[c]):
                    total += x * y
            own = delta(b, q[b]) * delta(c, q[c])
            if cap > 0.0 and total > 0.0:
                inv_c[(d, k)] = total / (own * cap)
    m = len(faces)
    rows = [[((e, index(n, en)), float(s)) for (e, en), s in zip(face_edges(d, p), SIGNS)]
            for (d, p, _) in faces]
    a = np.zeros((m, m))
    for i in range(m):
        for j in range(m):
            s = 0.0
            for ki, si in rows[i]:
                ic = inv_c.get(ki)
                if ic is None:
                    continue
                for kj, sj in rows[j]:
                    if ki == kj:
                        s += si * sj * ic
            a[i, j] = s * math.sqrt(faces[i][2] * faces[j][2])
    lam = float(np.linalg.eigvalsh(a).max()) if m else 0.0
    return 0.25 * dt * dt * lam


def _cell_incidence() -> Tuple[np.ndarray, list, list]:
    """The signed face-edge incidence of one cell (6 faces, 12 edges) in the
    order of cell_stability: faces (d, o) at q + o e_d, edges (d, ob, oc) at
    q + ob e_b + oc e_c."""
    faces = [(d, o) for d in range(3) for o in range(2)]
    edges = [(d, ob, oc) for d in range(3) for ob, oc in ((0, 0), (1, 0), (0, 1), (1, 1))]
    K = np.zeros((6, 12))
    for fi, (d, o) in enumerate(faces):
        p = [0, 0, 0]
        p[d] += o
        for (e, en), sgn in zip(face_edges(d, p), SIGNS):
            b, c = others(e)
            K[fi, edges.index((e, en[b], en[c]))] += sgn
    return K, faces, edges


def cell_bounds(co, grid, dt: float, cells: np.ndarray, fractions: Dict[Key, float]) -> np.ndarray:
    """cell_stability of many cells at once (cells: (m, 3) lowest nodes),
    as the largest eigenvalue of Y^1/2 K C^-1 K^T Y^1/2 per cell with the
    cell's fixed face-edge incidence K; a face or edge without a share
    enters with zero (its row and column vanish, which leaves the largest
    eigenvalue of the positive semidefinite matrix as it is)."""
    n = grid.nodes()
    N = n[0] * n[1] * n[2]
    m = len(cells)
    out = np.zeros(m)
    if m == 0:
        return out
    inside = np.all(cells + 1 < np.array(n), axis=1)
    K, faces, edges = _cell_incidence()
    dl = [np.diff(np.asarray(grid.lines[a], np.float64)) for a in range(3)]
    base = co.base if isinstance(co, Overlay) else co
    over = co.over if isinstance(co, Overlay) else {}
    okeys = np.array(sorted(l * N + k for (l, k) in over), np.int64)
    ovals = np.array([over[(int(x // N), int(x % N))] for x in okeys], np.float32)
    fkeys = np.array(sorted(a * N + k for (a, k) in fractions), np.int64)
    fvals = np.array([fractions[(int(x // N), int(x % N))] for x in fkeys], np.float64)

    def lane(l, k):
        v = base[l][k].astype(np.float64)
        if okeys.size:
            key = l * N + k
            pos = np.clip(np.searchsorted(okeys, key), 0, okeys.size - 1)
            hit = okeys[pos] == key
            v[hit] = ovals[pos[hit]]
        return v
    q = np.where(inside[:, None], cells, 0).astype(np.int64)
    Y = np.zeros((m, 6))
    for fi, (d, o) in enumerate(faces):
        p = q.copy()
        p[:, d] += o
        k = (p[:, 0] * n[1] + p[:, 1]) * n[2] + p[:, 2]
        da, db = lane(DA + d, k), lane(DB + d, k)
        with np.errstate(divide="ignore", invalid="ignore"):
            l = np.where(db > 0.0, dt * (1.0 + da) / (2.0 * db), 0.0)
        f = np.ones(m)
        if fkeys.size:
            key = d * N + k
            pos = np.clip(np.searchsorted(fkeys, key), 0, fkeys.size - 1)
            hit = fkeys[pos] == key
            f[hit] = fvals[pos[hit]]
        i = q[:, d]
        di = dl[d][np.minimum(i, len(dl[d]) - 1)]
        if o == 0:
            has = i > 0
            other = dl[d][np.maximum(i - 1, 0)]
        else:
            has = i + 2 < n[d]
            other = dl[d][np.minimum(i + 1, len(dl[d]) - 1)]
        share = np.where(has, di / (di + other), 1.0)
        valid = (db > 0.0) & (l > 0.0) & (f > 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            Y[:, fi] = np.where(valid, share / (f * l), 0.0)
    IC = np.zeros((m, 12))
    for ei, (d, ob, oc) in enumerate(edges):
        b, c = others(d)
        p = q.copy()
        p[:, b] += ob
        p[:, c] += oc
        k = (p[:, 0] * n[1] + p[:, 1]) * n[2] + p[:, 2]
        ca, cb = lane(CA + d, k), lane(CB + d, k)
        with np.errstate(divide="ignore", invalid="ignore"):
            cap = np.where(cb > 0.0, dt * (1.0 + ca) / (2.0 * cb), 0.0)

        def side(a, i):
            lo = np.where(i > 0, dl[a][np.maximum(i - 1, 0)], 0.0)
            hi = np.where(i + 1 < n[a], dl[a][np.minimum(i, len(dl[a]) - 1)], 0.0)
            return lo, hi
        bl, bh = side(b, p[:, b])
        cl, ch = side(c, p[:, c])
        total = (bl + bh) * (cl + ch)
        own = dl[b][np.minimum(q[:, b], len(dl[b]) - 1)] * dl[c][np.minimum(q[:, c], len(dl[c]) - 1)]
        valid = (cb > 0.0) & (cap > 0.0) & (total > 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            IC[:, ei] = np.where(valid, total / (own * cap), 0.0)
    A = np.einsum("fe,me,ge->mfg", K, IC, K)
    sy = np.sqrt(Y)
    A *= sy[:, :, None] * sy[:, None, :]
    lam = np.linalg.eigvalsh(A)[:, -1]
    out[:] = np.where(inside, 0.25 * dt * dt * lam, 0.0)
    return out


def stable_faces(faces, co, grid, dt: float, max_stability: float,
                 keep_step: Optional[float], absorbing) -> Tuple[Set[Key], bool]:
    """ConformalPec::stable_faces (conformal.rs:967-1015): the cut faces
    both of whose cells pass the cell-wise bound at the time step (every
    cell evaluated once, all at once)."""
    n = grid.nodes()
    fractions = {(f.axis, index(n, f.node)): f.fraction for f in faces}
    limit = max_stability * max_stability * (1.0 if keep_step is None else (dt / keep_step) *
                                              (dt / keep_step))
    plan = []
    want = set()
    every = True
    for f in faces:
        a = f.axis
        if f.fraction <= 0.0:
            continue
        if touches_layer(absorbing, a, f.node, n):
            every = False
            continue
        above = None if f.node[a] + 1 >= n[a] else tuple(f.node)
        below = None
        if f.node[a] > 0:
            q = list(f.node)
            q[a] -= 1
            below = tuple(q)
        for c in (above, below):
            if c is not None:
                want.add(c)
        plan.append((f, above, below))
    cells = sorted(want)
    bound = cell_bounds(co, grid, dt, np.array(cells, np.int64).reshape(-1, 3), fractions)
    ok = {c: b <= limit for c, b in zip(cells, bound)}
    out: Set[Key] = set()
    for f, above, below in plan:
        if (above is None or ok[above]) and (below is None or ok[below]):
            out.add((f.axis, index(n, f.node)))
        else:
            every = False
    return out, every


def cpml_h_at(cpml_h, a: int, p, n):
    """ConformalPec::cpml_h_at (conformal.rs:576-598): (d, coefficients) of
    each layer whose H positions hold the face along its derivative
    direction d != a; cpml_h lists (face index, cells, H profile)."""
    out = []
    for face, cells, h in cpml_h:
        d = face // 2
        if d == a:
            continue
        np_, k = n[d], p[d]
        if face % 2 == 1:
            inner = np_ - 1 - cells
            pos = k - inner if k >= inner else None
        else:
            pos = k if k < cells else None
        if pos is not None and pos < len(h):
            out.append((d, h[pos]))
    return out


@dataclass
class PmlTerm:
    """The CPML correction of one driven face in a layer, one derivative
    direction (conformal.rs:2398-2412): psi <- b psi + a T, circulation
    += (k T + psi) w for every (group, w) target, T = sign (u_other(hi) -
    u_other(p))."""
    targets: list
    other: int
    nodes: tuple
    sign: float
    b: float
    a: float
    k: float


def construct(faces, centroids, edges, walls, co: np.ndarray, grid, dt: float, tau: float,
              kept: Set[Key], deviation: float, absorbing, cpml_h=()) -> Built:
    """ConformalPec::construct (conformal.rs:1041-1589) without thin sheets:
    `co` is the lane array (12, N) flat with the edge scaling applied;
    cpml_h the Cartesian layers' H grading (set_cpml)."""
    n = grid.nodes()
    idx = lambda p: index(n, p)  # noqa: E731
    patch: Dict[Tuple[int, int], np.float32] = {}
    face_info: Dict[Key, Tuple[float, Optional[Tuple[float, float]]]] = {}
    for f, c in zip(faces, centroids):
        face_info[(f.axis, idx(f.node))] = (f.fraction, c)
    edge_frac = {(e.axis, idx(e.node)): e.fraction for e in edges}

    def fa_of(a, p):
        x = face_info.get((a, idx(p)))
        return 1.0 if x is None else x[0]
    wall_keys = {(a, idx(p)) for a, p in walls
                 if all(p[d] < n[d] for d in range(3)) and p[a] + 1 < n[a]}

    def is_open(a, p):
        k = idx(p)
        return co[CB + a, k] > 0.0 and (a, k) not in wall_keys

    def fe_of(a, p):
        if is_open(a, p):
            return edge_frac.get((a, idx(p)), 1.0)
        return 0.0

    live_cache: Dict[Key, bool] = {}

    def live(a, p):
        k = (a, idx(p))
        v = live_cache.get(k)
        if v is None:
            v = any(is_open(e, q) for e, q in face_edges(a, p))
            live_cache[k] = v
        return v

    def lr_full(a, k):
        da, db = float(co[DA + a, k]), float(co[DB + a, k])
        if db > 0.0:
            return dt * (1.0 + da) / (2.0 * db), (1.0 - da) / db
        return 0.0, 0.0

    def pos(a, p):
        x = face_info.get((a, idx(p)))
        t = x[1] if (x is not None and x[1] is not None) else (0.5, 0.5)
        b, c = others(a)

        def at(d, tt):
            ln = grid.lines[d]
            i = p[d]
            return float(ln[i]) + tt * (float(ln[i + 1]) - float(ln[i]))
        return (at(b, t[0]), at(c, t[1]))

    def in_abs(a, p):
        return touches_layer(absorbing, a, p, n)

    small = [f for f in faces
             if f.fraction < tau and live(f.axis, f.node) and (f.axis, idx(f.node)) not in kept]

    def stencil(f, uf: Groups):
        a = f.axis
        k0 = (a, idx(f.node))
        info = face_info.get(k0)
        if info is None or info[1] is None:
            return None
        if in_abs(a, f.node):
            return None

        def free(p):
            k = (a, idx(p))
            return (live(a, p) and fa_of(a, p) >= tau and co[DB + a, k[1]] > 0.0
                    and not uf.contains(k) and not in_abs(a, p))
        host = None
        for e, q, nb in in_plane_neighbours(a, f.node, n):
            fe = fe_of(e, q)
            if fe <= 0.0 or not free(nb):
                continue
            score = (fe, fa_of(a, nb))
            if host is None or score > host[0]:
                host = (score, nb)
        if host is None:
            return None
        host = host[1]
        b, c = others(a)
        scale = (grid.delta(b, f.node[b]), grid.delta(c, f.node[c]))
        xs = pos(a, f.node)

        def rel(p):
            x = pos(a, p)
            return ((x[0] - xs[0]) / scale[0], (x[1] - xs[1]) / scale[1])
        dh = rel(host)
        pts = [(host, dh)]
        for qi, d in enumerate((b, c)):
            pick = None
            for e, en, nb in in_plane_neighbours(a, host, n):
                if nb[d] == host[d] or nb == tuple(f.node) or fe_of(e, en) <= 0.0 or not free(nb):
                    continue
                r = rel(nb)
                toward = abs(r[qi]) < abs(dh[qi])
                if pick is None or (pick[0] and not toward):
                    pick = (toward, nb, r)
            if pick is not None:
                pts.append((pick[1], pick[2]))
            elif abs(dh[qi]) > 1e-9:
                return None
        w = interpolation_weights([p[1] for p in pts])
        if w is None:
            return None
        ls = grid.dual_delta(a, f.node[a])  # centroid_dual_length, Cartesian
        return [((a, idx(p[0])), p[0], wq * ls / grid.dual_delta(a, p[0][a]))
                for p, wq in zip(pts, w)]

    uf = Groups()
    merge_seeds = []
    while True:
        stencils = []
        failing = []
        for f in small:
            k = (f.axis, idx(f.node))
            if uf.contains(k):
                continue
            s = stencil(f, uf)
            if s is not None:
                stencils.append((k, f.axis, tuple(f.node), s))
            else:
                failing.append(f)
        if not failing:
            break
        for f in failing:
            uf.add((f.axis, idx(f.node)), tuple(f.node), f.fraction)
        merge_seeds.extend(failing)
        for _ in range(max(len(merge_seeds), 1)):
            changed = False
            for f in merge_seeds:
                k0 = (f.axis, idx(f.node))
                root = uf.find(k0)
                if uf.area[root] >= tau or uf.members[root][0] != k0:
                    continue
                best = None
                for m in list(uf.members[root]):
                    p = uf.node[m]
                    for e, q, nb in in_plane_neighbours(f.axis, p, n):
                        fe = fe_of(e, q)
                        if fe <= 0.0 or not live(f.axis, nb):
                            continue
                        kn = (f.axis, idx(nb))
                        if uf.contains(kn) and uf.find(kn) == root:
                            continue
                        score = (fe, fa_of(f.axis, nb))
                        if best is None or score > best[0]:
                            best = (score, kn, nb)
                if best is None:
                    continue
                (_, fa_nb), kn, nb = best
                if not uf.contains(kn):
                    uf.add(kn, nb, fa_nb)
                uf.union(root, kn)
                changed = True
            if not changed:
                break

    grouped: Set[Key] = set()
    groups: List[Group] = []
    driven = []

    def freeze(k):
        patch[(DA + k[0], k[1])] = np.float32(1.0)
        patch[(DB + k[0], k[1])] = np.float32(0.0)
        grouped.add(k)
    out = Built(patch, None)
    soft = []
    for _, members in uf.groups():
        l = r = 0.0
        terms: Dict[Key, list] = {}
        fl = []
        for k in members:
            a = k[0]
            p = uf.node[k]
            lf, rf = lr_full(a, k[1])
            fa = fa_of(a, p)
            l += fa * lf
            r += fa * rf
            add_terms(terms, a, p, 1.0, n)
            fl.append((a, p, fa, lf, rf))
            freeze(k)
        out.merged_faces += len(members)
        if l <= 0.0:
            continue
        if deviation > 0.0:
            soft.append((sorted_terms(terms, n), l, r, fl))
            continue
        g = len(groups)
        groups.append(Group.new(sorted_terms(terms, n), l, r, dt))
        driven.extend((a, p, [(g, 1.0)]) for (a, p, *_rest) in fl)
    out.groups = len(groups) + len(soft)
    scalar = len(groups)

    hosts: Dict[Key, list] = {}
    for _, a, p, st in stencils:
        for kh, ph, w in st:
            h = hosts.get(kh)
            if h is None:
                terms = {}
                add_terms(terms, a, ph, 1.0, n)
                h = [a, ph, terms]
                hosts[kh] = h
            add_terms(h[2], a, p, w, n)
    host_keys = sorted(hosts)
    local = {k: q for q, k in enumerate(host_keys)}
    ml: List[Dict[int, List[float]]] = [dict() for _ in host_keys]
    for q, kh in enumerate(host_keys):
        a, ph, _t = hosts[kh]
        lh, rh = lr_full(a, kh[1])
        fh = fa_of(a, ph)
        ml[q][q] = [fh * lh, fh * rh]
    link = Groups()
    for q in range(len(host_keys)):
        link.add((0, q), (0, 0, 0), 0.0)
    nh = len(host_keys)
    for ns, (k, a, p, st) in enumerate(stencils):
        lf, rf = lr_full(a, k[1])
        fa = fa_of(a, p) * grid.dual_delta(a, p[a]) / grid.dual_delta(a, p[a])
        for kj, _pj, wj in st:
            j = local[kj]
            for kk, _pk, wk in st:
                kq = local[kk]
                e = ml[j].setdefault(kq, [0.0, 0.0])
                e[0] += wj * wk * fa * lf
                e[1] += wj * wk * fa * rf
            link.union((0, local[st[0][0]]), (0, j))
        if deviation > 0.0:
            qd = nh + ns
            row = {qd: [(fa + deviation) * lf, fa * rf]}
            for kj, _pj, wj in st:
                j = local[kj]
                cross = (wj * fa * lf, wj * fa * rf)
                e = row.setdefault(j, [0.0, 0.0])
                e[0] += cross[0]
                e[1] += cross[1]
                e = ml[j].setdefault(qd, [0.0, 0.0])
                e[0] += cross[0]
                e[1] += cross[1]
            ml.append(row)
            link.add((0, qd), (0, 0, 0), 0.0)
            link.union((0, local[st[0][0]]), (0, qd))
    first_host = len(groups)
    for kh in host_keys:
        a, ph, terms = hosts[kh]
        q = local[kh]
        g = first_host + q
        l, r = ml[q][q]
        groups.append(Group.new(sorted_terms(terms, n), l, r, dt))
        driven.append((a, ph, [(g, 1.0)]))
        freeze(kh)
    with_deviation = len(ml) > nh
    if with_deviation:
        for ns, (_k, a, p, _st) in enumerate(stencils):
            terms = {}
            add_terms(terms, a, p, 1.0, n)
            l, r = ml[nh + ns][nh + ns]
            groups.append(Group.new(sorted_terms(terms, n), l, r, dt))
    for (sterms, sl, sr, fl) in soft:
        qg = len(ml)
        g = first_host + qg
        ml.append({qg: [sl, sr]})
        link.add((0, qg), (0, 0, 0), 0.0)
        groups.append(Group.new(sterms, sl, sr, dt))
        for (a, p, fa, lf, rf) in fl:
            w = [(g, 1.0)]
            if fa < tau:
                qd = len(ml)
                cross = [fa * lf, fa * rf]
                ml.append({qd: [(fa + deviation) * lf, fa * rf], qg: list(cross)})
                ml[qg][qd] = list(cross)
                link.add((0, qd), (0, 0, 0), 0.0)
                link.union((0, qg), (0, qd))
                terms = {}
                add_terms(terms, a, p, 1.0, n)
                l, r = ml[qd][qd]
                groups.append(Group.new(sorted_terms(terms, n), l, r, dt))
                w.append((first_host + qd, 1.0))
            driven.append((a, p, w))
    comps: List[List[int]] = []
    seen: Dict = {}
    for q in range(len(ml)):
        rt = link.find((0, q))
        if rt not in seen:
            seen[rt] = len(comps)
            comps.append([])
        comps[seen[rt]].append(q)
    mlt = [{k: (v[0], v[1]) for k, v in row.items()} for row in ml]
    blocks = [Block.build(comp, mlt, first_host, dt) for comp in comps]
    out.hosts = nh
    out.interpolated = len(stencils)
    for ns, (k, a, p, st) in enumerate(stencils):
        w = [(first_host + local[s[0]], s[2]) for s in st]
        if with_deviation:
            w.append((first_host + nh + ns, 1.0))
        driven.append((a, p, w))
        freeze(k)
    # the CPML's H-side correction of the driven faces in the layers, on
    # every reduced current driving the face (conformal.rs:1502-1539)
    pml = []
    for (a, p, targets) in driven:
        if not in_abs(a, p):
            continue
        for d, c in cpml_h_at(cpml_h, a, p, n):
            sign, other = (1.0, (a + 2) % 3) if d == (a + 1) % 3 else (-1.0, (a + 1) % 3)
            hi = list(p)
            hi[d] += 1
            pml.append(PmlTerm(list(targets), other, (tuple(hi), tuple(p)), sign,
                               float(c["b"]), float(c["a"]), float(c["direct"])))

    for f in faces:
        k = idx(f.node)
        if (f.axis, k) in grouped or not live(f.axis, f.node):
            continue
        fa = f.fraction
        if fa < tau:
            if (f.axis, k) in kept:
                out.kept += 1
            else:
                fa = max(tau, 1e-6)
                out.floored += 1
        if fa < 1.0:
            key = (DB + f.axis, k)
            db = patch.get(key, co[DB + f.axis, k])
            patch[key] = np.float32(float(db) / fa)
        out.faces += 1
    out.merged = MergedFaces(tuple(n), dt, groups, scalar, blocks, driven, pml)
    return out


_dp, _u32p, _u64p, _i32p = (C.POINTER(C.c_double), C.POINTER(C.c_uint32),
                            C.POINTER(C.c_uint64), C.POINTER(C.c_int32))


class MergedDesc(C.Structure):
    """ems_merged_desc (include/ems_gpu.h)."""
    _fields_ = [("n_groups", C.c_int32), ("n_scalar", C.c_int32), ("n_blocks", C.c_int32),
                ("n_faces", C.c_int32), ("dt", C.c_double),
                ("da", _dp), ("db", _dp), ("l", _dp),
                ("term_start", _u32p), ("term_axis", _i32p), ("term_idx", _u64p), ("term_w", _dp),
                ("blk_start", _u32p), ("members", _u32p), ("env_first", _u32p),
                ("env_off", _u64p), ("env_val", _dp),
                ("brow_start", _u32p), ("bcol", _u32p), ("bval", _dp),
                ("mrow_start", _u32p), ("mcol", _u32p), ("mval", _dp),
                ("col_start", _u32p), ("col_row", _u32p), ("col_val", _dp),
                ("n_pml", C.c_int32), ("pml_axis", _i32p), ("pml_idx0", _u64p),
                ("pml_idx1", _u64p), ("pml_coef", _dp), ("pml_start", _u32p),
                ("pml_term", _u32p), ("pml_w", _dp),
                ("face_axis", _i32p), ("face_idx", _u64p), ("face_start", _u32p),
                ("face_group", _u32p), ("face_w", _dp)]


def make_desc(op: MergedFaces):
    """The descriptor of `op` and the arrays it points into (keep them
    alive while the descriptor is used)."""
    arr = {k: np.ascontiguousarray(v) for k, v in op.desc_arrays().items()}
    d = MergedDesc()
    d.n_groups, d.n_scalar = len(op.groups), op.scalar
    d.n_blocks, d.n_faces, d.dt = len(op.blocks), len(op.faces), op.dt
    d.n_pml = len(op.pml or [])
    types = {"da": _dp, "db": _dp, "l": _dp, "term_w": _dp, "env_val": _dp, "bval": _dp,
             "mval": _dp, "col_val": _dp, "face_w": _dp, "term_axis": _i32p, "face_axis": _i32p,
             "term_idx": _u64p, "env_off": _u64p, "face_idx": _u64p, "pml_axis": _i32p,
             "pml_idx0": _u64p, "pml_idx1": _u64p, "pml_coef": _dp, "pml_w": _dp}
    for k, v in arr.items():
        setattr(d, k, v.ctypes.data_as(types.get(k, _u32p)))
    return d, arr


def grid_bound(co, grid, dt: float, limit: float, chunk: int = 1 << 18) -> Optional[float]:
    """grid_bound (conformal.rs:1707-1731): the largest cell-wise bound over
    every cell of the grid on the round's coefficients, or None as soon as
    a cell reaches the limit (taken a part in 1e9 lower). Each cell enters
    with its own largest eigenvalue (the reference first tries the row
    sums, an upper bound, and reports those where they pass)."""
    n = grid.nodes()
    lim = limit * (1.0 - 1e-9)
    c = [n[a] - 1 for a in range(3)]
    total = c[0] * c[1] * c[2]
    worst = 0.0
    for s0 in range(0, total, chunk):
        ids = np.arange(s0, min(s0 + chunk, total), dtype=np.int64)
        cells = np.stack([ids // (c[1] * c[2]), (ids // c[2]) % c[1], ids % c[2]], axis=1)
        v = cell_bounds(co, grid, dt, cells, {})
        if not np.all(v < lim):
            return None
        worst = max(worst, float(v.max()))
    return worst


_fp = C.POINTER(C.c_float)


class StabilityDesc(C.Structure):
    """ems_stability_desc (include/ems_host.h)."""
    _fields_ = [("nodes", C.c_int32 * 3), ("cb", _fp * 3), ("db", _fp * 3),
                ("n_groups", C.c_int32), ("g_dt_over_l", _dp), ("g_term_start", _u32p),
                ("g_term_axis", _i32p), ("g_term_idx", _u64p), ("g_term_w", _dp),
                ("n_blocks", C.c_int32), ("blk_start", _u32p), ("blk_dt", _dp),
                ("row_term_start", _u32p), ("row_term_axis", _i32p), ("row_term_idx", _u64p),
                ("row_term_w", _dp), ("env_first", _u32p), ("env_off", _u64p),
                ("env_val", _dp), ("n_seeds", C.c_int32), ("seed_axis", _i32p),
                ("seed_idx", _u64p), ("iterations", C.c_int32)]


def stability(lanes, n, op: Optional[MergedFaces], seeds, iterations: int) -> float:
    """The stability estimate s of the round's coefficients `lanes` (the 12
    lanes (12, N) flat, the patch applied) with the reduced currents of
    `op` (ems_conformal_stability, the reference's power iteration)."""
    from . import _native as nat
    groups = op.groups[:op.scalar] if op is not None else []
    blocks = op.blocks if op is not None else []

    def terms_arrays(lists):
        st, ax, ix, w = [0], [], [], []
        for terms in lists:
            for a, p, x in terms:
                ax.append(a)
                ix.append(index(n, p))
                w.append(x)
            st.append(len(ax))
        return (np.array(st, np.uint32), np.array(ax, np.int32), np.array(ix, np.uint64),
                np.array(w, np.float64))
    keep = []
    d = StabilityDesc()
    d.nodes[:] = list(n)
    for a in range(3):
        cb = np.ascontiguousarray(lanes[CB + a], np.float32)
        db = np.ascontiguousarray(lanes[DB + a], np.float32)
        keep += [cb, db]
        d.cb[a] = cb.ctypes.data_as(_fp)
        d.db[a] = db.ctypes.data_as(_fp)
    gs, ga, gi, gw = terms_arrays([g.terms for g in groups])
    gd = np.array([op.dt / g.l for g in groups], np.float64)
    rows, bs, bdt, ef, eo, ev = [], [0], [], [], [0], []
    for b in blocks:
        fac = b.mfactor if b.mfactor is not None else b.factor
        for i, q in enumerate(b.members):
            rows.append(op.groups[q].terms)
            ef.append(fac.first[i])
            ev += fac.rows[i]
            eo.append(len(ev))
        bs.append(len(rows))
        bdt.append(b.dt)
    rs, ra, ri, rw = terms_arrays(rows)
    sa = np.array([a for a, _ in seeds], np.int32)
    si = np.array([k for _, k in seeds], np.uint64)
    arrs = dict(g_dt_over_l=gd, g_term_start=gs, g_term_axis=ga, g_term_idx=gi, g_term_w=gw,
                blk_start=np.array(bs, np.uint32), blk_dt=np.array(bdt, np.float64),
                row_term_start=rs, row_term_axis=ra, row_term_idx=ri, row_term_w=rw,
                env_first=np.array(ef, np.uint32), env_off=np.array(eo, np.uint64),
                env_val=np.array(ev, np.float64), seed_axis=sa, seed_idx=si)
    types = {"g_dt_over_l": _dp, "g_term_w": _dp, "blk_dt": _dp, "row_term_w": _dp,
             "env_val": _dp, "g_term_axis": _i32p, "row_term_axis": _i32p, "seed_axis": _i32p,
             "g_term_idx": _u64p, "row_term_idx": _u64p, "env_off": _u64p, "seed_idx": _u64p}
    for k, v in arrs.items():
        v = np.ascontiguousarray(v)
        keep.append(v)
        setattr(d, k, v.ctypes.data_as(types.get(k, _u32p)))
    d.n_groups, d.n_blocks, d.n_seeds = len(groups), len(blocks), len(seeds)
    d.iterations = int(iterations)
    r = nat.ems_conformal_stability(C.byref(d))
    if r < 0:
        raise ValueError("stability estimate: invalid tables")
    return r
