"""Conformal (sub-cell) PEC (src/engine/conformal.rs).

The Dey-Mittra scheme in the integral form of the engine is a change of
coefficients (conformal.rs:1-30): a cut E edge with open length fraction
f_e gets cb -> f_e cb, a cut face with open area fraction f_A gets
db -> db / f_A; ca and da stay. The patches are written into the
coefficient lanes, so the device step kernels run them unchanged.

A face below the merge threshold loses its own current (conformal.rs:
34-95): interpolated from neighbouring faces at its open-area centroid, or
merged with a neighbour where no stencil exists, each with a deviation
current; the reduced currents are updated after every H half-step by the
MergedFaces operator (ems_amd/merged.py builds it, the device runs it). A
small face whose two cells pass the cell-wise stability bound keeps its
current at its true area (keep_stable); one without an open neighbour is
raised to the threshold.

Arithmetic as ConformalPec::apply_known (conformal.rs:770-954): the edge
scaling (cb as f64 * f_e) as f32 in edge order, the stable faces from the
scaled lanes, then rounds: construct at the threshold, the patch on top of
the scaled lanes (changed, :2704-2715), the stability check (the cell-wise
bound where every current is a grid face's, else the power-iteration
estimate, ems_conformal_stability), the threshold raised by 1.25 while the
check fails and rounds remain.

Inside CPML layers (set_cpml) small faces are merged and the merge groups
apply the layer's H-side correction themselves. Scope: Cartesian grids;
no thin PEC sheets; no cylindrical PML.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from . import merged as M
from ._native import EmsError

CB, DB = 3, 9  # first cb and db lanes (x, y, z follow)


@dataclass
class ConformalEdge:
    """An E edge partly inside metal (conformal.rs:259-267): open length
    fraction in [0, 1] (0: PEC)."""
    axis: int
    node: Tuple[int, int, int]
    fraction: float


@dataclass
class ConformalFace:
    """A cell face (H edge) partly inside metal (conformal.rs:271-279):
    open area fraction in [0, 1]; node is its lowest corner."""
    axis: int
    node: Tuple[int, int, int]
    fraction: float


@dataclass
class ConformalOptions:
    """ConformalOptions (conformal.rs:283-331): iterations of the stability
    estimate per round (0 skips the check), rounds of threshold raising."""
    merge_below: float = 0.5
    max_stability: float = 0.999
    iterations: int = 300
    max_rounds: int = 4
    deviation_area: float = 0.5
    keep_stable: bool = True
    keep_step: Optional[float] = None


@dataclass
class ConformalReport:
    """ConformalReport (conformal.rs:350-424), the fields that apply here,
    plus pec_edges (edges made PEC by a zero fraction)."""
    edges: int = 0
    faces: int = 0
    interpolated: int = 0
    hosts: int = 0
    merged: int = 0
    groups: int = 0
    floored: int = 0
    kept: int = 0
    bounded: bool = False
    merge_below: float = 0.5
    stability: float = float("nan")
    proven: bool = False
    verified: bool = True
    pec_edges: int = 0

    def warning(self) -> Optional[str]:
        """conformal.rs:425-434."""
        if self.verified:
            return None
        return (f"conformal PEC: estimated stability factor {self.stability:.4f} (> 1 is "
                f"unstable) with merge threshold {self.merge_below:.3f}; reduce the time step "
                "factor")


class ConformalPec:
    """Cut edges and faces of curved metal (ConformalPec, conformal.rs:440-755)."""

    def __init__(self, options: ConformalOptions = None):
        self.options = options or ConformalOptions()
        self.edges: List[ConformalEdge] = []
        self.faces: List[ConformalFace] = []
        self.centroids: List[Optional[Tuple[float, float]]] = []
        self.walls: List[Tuple[int, Tuple[int, int, int]]] = []
        self.absorbing = [0] * 6
        self.cpml_h: list = []

    def push_edge(self, e: ConformalEdge) -> None:
        self.edges.append(e)

    def push_face(self, f: ConformalFace) -> None:
        """A cut face without a centroid: small, it is merged, not
        interpolated (conformal.rs:518-524)."""
        self.faces.append(f)
        self.centroids.append(None)

    def push_face_with_centroid(self, f: ConformalFace, centroid: Sequence[float]) -> None:
        """A cut face with the centroid of its open area, [t_b, t_c] in face
        sides along (b, c) = (next, prev) of its axis (conformal.rs:527-533)."""
        self.faces.append(f)
        self.centroids.append((float(centroid[0]), float(centroid[1])))

    def push_wall(self, axis: int, node) -> None:
        """An edge closed to interpolation and merging, not PEC
        (conformal.rs:509-515)."""
        self.walls.append((int(axis), tuple(int(x) for x in node)))

    def set_cpml(self, cpml) -> None:
        """ConformalPec::set_cpml (conformal.rs:555-571): the CPML's layers
        become absorbing layers (small faces there are merged) and their
        H-side grading is kept for the merge groups' own correction."""
        for face, params, _nf in cpml.faces:
            f = int(face)
            self.set_absorbing_layers([(f, params.cells)])
            h = cpml.h_profile(face)
            if h is not None:
                self.cpml_h = [l for l in self.cpml_h if l[0] != f]
                self.cpml_h.append((f, params.cells, h))

    def set_absorbing_layers(self, layers) -> None:
        """(face index 0..5 = x-, x+, y-, y+, z-, z+, cells) pairs
        (conformal.rs:537-541)."""
        for f, cells in layers:
            self.absorbing[f] = max(self.absorbing[f], int(cells))

    def validate(self, n) -> None:
        """conformal.rs:1017-1036: fractions in [0, 1], edges and faces
        inside the grid, centroids in [0, 1]^2."""
        for e in self.edges:
            a = e.axis
            if not (0.0 <= e.fraction <= 1.0) or any(e.node[d] >= n[d] for d in range(3)) \
                    or e.node[a] + 1 >= n[a]:
                raise EmsError(nat.EMS_E_INVALID, f"invalid conformal edge {e}")
        for f, c in zip(self.faces, self.centroids):
            bad_c = c is not None and not all(0.0 <= t <= 1.0 for t in c)
            if not (0.0 <= f.fraction <= 1.0) or not M.face_exists(f.axis, f.node, n) or bad_c:
                raise EmsError(nat.EMS_E_INVALID, f"invalid conformal face {f}")

    def build(self, lanes: np.ndarray, grid, dt: float):
        """The changed lanes (index -> new dense lane), the report and the
        MergedFaces operator (None without reduced currents) for the
        coefficient lanes `lanes` (12, nx, ny, nz) of a grid and time step."""
        n = tuple(lanes.shape[1:])
        self.validate(n)
        opt = self.options
        flat = lanes.reshape(12, -1)
        co = M.Overlay(flat)
        rep = ConformalReport(edges=len(self.edges))
        for e in self.edges:
            k = M.index(n, e.node)
            co[CB + e.axis, k] = np.float32(float(co[CB + e.axis, k]) * e.fraction)
            if e.fraction == 0.0:
                rep.pec_edges += 1
        if opt.keep_stable:
            at = opt.keep_step if (opt.keep_step is not None and 0.0 < opt.keep_step < dt) else None
            kept, bounded = M.stable_faces(self.faces, co, grid, dt, opt.max_stability, at,
                                           self.absorbing)
        else:
            kept, bounded = set(), False
        seeds = [(a, M.index(n, p)) for f in self.faces for a, p in M.face_edges(f.axis, f.node)]
        tau = min(max(opt.merge_below, 0.0), 1.0)
        rounds = max(opt.max_rounds, 1)
        limit = opt.max_stability * opt.max_stability
        for rnd in range(rounds):
            b = M.construct(self.faces, self.centroids, self.edges, self.walls, co, grid, dt,
                            tau, kept, min(max(opt.deviation_area, 0.0), 1.0), self.absorbing,
                            self.cpml_h)
            r = ConformalReport(edges=rep.edges, pec_edges=rep.pec_edges, faces=b.faces,
                                interpolated=b.interpolated, hosts=b.hosts,
                                merged=b.merged_faces, groups=b.groups, floored=b.floored,
                                kept=b.kept, bounded=bounded, merge_below=tau)
            cr = M.Overlay(flat)
            cr.over = dict(co.over)
            for (lane, k), v in b.patch.items():
                cr[lane, k] = v
            last = rnd + 1 == rounds or tau >= 1.0
            op = b.merged if len(b.merged) else None
            if opt.iterations > 0:
                # apply_known (conformal.rs:850-900): the cell-wise bound where
                # every current is a grid face's, else the power iteration
                bound = M.grid_bound(cr, grid, dt, limit) if op is None else None
                if bound is not None:
                    r.stability, r.verified, r.proven = math.sqrt(bound), True, True
                else:
                    dense = self._dense(flat, cr, (CB, CB + 1, CB + 2, DB, DB + 1, DB + 2))
                    r.stability = M.stability(dense, n, op, seeds, opt.iterations)
                    r.verified = r.stability <= opt.max_stability
            if r.verified or last:
                break
            tau = min(tau * 1.25, 1.0)
        rep = r
        # the values that differ from the lanes (bit patterns), by lane
        out: Dict[int, np.ndarray] = {}
        for (lane, k), v in cr.over.items():
            if np.float32(v).view(np.uint32) != flat[lane, k].view(np.uint32):
                if lane not in out:
                    out[lane] = flat[lane].copy()
                out[lane][k] = v
        out = {l: v.reshape(n) for l, v in out.items()}
        return out, rep, op

    @staticmethod
    def _dense(flat, cr, lanes):
        """The lanes `lanes` with the overlay's changes, dense (others None)."""
        out = [None] * 12
        for l in lanes:
            out[l] = flat[l].copy()
        for (lane, k), v in cr.over.items():
            if out[lane] is not None:
                out[lane][k] = v
        return out

    def apply(self, engine) -> ConformalReport:
        """ConformalPec::apply (conformal.rs:755): patches the engine's
        coefficient lanes and registers the reduced currents' operator
        (before operators that read the lanes are added)."""
        if getattr(engine, "_dense_lanes", None) is None:
            raise EmsError(nat.EMS_E_INVALID, "conformal PEC needs the host lanes")
        changed, rep, op = self.build(engine._dense_lanes, engine.grid, engine.dt)
        engine.patch_lanes(changed)
        if op is not None:
            engine.add_operator(op)
        return rep


# -- a PEC cylinder on a rectilinear grid -------------------------------------

def _disk_rect_area_exact(x0, x1, y0, y1, xc, yc, r) -> float:
    """Closed form: inclusion-exclusion over the quadrant function
    Q(x, y) = area of the disk with X <= x, Y <= y."""
    def G(u, v):
        return _corner(u, v, r)
    return (G(x1 - xc, y1 - yc) - G(x0 - xc, y1 - yc) - G(x1 - xc, y0 - yc) +
            G(x0 - xc, y0 - yc))


def _seg(u: float, r: float) -> float:
    """Area of the disk of radius r with X <= u."""
    u = min(max(u, -r), r)
    return r * r * math.acos(-u / r) + u * math.sqrt(r * r - u * u) if r > 0 else 0.0


def _corner(u: float, v: float, r: float) -> float:
    """Area of the disk of radius r (centre 0) with X <= u and Y <= v:
    the integral over X <= u of the chord part below v."""
    u = min(max(u, -r), r)
    if v <= -r:
        return 0.0
    if v >= r:
        return _seg(u, r)

    # integral over x in [-r, u] of clip(v - (-h(x)), 0, 2h(x)), h = sqrt(r^2 - x^2)
    def F(x):
        # antiderivative of h(x)
        x = min(max(x, -r), r)
        return 0.5 * (x * math.sqrt(r * r - x * x) + r * r * math.asin(x / r))
    xv = math.sqrt(max(r * r - v * v, 0.0))
    total = 0.0
    if v >= 0.0:
        # where |x| < xv the chord top h(x) > v: part below v is v + h(x);
        # elsewhere the whole chord 2h(x)
        a1, b1 = -r, min(u, -xv)
        if b1 > a1:
            total += 2.0 * (F(b1) - F(a1))
        a2, b2 = max(-xv, -r), min(u, xv)
        if b2 > a2:
            total += v * (b2 - a2) + (F(b2) - F(a2))
        a3, b3 = xv, u
        if b3 > a3:
            total += 2.0 * (F(b3) - F(a3))
    else:
        # v < 0: only |x| < xv has the chord bottom -h(x) below v
        a2, b2 = -xv, min(u, xv)
        if b2 > a2:
            total += v * (b2 - a2) + (F(b2) - F(a2))
    return total


_GL = np.polynomial.legendre.leggauss(48)


def _disk_rect_moments(x0, x1, y0, y1, xc, yc, r) -> Tuple[float, float]:
    """First moments (integrals of x and of y) of the disk-rectangle
    intersection, by Gauss-Legendre quadrature over the x intervals on which
    the chord's clipping to [y0, y1] keeps one form."""
    cuts = [x0, x1, xc - r, xc + r]
    for y in (y0, y1):
        d = r * r - (y - yc) ** 2
        if d > 0.0:
            cuts += [xc - math.sqrt(d), xc + math.sqrt(d)]
    cuts = sorted(min(max(c, x0), x1) for c in cuts)
    mx = my = 0.0
    for a, b in zip(cuts[:-1], cuts[1:]):
        if b <= a:
            continue
        x = 0.5 * (a + b) + 0.5 * (b - a) * _GL[0]
        h = np.sqrt(np.clip(r * r - (x - xc) ** 2, 0.0, None))
        lo = np.maximum(y0, yc - h)
        hi = np.minimum(y1, yc + h)
        ln = np.clip(hi - lo, 0.0, None)
        w = 0.5 * (b - a) * _GL[1]
        mx += float(np.sum(w * x * ln))
        my += float(np.sum(w * 0.5 * (hi * hi - lo * lo) * (ln > 0)))
    return mx, my


def _open_interval_centroid(a0, a1, c_lo, c_hi, metal_inside: bool) -> float:
    """Centroid (as a fraction of [a0, a1]) of the open part of the segment,
    metal on [c_lo, c_hi] (inside) or outside it."""
    if metal_inside:
        pieces = [(a0, min(a1, c_lo)), (max(a0, c_hi), a1)]
    else:
        pieces = [(max(a0, c_lo), min(a1, c_hi))]
    num = den = 0.0
    for u, v in pieces:
        if v > u:
            num += 0.5 * (u + v) * (v - u)
            den += v - u
    t = (num / den - a0) / (a1 - a0) if den > 0.0 else 0.5
    return min(max(t, 0.0), 1.0)


def cylinder_z(grid, center: Sequence[float], radius: float, mats=None,
               options: ConformalOptions = None, metal: str = "inside",
               z_nodes: Optional[Tuple[int, int]] = None, centroids: bool = True) -> ConformalPec:
    """A circular PEC boundary along z of `radius` about (x, y) = `center`:
    metal inside the circle (a rod) or outside it (a cavity), over all z or
    between node planes z_nodes = (k0, k1) (flat end caps on those planes).
    Edges wholly in metal become PEC in `mats` (CellMaterials, per-edge
    mask); cut x/y edges and the faces the circle cuts become conformal
    edges and faces. Faces normal to x (y) take the open fraction of the y
    (x) edge they stand on (the geometry is z-invariant between the caps).
    With `centroids` every cut face carries the centroid of its open area
    (small ones are then interpolated, else merged)."""
    if metal not in ("inside", "outside"):
        raise EmsError(nat.EMS_E_INVALID, f"metal must be inside or outside, not {metal!r}")
    xs, ys, zs = (np.asarray(grid.lines[a], np.float64) for a in range(3))
    nx, ny, nz = len(xs), len(ys), len(zs)
    k0, k1 = (0, nz - 1) if z_nodes is None else (int(z_nodes[0]), int(z_nodes[1]))
    if not 0 <= k0 < k1 <= nz - 1:
        raise EmsError(nat.EMS_E_INVALID, f"cylinder z nodes {z_nodes} outside 0..{nz - 1}")
    xc, yc, r = float(center[0]), float(center[1]), float(radius)
    pec = ConformalPec(options)
    in_disk = ((xs[:, None] - xc) ** 2 + (ys[None, :] - yc) ** 2) < r * r

    def disk_frac(a0, a1, c):
        """Fraction of the segment [a0, a1] (relative to the centre) inside
        the disk, at transverse offset c from the centre."""
        d = r * r - c * c
        if d <= 0.0:
            return 0.0
        s = math.sqrt(d)
        return max(0.0, min(a1, s) - max(a0, -s)) / (a1 - a0)

    # x edges (i, j): from xs[i] to xs[i+1] at ys[j]; y edges (i, j) likewise
    gx = np.array([[disk_frac(xs[i] - xc, xs[i + 1] - xc, ys[j] - yc) for j in range(ny)]
                   for i in range(nx - 1)])
    gy = np.array([[disk_frac(ys[j] - yc, ys[j + 1] - yc, xs[i] - xc) for j in range(ny - 1)]
                   for i in range(nx)])
    # z faces (i, j): the cell [xs[i], xs[i+1]] x [ys[j], ys[j+1]]
    gz = np.zeros((nx - 1, ny - 1))
    for i in range(nx - 1):
        for j in range(ny - 1):
            a = _disk_rect_area_exact(xs[i], xs[i + 1], ys[j], ys[j + 1], xc, yc, r)
            gz[i, j] = a / ((xs[i + 1] - xs[i]) * (ys[j + 1] - ys[j]))
    if metal == "inside":
        fx, fy, fz, zmetal = 1.0 - gx, 1.0 - gy, 1.0 - gz, in_disk
    else:
        fx, fy, fz, zmetal = gx, gy, gz, ~in_disk
    # fractions within 1e-9 of 0 or 1 are rounding of whole cells
    fx, fy, fz = (np.where(f < 1e-9, 0.0, np.where(f > 1.0 - 1e-9, 1.0, f)) for f in (fx, fy, fz))
    inside = metal == "inside"

    def chord(c):
        d = r * r - c * c
        s = math.sqrt(d) if d > 0.0 else 0.0
        return -s, s
    cz: Dict[Tuple[int, int], Tuple[float, float]] = {}
    cx: Dict[Tuple[int, int], float] = {}
    cy: Dict[Tuple[int, int], float] = {}
    if centroids:
        for (i, j) in zip(*np.nonzero((fz < 1.0) & (fz > 0.0))):
            x0, x1, y0, y1 = xs[i], xs[i + 1], ys[j], ys[j + 1]
            area = (x1 - x0) * (y1 - y0)
            a_in = _disk_rect_area_exact(x0, x1, y0, y1, xc, yc, r)
            mx, my = _disk_rect_moments(x0, x1, y0, y1, xc, yc, r)
            if inside:
                a_open = area - a_in
                mx = 0.5 * (x0 + x1) * area - mx
                my = 0.5 * (y0 + y1) * area - my
            else:
                a_open = a_in
            t = ((mx / a_open - x0) / (x1 - x0), (my / a_open - y0) / (y1 - y0))
            cz[(int(i), int(j))] = tuple(min(max(v, 0.0), 1.0) for v in t)
        for (i, j) in zip(*np.nonzero((fy < 1.0) & (fy > 0.0))):
            lo, hi = chord(xs[i] - xc)
            cy[(int(i), int(j))] = _open_interval_centroid(ys[j], ys[j + 1], yc + lo, yc + hi,
                                                           inside)
        for (i, j) in zip(*np.nonzero((fx < 1.0) & (fx > 0.0))):
            lo, hi = chord(ys[j] - yc)
            cx[(int(i), int(j))] = _open_interval_centroid(xs[i], xs[i + 1], xc + lo, xc + hi,
                                                           inside)

    def push(axis, node, frac, cen):
        f = ConformalFace(axis, node, frac)
        if centroids:
            pec.push_face_with_centroid(f, cen)
        else:
            pec.push_face(f)
    if mats is not None:
        for i, j in zip(*np.nonzero(zmetal)):
            for k in range(k0, k1):
                mats.set_pec_edge(2, (int(i), int(j), k))
    for k in range(k0, k1 + 1):
        for axis, f_e in ((0, fx), (1, fy)):
            for (i, j) in zip(*np.nonzero(f_e < 1.0)):
                f = float(f_e[i, j])
                if f <= 0.0:
                    if mats is not None:
                        mats.set_pec_edge(axis, (int(i), int(j), k))
                else:
                    pec.push_edge(ConformalEdge(axis, (int(i), int(j), k), f))
        for (i, j) in zip(*np.nonzero((fz < 1.0) & (fz > 0.0))):
            push(2, (int(i), int(j), k), float(fz[i, j]), cz.get((int(i), int(j))))
        if k < k1:
            # x-normal faces span y [j, j+1] at xs[i]: the y edge's fraction;
            # centroids along (b, c) = (y, z) for x faces, (z, x) for y faces
            for (i, j) in zip(*np.nonzero((fy < 1.0) & (fy > 0.0))):
                push(0, (int(i), int(j), k), float(fy[i, j]), (cy.get((int(i), int(j))), 0.5))
            for (i, j) in zip(*np.nonzero((fx < 1.0) & (fx > 0.0))):
                push(1, (int(i), int(j), k), float(fx[i, j]), (0.5, cx.get((int(i), int(j)))))
    return pec
