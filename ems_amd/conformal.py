"""Conformal (sub-cell) PEC by coefficient patches (src/engine/conformal.rs).

The Dey-Mittra scheme in the integral form of the engine is a change of
coefficients (conformal.rs:1-30): a cut E edge with open length fraction
f_e gets cb -> f_e cb, a cut face with open area fraction f_A gets
db -> db / f_A; ca and da stay. The patches are written into the
coefficient lanes, so the device step kernels run them unchanged.

Arithmetic as ConformalPec::apply_known (conformal.rs:786-793: the edge
scaling, (cb as f64 * f_e) as f32) and the face loop of construct
(conformal.rs:1542-1566: (db as f64 / f_A) as f32, f_A below the merge
threshold raised to it, the threshold at least 1e-6).

Scope: every cut face keeps its own current. The reference instead takes
the current of a small face (f_A below the threshold) from its neighbours
(interpolation with deviation currents, or merging, solved per step with
envelope-Cholesky blocks by its MergedFaces operator, conformal.rs:2313)
wherever a stencil exists, and raises f_A to the threshold only for faces
without one. Here every small face is raised to the threshold (the
reference's fallback), or kept at its true area with merge_below = 0
(plain Dey-Mittra, which needs a shorter time step). The stability check
(conformal.rs:1698-1760, power iteration) is not run.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
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
    """The merge threshold of ConformalOptions (conformal.rs:283-331):
    faces below it are raised to it."""
    merge_below: float = 0.5


@dataclass
class ConformalReport:
    """Counts of ConformalReport (conformal.rs:350-424) that apply here."""
    edges: int = 0
    faces: int = 0
    floored: int = 0
    pec_edges: int = 0


def _others(a: int) -> Tuple[int, int]:
    return ((1, 2), (0, 2), (0, 1))[a]


def _face_edges(a: int, p) -> List[Tuple[int, Tuple[int, int, int]]]:
    """The four edges around the face normal to a with lowest corner p."""
    b, c = _others(a)
    pb = list(p)
    pb[b] += 1
    pc = list(p)
    pc[c] += 1
    return [(b, tuple(p)), (b, tuple(pc)), (c, tuple(p)), (c, tuple(pb))]


class ConformalPec:
    """Cut edges and faces of curved metal (ConformalPec, conformal.rs:440-755)."""

    def __init__(self, options: ConformalOptions = None):
        self.options = options or ConformalOptions()
        self.edges: List[ConformalEdge] = []
        self.faces: List[ConformalFace] = []

    def push_edge(self, e: ConformalEdge) -> None:
        self.edges.append(e)

    def push_face(self, f: ConformalFace) -> None:
        self.faces.append(f)

    def validate(self, n) -> None:
        """conformal.rs:1019-1035: fractions in [0, 1], edges and faces
        inside the grid."""
        for e in self.edges:
            a = e.axis
            if not (0.0 <= e.fraction <= 1.0) or any(e.node[d] >= n[d] for d in range(3)) \
                    or e.node[a] + 1 >= n[a]:
                raise EmsError(nat.EMS_E_INVALID, f"invalid conformal edge {e}")
        for f in self.faces:
            b, c = _others(f.axis)
            if not (0.0 <= f.fraction <= 1.0) or any(f.node[d] >= n[d] for d in range(3)) \
                    or f.node[b] + 1 >= n[b] or f.node[c] + 1 >= n[c]:
                raise EmsError(nat.EMS_E_INVALID, f"invalid conformal face {f}")

    def patched(self, lanes: np.ndarray) -> Tuple[Dict[int, np.ndarray], ConformalReport]:
        """The changed lanes (index -> new dense lane) for coefficient lanes
        `lanes` (12, nx, ny, nz), and the report."""
        n = lanes.shape[1:]
        self.validate(n)
        out: Dict[int, np.ndarray] = {}

        def lane(l):
            if l not in out:
                out[l] = lanes[l].copy()
            return out[l]

        rep = ConformalReport(edges=len(self.edges))
        for e in self.edges:
            cb = lane(CB + e.axis)
            cb[e.node] = np.float32(float(cb[e.node]) * e.fraction)
            if e.fraction == 0.0:
                rep.pec_edges += 1
        # whether an edge is open after the edge scaling (cb > 0)
        cbs = [out.get(CB + a, lanes[CB + a]) for a in range(3)]
        tau = min(max(self.options.merge_below, 0.0), 1.0)
        for f in self.faces:
            if not any(cbs[b][q] > 0.0 for b, q in _face_edges(f.axis, f.node)):
                continue  # not live: every edge closed
            fa = f.fraction
            if fa < tau:
                fa = max(tau, 1e-6)
                rep.floored += 1
            if fa < 1.0:
                db = lane(DB + f.axis)
                db[f.node] = np.float32(float(db[f.node]) / fa)
            rep.faces += 1
        return out, rep

    def apply(self, engine) -> ConformalReport:
        """Patches the engine's coefficient lanes (before operators that read
        them are added, as ConformalPec::apply on a fresh engine)."""
        if getattr(engine, "_dense_lanes", None) is None:
            raise EmsError(nat.EMS_E_INVALID, "conformal PEC needs the host lanes")
        changed, rep = self.patched(engine._dense_lanes)
        engine.patch_lanes(changed)
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


def cylinder_z(grid, center: Sequence[float], radius: float, mats=None,
               options: ConformalOptions = None, metal: str = "inside",
               z_nodes: Optional[Tuple[int, int]] = None) -> ConformalPec:
    """A circular PEC boundary along z of `radius` about (x, y) = `center`:
    metal inside the circle (a rod) or outside it (a cavity), over all z or
    between node planes z_nodes = (k0, k1) (flat end caps on those planes).
    Edges wholly in metal become PEC in `mats` (CellMaterials, per-edge
    mask); cut x/y edges and the faces the circle cuts become conformal
    edges and faces. Faces normal to x (y) take the open fraction of the y
    (x) edge they stand on (the geometry is z-invariant between the caps)."""
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
    fx = np.clip(fx, 0.0, 1.0)
    fy = np.clip(fy, 0.0, 1.0)
    fz = np.clip(fz, 0.0, 1.0)
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
            pec.push_face(ConformalFace(2, (int(i), int(j), k), float(fz[i, j])))
        if k < k1:
            # x-normal faces span y [j, j+1] at xs[i]: the y edge's fraction
            for (i, j) in zip(*np.nonzero((fy < 1.0) & (fy > 0.0))):
                pec.push_face(ConformalFace(0, (int(i), int(j), k), float(fy[i, j])))
            for (i, j) in zip(*np.nonzero((fx < 1.0) & (fx > 0.0))):
                pec.push_face(ConformalFace(1, (int(i), int(j), k), float(fx[i, j])))
    return pec
