"""The first-order Mur absorbing boundary (MurBoundary, src/engine/boundary/
mur.rs): on outer faces built as PEC, every tangential face edge follows
the one-way wave equation discretised between the boundary node and its
interior neighbour,

    u_b^{n+1} = u_1^n + k (u_1^{n+1} - u_b^n),  k = (v dt - d)/(v dt + d),

with d the boundary cell's size along the face normal; an edge on two Mur
faces (a domain edge) takes the average of its two 1-D updates. The host
builds the edge lists (as MurBoundary::build does, mur.rs:395-585) and the
device or the oracle runs the update (ems_gpu_add_mur).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError
from .engine import Axis, Face


class MurDesc(C.Structure):
    """ems_mur_desc (include/ems_gpu.h)."""
    _fields_ = [("n_edges", C.c_int32), ("edge_axis", C.POINTER(C.c_int32)),
                ("edge_node", C.POINTER(C.c_int32)), ("edge_inner", C.POINTER(C.c_int32)),
                ("edge_k", C.POINTER(C.c_float)), ("edge_hold", C.POINTER(C.c_uint64)),
                ("n_corners", C.c_int32), ("corner_axis", C.POINTER(C.c_int32)),
                ("corner_node", C.POINTER(C.c_int32)), ("corner_inner0", C.POINTER(C.c_int32)),
                ("corner_inner1", C.POINTER(C.c_int32)), ("corner_k0", C.POINTER(C.c_float)),
                ("corner_k1", C.POINTER(C.c_float)), ("corner_hold", C.POINTER(C.c_uint64)),
                ("openems_order", C.c_int32), ("openems_corners", C.c_int32)]


class MurBoundary:
    """MurBoundary (mur.rs:207-640) on one or more faces, each with its
    phase velocity; domain edges where two of its faces meet are its
    corners. Register it before sources, probes and lumped elements."""

    def __init__(self):
        self.faces: List[Tuple[Face, float]] = []
        # (axis, node, inner, k, hold)
        self.edges: List[Tuple[int, Tuple[int, int, int], Tuple[int, int, int], np.float32, int]] = []
        # (axis, node, inner0, inner1, k0, k1, hold)
        self.corners: list = []
        self._openems_order = False
        self._openems_corners = False

    @staticmethod
    def coefficient(v: float, dt: float, delta: float) -> float:
        """MurBoundary::coefficient (mur.rs:732-736)."""
        return (v * dt - delta) / (v * dt + delta)

    @staticmethod
    def new(engine, face: Face, phase_velocity: float) -> "MurBoundary":
        return MurBoundary.with_faces(engine, [(face, phase_velocity)])

    @staticmethod
    def with_faces(engine, faces: Sequence[Tuple[Face, float]]) -> "MurBoundary":
        """MurBoundary::with_faces / build (mur.rs:395-585): the face edges
        of every listed face (tangential axes t, then the third axis o, in
        index order), a domain edge on two listed faces kept once, from the
        lower-index face, as a corner."""
        grid, dt = engine.grid, float(engine.dt)
        lanes = engine.coefficients()
        n = grid.nodes()
        k_face: List[Optional[np.float32]] = [None] * 6
        m = MurBoundary()
        for face, v in faces:
            face = Face(face)
            v = float(v)
            if not (np.isfinite(v) and v > 0.0):
                raise EmsError(nat.EMS_E_INVALID,
                               f"Mur boundary on {face.name}: phase velocity {v} must be finite and > 0")
            if k_face[face] is not None:
                raise EmsError(nat.EMS_E_INVALID, f"Mur boundary: face {face.name} listed twice")
            a = int(face.axis)
            if n[a] < 3:
                raise EmsError(nat.EMS_E_INVALID,
                               f"Mur boundary on {face.name}: needs at least 3 mesh lines along "
                               f"axis {a}, got {n[a]}")
            d = grid.delta(a, n[a] - 2) if face.is_max else grid.delta(a, 0)
            k_face[face] = np.float32(MurBoundary.coefficient(v, dt, d))
            m.faces.append((face, v))

        def plane(f: Face) -> int:
            return n[int(f.axis)] - 1 if f.is_max else 0

        def inner(f: Face, node):
            node = list(node)
            a = int(f.axis)
            node[a] = node[a] - 1 if f.is_max else node[a] + 1
            return tuple(node)

        for face, _ in m.faces:
            a = int(face.axis)
            for t in [x for x in range(3) if x != a]:
                o = 3 - a - t
                ca, cb = lanes[t], lanes[3 + t]
                for p in range(n[t] - 1):
                    for q in range(n[o]):
                        node = [0, 0, 0]
                        node[a], node[t], node[o] = plane(face), p, q
                        node = tuple(node)
                        if ca[node] != 0.0 or cb[node] != 0.0:
                            raise EmsError(nat.EMS_E_INVALID,
                                           f"Mur boundary on {face.name}: the face must be built "
                                           f"with FaceCondition.Pec (edge {t} at {node} is not)")
                        other = None
                        for g in (Face(2 * o), Face(2 * o + 1)):
                            if plane(g) == q and k_face[g] is not None:
                                other = g
                                break
                        if other is None:
                            m.edges.append((t, node, inner(face, node), k_face[face], 0))
                        elif face < other:
                            m.corners.append((t, node, inner(face, node), inner(other, node),
                                              k_face[face], k_face[other], 0))
        return m

    def openems_order(self, on: bool = True) -> "MurBoundary":
        """MurBoundary::openems_order (mur.rs:655-667)."""
        self._openems_order = bool(on)
        return self

    def openems_corners(self, on: bool = True) -> "MurBoundary":
        """MurBoundary::openems_corners (mur.rs:640-653)."""
        self._openems_corners = bool(on)
        return self

    def hold_faces(self, faces: Sequence[Face], until: int) -> "MurBoundary":
        """MurBoundary::hold_faces (mur.rs:587-620): no update of the edges
        on `faces` before step `until`."""
        def on(node, inn):
            for f in faces:
                a = int(Face(f).axis)
                if node[a] != inn[a] and (node[a] == 0) != Face(f).is_max:
                    return True
            return False
        self.edges = [(t, nd, i, k, until if on(nd, i) else h) for t, nd, i, k, h in self.edges]
        self.corners = [(t, nd, i0, i1, k0, k1, until if (on(nd, i0) or on(nd, i1)) else h)
                        for t, nd, i0, i1, k0, k1, h in self.corners]
        return self

    def edge_count(self) -> int:
        return len(self.edges) + len(self.corners)

    def arrays(self) -> dict:
        """Flat arrays of the edges and corners (the descriptor's)."""
        e, c = self.edges, self.corners
        i32 = lambda x: np.ascontiguousarray(np.array(x, dtype=np.int32).reshape(-1))  # noqa: E731
        return {
            "eax": i32([t for t, *_ in e]), "enode": i32([x[1] for x in e]),
            "einner": i32([x[2] for x in e]),
            "ek": np.ascontiguousarray(np.array([x[3] for x in e], dtype=np.float32)),
            "ehold": np.ascontiguousarray(np.array([x[4] for x in e], dtype=np.uint64)),
            "cax": i32([x[0] for x in c]), "cnode": i32([x[1] for x in c]),
            "cin0": i32([x[2] for x in c]), "cin1": i32([x[3] for x in c]),
            "ck0": np.ascontiguousarray(np.array([x[4] for x in c], dtype=np.float32)),
            "ck1": np.ascontiguousarray(np.array([x[5] for x in c], dtype=np.float32)),
            "chold": np.ascontiguousarray(np.array([x[6] for x in c], dtype=np.uint64)),
        }

    def desc(self):
        """The ems_mur_desc and the arrays it points into."""
        a = self.arrays()
        P = nat.ptr
        d = MurDesc(len(self.edges), P(a["eax"], C.c_int32), P(a["enode"], C.c_int32),
                    P(a["einner"], C.c_int32), P(a["ek"], C.c_float), P(a["ehold"], C.c_uint64),
                    len(self.corners), P(a["cax"], C.c_int32), P(a["cnode"], C.c_int32),
                    P(a["cin0"], C.c_int32), P(a["cin1"], C.c_int32), P(a["ck0"], C.c_float),
                    P(a["ck1"], C.c_float), P(a["chold"], C.c_uint64),
                    1 if self._openems_order else 0, 1 if self._openems_corners else 0)
        return d, a


__all__ = ["MurBoundary", "MurDesc", "Axis"]
