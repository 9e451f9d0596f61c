"""openEMS's ``PML_<n>`` layers (OpenEmsPml, src/engine/boundary/upml.rs).

The uniaxial PML of Gedney (IEEE TAP 44(12), 1996) in the two-stage form
openEMS runs: per layer edge a flux F with

    F' = (1 - a_p)/(1 + a_p) F + dt/((A/l)(1 + a_p)) (circulation)
    u' = ((1 - a_q)/(1 + a_q) u - (1 - a_a)/((1 + a_q) m) F) + (1 + a_a)/((1 + a_q) m) F'

(p, q the next and previous axes of the edge's axis a, a_w = sigma_w dt /
(2 eps0), m = eps0' eps_r for E, mu0' mu_r for H). The host grades the
layers (Profile, upml.rs:153-199), rewrites the coefficients of the layer
edges so that the ordinary update leaves the flux increment in the field
(ca = 0, cb = dt/((A/l)(1 + a_p))) and lists the edges with their four
update coefficients (OpenEmsPml::build, upml.rs:602-700); the device (or
the oracle) runs the second stage after each half-step (ems_gpu_add_upml).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError
from .engine import EPS0, Axis, Face, Z0

# upml.rs:126-133: the permittivity and permeability of vacuum in the second
# stage of openEMS's layer update.
OPENEMS_PML_EPS0 = 8.854_187_817_620_389e-12
OPENEMS_PML_MU0 = 1.256_637_062e-6


def openems_pml_sigma(depth: float, width: float, cells: int) -> float:
    """openems_pml_sigma (upml.rs:143-150): the conductivity at physical
    depth `depth` below the interface of a layer of `cells` cells and width
    `width`: ln(1e6) ln(2.5) / (2 D Z0 (2.5^n - 1)) 2.5^(d/D), D = width/n;
    0 at and outside the interface."""
    if depth <= 0.0 or cells == 0 or width <= 0.0:
        return 0.0
    dl = width / cells
    g = 2.5
    return math.log(1e6) * math.log(g) / (2.0 * dl * Z0 * (g ** int(cells) - 1.0)) * g ** (depth / dl)


class UpmlDesc(C.Structure):
    """ems_upml_desc (include/ems_gpu.h)."""
    _fields_ = [("n_e", C.c_int32), ("e_axis", C.POINTER(C.c_int32)),
                ("e_node", C.POINTER(C.c_int32)), ("e_coef", C.POINTER(C.c_float)),
                ("n_h", C.c_int32), ("h_axis", C.POINTER(C.c_int32)),
                ("h_node", C.POINTER(C.c_int32)), ("h_coef", C.POINTER(C.c_float))]


class PmlProfile:
    """Profile (upml.rs:153-199): per axis a = sigma dt/(2 eps0) at the
    mesh lines (`node`) and at the cell midpoints (`mid`, one extra 0)."""

    def __init__(self, grid, dt: float, faces: Sequence[Tuple[Face, int]]):
        n = grid.nodes()
        self.node = [np.zeros(n[a]) for a in range(3)]
        self.mid = [np.zeros(n[a]) for a in range(3)]
        for face, cells in faces:
            ax = int(face.axis)
            l = grid.lines[ax]
            last = l.size - 1
            inner, outer = (float(l[last - cells]), float(l[last])) if face.is_max else \
                (float(l[cells]), float(l[0]))
            width = abs(outer - inner)

            def a(x, inner=inner, mx=face.is_max, width=width, cells=cells):
                d = x - inner if mx else inner - x
                return openems_pml_sigma(d, width, cells) * dt / (2.0 * EPS0)

            for k in range(last + 1):
                self.node[ax][k] += a(float(l[k]))
                if k < last:
                    self.mid[ax][k] += a(0.5 * (float(l[k]) + float(l[k + 1])))

    def at(self, e_side: bool, comp: int):
        """(a_a, a_p, a_q) of every E edge (e_side) or H dual edge along
        `comp`, as arrays broadcasting over (nx, ny, nz) (Profile::at)."""
        p, q = (comp + 1) % 3, (comp + 2) % 3
        own, other = (self.mid, self.node) if e_side else (self.node, self.mid)

        def along(arr, axis):
            shape = [1, 1, 1]
            shape[axis] = arr.size
            return arr.reshape(shape)

        return along(own[comp], comp), along(other[p], p), along(other[q], q)


class PmlLayers:
    """PmlLayers (upml.rs:255-268): which edges lie inside the layers."""

    def __init__(self, profile: PmlProfile):
        self._p = profile

    def contains(self, e_side: bool, comp: int, node) -> bool:
        aa, ap, aq = self._p.at(e_side, comp)
        i, j, k = node
        pick = lambda x: float(x[min(i, x.shape[0] - 1), min(j, x.shape[1] - 1),  # noqa: E731
                                 min(k, x.shape[2] - 1)])
        return pick(aa) != 0.0 or pick(ap) != 0.0 or pick(aq) != 0.0


class OpenEmsPml:
    """openEMS's PML_n layers on any set of faces (OpenEmsPml, upml.rs:541-760).

    Build the engine with PEC on the layer faces, then create this (it
    rewrites the coefficients of the layer edges on the engine, host and
    device) and register it with ``engine.add_operator`` before dispersive
    media, Mur boundaries, plane waves and cut-face currents."""

    def __init__(self):
        self.faces: list = []
        self.e = None  # (axis, nodes (n, 3), coef (n, 4)) of the E edges
        self.h = None
        self.layers: Optional[PmlLayers] = None
        self.nodes = None
        self.dt = None

    @staticmethod
    def new(engine, faces: Iterable[Tuple[Face, int]]) -> "OpenEmsPml":
        """OpenEmsPml::new (upml.rs:563-576): m from the engine's
        coefficients (the capacitance dt (1 + ca)/(2 cb) over A/l)."""
        return OpenEmsPml._build(engine, faces, None, None)

    @staticmethod
    def with_materials(engine, faces: Iterable[Tuple[Face, int]], eps_r, mu_r) -> "OpenEmsPml":
        """OpenEmsPml::with_materials (upml.rs:578-600): m = eps0' eps_r of
        the E edge (mu0' mu_r of the H dual edge). eps_r, mu_r: a scalar or
        per-edge values of shape (3, nx, ny, nz) (EdgeMaterials::eps_r /
        mu_r by the edge's axis)."""
        n = engine.grid.nodes()
        mats = []
        for v in (eps_r, mu_r):
            a = np.asarray(v, dtype=np.float64)
            if a.ndim != 0 and a.shape != (3,) + tuple(n):
                raise EmsError(nat.EMS_E_INVALID, "PML: the edge materials do not match the grid")
            mats.append(a)
        return OpenEmsPml._build(engine, faces, mats[0], mats[1])

    @staticmethod
    def _build(engine, faces, eps_r, mu_r) -> "OpenEmsPml":
        grid, dt = engine.grid, float(engine.dt)
        n = grid.nodes()
        seen = set()
        lst = []
        for face, cells in faces:
            face, cells = Face(face), int(cells)
            if face in seen:
                raise EmsError(nat.EMS_E_INVALID, f"PML: face {face.name} given twice")
            seen.add(face)
            total = n[int(face.axis)] - 1
            if cells == 0 or cells > total:
                raise EmsError(nat.EMS_E_INVALID,
                               f"PML on {face.name}: {cells} cells, the mesh has {total} along "
                               f"{Axis(int(face.axis)).name}")
            lst.append((face, cells))
        prof = PmlProfile(grid, dt, lst)
        lanes = engine.coefficients()
        if lanes is None:
            raise EmsError(nat.EMS_E_INVALID, "PML: the engine's host lanes were released")
        d = [np.append(np.diff(grid.lines[a]), np.nan) for a in range(3)]  # delta, NaN past the end
        dd = [np.array([grid.dual_delta(a, i) for i in range(n[a])]) for a in range(3)]
        changed = {}
        out = {}
        for e_side in (True, False):
            axs, nds, cos = [], [], []
            for comp in range(3):
                p, q = (comp + 1) % 3, (comp + 2) % 3
                l0 = comp if e_side else 6 + comp
                ca, cb = lanes[l0], lanes[l0 + 3]
                aa, ap, aq = prof.at(e_side, comp)
                mask = (cb != 0.0) & ((aa != 0.0) | (ap != 0.0) | (aq != 0.0))
                if not mask.any():
                    continue
                sel = np.nonzero(mask)
                aa_s = np.broadcast_to(aa, n)[sel]
                ap_s = np.broadcast_to(ap, n)[sel]
                aq_s = np.broadcast_to(aq, n)[sel]
                # A/l of the edge (E) or dual edge (H) as the coefficients take it.
                if e_side:
                    geo = dd[p][sel[p]] * dd[q][sel[q]] / d[comp][sel[comp]]
                else:
                    geo = d[p][sel[p]] * d[q][sel[q]] / dd[comp][sel[comp]]
                if not np.all(np.isfinite(geo)):
                    raise EmsError(nat.EMS_E_INVALID, "PML: a layer edge outside the mesh")
                mat = eps_r if e_side else mu_r
                if mat is None:
                    ca_s = ca[sel].astype(np.float64)
                    cb_s = cb[sel].astype(np.float64)
                    m = dt * (1.0 + ca_s) / (2.0 * cb_s) / geo
                else:
                    rel = mat if mat.ndim == 0 else mat[comp][sel]
                    m = (OPENEMS_PML_EPS0 if e_side else OPENEMS_PML_MU0) * rel
                new_ca, new_cb = ca.copy(), cb.copy()
                new_ca[sel] = 0.0
                new_cb[sel] = (dt / ((1.0 + ap_s) * geo)).astype(np.float32)
                changed[l0], changed[l0 + 3] = new_ca, new_cb
                co = np.empty((sel[0].size, 4), dtype=np.float32)
                co[:, 0] = (1.0 - ap_s) / (1.0 + ap_s)
                co[:, 1] = (1.0 - aq_s) / (1.0 + aq_s)
                co[:, 2] = (1.0 + aa_s) / ((1.0 + aq_s) * m)
                co[:, 3] = (1.0 - aa_s) / ((1.0 + aq_s) * m)
                axs.append(np.full(sel[0].size, comp, dtype=np.int32))
                nds.append(np.stack(sel, axis=1).astype(np.int32))
                cos.append(co)
            if axs:
                out[e_side] = (np.concatenate(axs), np.ascontiguousarray(np.concatenate(nds)),
                               np.ascontiguousarray(np.concatenate(cos)))
            else:
                out[e_side] = (np.zeros(0, np.int32), np.zeros((0, 3), np.int32),
                               np.zeros((0, 4), np.float32))
        if changed:
            engine.patch_lanes(changed)
        pml = OpenEmsPml()
        pml.faces = lst
        pml.e, pml.h = out[True], out[False]
        pml.layers = PmlLayers(prof)
        pml.nodes, pml.dt = n, dt
        return pml

    def edge_counts(self) -> Tuple[int, int]:
        """OpenEmsPml::edge_counts: E edges and H dual edges inside the layers."""
        return int(self.e[0].size), int(self.h[0].size)

    def state_counts(self) -> Tuple[int, int]:
        """OpenEmsPml::state_counts: every layer edge carries state."""
        return self.edge_counts()

    def desc(self):
        """The ems_upml_desc and the arrays it points into."""
        P = nat.ptr
        (ea, en, ec), (ha, hn, hc) = self.e, self.h
        keep = (ea, en, ec, ha, hn, hc)
        d = UpmlDesc(ea.size, P(ea, C.c_int32), P(en, C.c_int32), P(ec, C.c_float),
                     ha.size, P(ha, C.c_int32), P(hn, C.c_int32), P(hc, C.c_float))
        return d, keep


__all__ = ["OpenEmsPml", "PmlLayers", "PmlProfile", "UpmlDesc", "openems_pml_sigma",
           "OPENEMS_PML_EPS0", "OPENEMS_PML_MU0"]
