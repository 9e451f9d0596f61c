"""Dispersive (Drude, Lorentz, Debye) media by the auxiliary differential
equation method (src/engine/dispersive.rs).

Every pole is a series R-L-C branch across its edge (dispersive.rs:25-45),
integrated with the semi-implicit trapezoidal rule or the explicit
leapfrog (dispersive.rs:428-461). The host side of `Dispersive::install`
(dispersive.rs:2063-2319) runs here: edges in (plane, field, axis, j, k)
order, each pole's constants from its strength and the edge's geometry
m = X0 * g, edges whose cb (db) is 0 skipped. The device keeps the branch
state and corrects the pole edges after every E (H) update
(dispersive.rs:1183-1343, 1392-1470).
"""
from __future__ import annotations

import ctypes as C
import math
from enum import IntEnum
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError
from .engine import EPS0, MU0, Axis


class Field(IntEnum):
    """Field (dispersive.rs:463-470)."""
    Electric = 0
    Magnetic = 1


class Scheme(IntEnum):
    """Scheme (dispersive.rs:428-461)."""
    Trapezoidal = 0
    Leapfrog = 1


KIND_TRAPEZOIDAL, KIND_LEAPFROG = 0, 1
POLE_KINDS = ("lorentz", "drude", "debye")  # ems_pole_entry.kind


class Pole:
    """Pole (dispersive.rs:218-426). Frequencies in rad/s, times in s."""

    def __init__(self, kind: str, a: float, b: float, c: float = 0.0):
        self.kind = kind
        self.p = (float(a), float(b), float(c))

    @staticmethod
    def lorentz(omega_p: float, omega_0: float, gamma: float) -> "Pole":
        return Pole("lorentz", omega_p, omega_0, gamma)

    @staticmethod
    def drude(omega_p: float, gamma: float) -> "Pole":
        return Pole("drude", omega_p, gamma)

    @staticmethod
    def debye(delta_eps: float, tau: float) -> "Pole":
        return Pole("debye", delta_eps, tau)

    def __eq__(self, o):
        return isinstance(o, Pole) and (self.kind, self.p) == (o.kind, o.p)

    def __hash__(self):
        return hash((self.kind, self.p))

    def __repr__(self):
        return f"Pole.{self.kind}{self.p if self.kind == 'lorentz' else self.p[:2]}"

    def susceptibility(self, omega: float) -> complex:
        """chi(omega) (e^{jwt} convention, losses Im chi < 0)."""
        if self.kind == "lorentz":
            wp, w0, g = self.p
            return wp * wp / complex(w0 * w0 - omega * omega, omega * g)
        if self.kind == "drude":
            wp, g = self.p[:2]
            return -(wp * wp) / complex(omega * omega, -(omega * g))
        de, tau = self.p[:2]
        return de / complex(1.0, omega * tau)

    def discrete_susceptibility(self, f: float, dt: float) -> complex:
        """chi at the bilinear frequency (2/dt) tan(pi f dt) (dispersive.rs:288-290)."""
        return self.susceptibility(2.0 / dt * math.tan(math.pi * f * dt))

    def leapfrog_susceptibility(self, f: float, dt: float) -> complex:
        """dispersive.rs:409-425."""
        w = 2.0 * math.pi * f
        c = math.cos(0.5 * w * dt)
        om = 2.0 / dt * math.sin(0.5 * w * dt)
        if self.kind == "lorentz":
            wp, w0, g = self.p
            return wp * wp / complex(w0 * w0 - om * om, om * g * c)
        if self.kind == "drude":
            wp, g = self.p[:2]
            return wp * wp / complex(-(om * om), om * g * c)
        return self.discrete_susceptibility(f, dt)

    def scaled(self, s: float) -> "Pole":
        """Susceptibility times s (dispersive.rs:295-316)."""
        if self.kind == "lorentz":
            return Pole("lorentz", self.p[0] * math.sqrt(s), self.p[1], self.p[2])
        if self.kind == "drude":
            return Pole("drude", self.p[0] * math.sqrt(s), self.p[1])
        return Pole("debye", self.p[0] * s, self.p[1])

    def is_null(self) -> bool:
        return self.p[0] == 0.0

    def validate(self) -> None:
        n = 3 if self.kind == "lorentz" else 2
        for v in self.p[:n]:
            if not (math.isfinite(v) and v >= 0.0):
                raise EmsError(nat.EMS_E_INVALID,
                               f"dispersive pole {self!r}: {v} must be finite and >= 0")

    def _lrc(self, m: float):
        if self.kind == "lorentz":
            wp, w0, g = self.p
            l = 1.0 / (m * wp * wp)
            return l, g * l, w0 * w0 * l
        if self.kind == "drude":
            wp, g = self.p[:2]
            l = 1.0 / (m * wp * wp)
            return l, g * l, 0.0
        de, tau = self.p[:2]
        c = m * de
        return 0.0, tau / c, 1.0 / c

    def branch(self, m: float, dt: float):
        """(2L/dt, 1/C_p, z) (dispersive.rs:348-371)."""
        l, r, inv_c = self._lrc(m)
        a = 2.0 * l / dt
        return a, inv_c, a + r + 0.5 * dt * inv_c

    def leapfrog(self, m: float, dt: float):
        """(k_a, k_b, 1/C_p, L/2) or None for Debye (dispersive.rs:375-393)."""
        if self.kind == "debye":
            return None
        l, r, inv_c = self._lrc(m)
        d = l / dt + 0.5 * r
        return (l / dt - 0.5 * r) / d, 1.0 / d, inv_c, 0.5 * l


def _entry(field: int, axis: int, node, pole: "Pole"):
    """The ems_pole_entry of one pole on one edge."""
    return nat.PoleEntryC(int(field), int(axis), (C.c_int32 * 3)(*node),
                          POLE_KINDS.index(pole.kind), (C.c_double * 3)(*pole.p))


class Dispersive:
    """Dispersive (dispersive.rs:1790-2700): poles on E edges (and H dual
    edges); install with ``engine.add_operator``."""

    def __init__(self):
        self.entries: List[Tuple[int, int, Tuple[int, int, int], Pole]] = []
        self._scheme = Scheme.Trapezoidal
        self._installed = None
        self._skipped = 0

    @staticmethod
    def new() -> "Dispersive":
        return Dispersive()

    def with_scheme(self, s: Scheme) -> "Dispersive":
        self._scheme = Scheme(s)
        return self

    def scheme(self) -> Scheme:
        return self._scheme

    def push(self, field: Field, axis: int, node, pole: Pole) -> None:
        """dispersive.rs:1870-1880: null poles are ignored."""
        if not pole.is_null():
            self.entries.append((int(field), int(axis), tuple(int(v) for v in node), pole))

    def fill_cells(self, grid, cells, field: Field, poles: Sequence[Pole]) -> None:
        """dispersive.rs:1896-1986 (ems_dispersive_fill_cells): every edge
        touching the cell box gets the poles scaled by the share of its
        averaging volume inside."""
        for p in poles:
            p.validate()
        nc = grid.cells()
        cells = [(int(r[0]), int(r[1])) for r in cells]
        if any(cells[a][1] > nc[a] or cells[a][0] > cells[a][1] for a in range(3)):
            raise EmsError(nat.EMS_E_INVALID,
                           f"dispersive cell box {cells} does not fit the grid ({nc} cells)")
        if any(r[0] == r[1] for r in cells) or not poles:
            return
        src = (nat.PoleEntryC * len(poles))(*[_entry(int(field), 0, (0, 0, 0), p) for p in poles])
        box = (C.c_int32 * 6)(*[v for r in cells for v in r])
        cnt = C.c_uint64()
        args = (*grid._args(), box, int(field), src, len(poles))
        nat.check(nat.ems_dispersive_fill_cells(*args, None, 0, C.byref(cnt)))
        out = (nat.PoleEntryC * max(cnt.value, 1))()
        nat.check(nat.ems_dispersive_fill_cells(*args, out, cnt.value, C.byref(cnt)))
        for e in out[:cnt.value]:
            self.entries.append((int(e.field), int(e.axis), tuple(int(v) for v in e.node),
                                 Pole(POLE_KINDS[e.kind], e.p[0], e.p[1], e.p[2])))

    def fill_mask(self, grid, mask, field: Field, poles: Sequence[Pole]) -> None:
        """fill_cells for an arbitrary set of cells (mask: bool, cells-shaped):
        every electric edge gets the poles scaled by the share of its
        averaging area inside the set, summed in fill_cells' order (so a box
        mask gives fill_cells' entries bit for bit, in its order)."""
        if field != Field.Electric:
            raise EmsError(nat.EMS_E_UNSUPPORTED, "fill_mask: electric poles only")
        for p in poles:
            p.validate()
        nc = grid.cells()
        mask = np.asarray(mask, dtype=bool)
        if mask.shape != tuple(nc):
            raise EmsError(nat.EMS_E_INVALID, f"mask shape {mask.shape}, grid cells {nc}")
        if not mask.any():
            return
        n = grid.nodes()
        lo = [int(v.min()) for v in np.nonzero(mask)]
        hi = [int(v.max()) + 1 for v in np.nonzero(mask)]
        deltas = [np.array([grid.delta(a, q) for q in range(nc[a])]) for a in range(3)]
        for axis in Axis:
            ab, ac = axis.others()
            a, b, c = int(axis), int(ab), int(ac)
            # edge nodes: pa over the mask's cells along a, pb / pc over the
            # nodes that touch its cells
            pa = np.arange(lo[a], hi[a])
            pb = np.arange(lo[b], min(hi[b] + 1, n[b]))
            pc = np.arange(lo[c], min(hi[c] + 1, n[c]))
            m = np.transpose(mask, (a, b, c))

            def cell(d, q, t):
                ok = (q >= 0) & (q < nc[t])
                return np.where(ok, d[np.clip(q, 0, nc[t] - 1)], 0.0), ok
            w = np.zeros((len(pb), len(pc)))
            terms = []
            for cbo in (-1, 0):
                db_, okb = cell(deltas[b], pb + cbo, b)
                for cco in (-1, 0):
                    dc_, okc = cell(deltas[c], pc + cco, c)
                    area = db_[:, None] * dc_[None, :]
                    w = w + area
                    cbi = np.clip(pb + cbo, 0, nc[b] - 1)
                    cci = np.clip(pc + cco, 0, nc[c] - 1)
                    inside = m[np.ix_(pa, cbi, cci)] & (okb[:, None] & okc[None, :])[None]
                    terms.append(np.where(inside, area[None], 0.0))
            w_in = ((terms[0] + terms[1]) + terms[2]) + terms[3]
            share = np.where(w[None] > 0.0, w_in / np.where(w[None] > 0.0, w[None], 1.0), 0.0)
            for ia, ib, ic in zip(*np.nonzero(share > 0.0)):
                node = [0, 0, 0]
                node[a], node[b], node[c] = int(pa[ia]), int(pb[ib]), int(pc[ic])
                sh = float(share[ia, ib, ic])
                for p in poles:
                    self.push(field, axis, node, p.scaled(sh))

    def __len__(self) -> int:
        return len(self.entries)

    def _plan(self, grid, dt: float, lanes: np.ndarray) -> dict:
        """Dispersive::install (dispersive.rs:2063-2319) and the side layout
        of slab_hooks (:2609-2770), by the host library
        (ems_dispersive_plan)."""
        if self._installed is not None:
            raise EmsError(nat.EMS_E_INVALID, "dispersive medium already installed")
        n = grid.nodes()
        for _, _, _, p in self.entries:
            p.validate()
        for field, axis, node, _ in self.entries:
            a = axis
            b, c = (int(x) for x in Axis(axis).others())
            ok = all(node[t] < n[t] for t in range(3)) and (
                node[a] + 1 < n[a] if field == Field.Electric else
                (node[b] + 1 < n[b] and node[c] + 1 < n[c]))
            if not ok:
                raise EmsError(nat.EMS_E_INVALID,
                               f"dispersive {Field(field).name} edge {Axis(axis).name}:{node} "
                               f"is not in the grid ({n} nodes)")
        ne = len(self.entries)
        cap = max(ne, 1)
        ents = (nat.PoleEntryC * cap)(*[_entry(f, a, nd, p) for f, a, nd, p in self.entries])
        keep = {}

        def buf(name, dtype, size):
            keep[name] = np.zeros(size, dtype=dtype)
            return keep[name]
        plan = nat.PolePlanC()
        for f in (0, 1):
            plan.axis[f] = nat.ptr(buf(f"axis{f}", np.int32, cap), C.c_int32)
            plan.nodes[f] = nat.ptr(buf(f"nodes{f}", np.int32, 3 * cap), C.c_int32)
            plan.sum_inv_z[f] = nat.ptr(buf(f"siz{f}", np.float64, cap), C.c_double)
            plan.pole_end[f] = nat.ptr(buf(f"pe{f}", np.uint32, cap), C.c_uint32)
            plan.cidx[f] = nat.ptr(buf(f"cidx{f}", np.uint32, cap), C.c_uint32)
        for name in ("a", "inv_c", "inv_z", "half_l"):
            setattr(plan, name, nat.ptr(buf(name, np.float64, cap), C.c_double))
        plan.kind = nat.ptr(buf("kind", np.int32, cap), C.c_int32)
        dense = np.ascontiguousarray(lanes, dtype=np.float32)
        nat.check(nat.ems_dispersive_plan(*grid._args(), float(dt), nat.ptr(dense, C.c_float),
                                          int(self._scheme), ents, ne, C.byref(plan)))
        sides: Dict[int, dict] = {}
        for f in (0, 1):
            m, q = plan.n_edges[f], plan.n_poles[f]
            nd = keep[f"nodes{f}"][:3 * m].reshape(-1, 3)
            sides[f] = {"axis": [int(v) for v in keep[f"axis{f}"][:m]],
                        "nodes": [tuple(int(v) for v in r) for r in nd],
                        "sum_inv_z": [float(v) for v in keep[f"siz{f}"][:m]],
                        "pole_end": [int(v) for v in keep[f"pe{f}"][:m]],
                        "cidx": [int(v) for v in keep[f"cidx{f}"][:q]]}
        k = plan.n_consts
        consts = [(float(keep["a"][i]), float(keep["inv_c"][i]), float(keep["inv_z"][i]),
                   float(keep["half_l"][i]), int(keep["kind"][i])) for i in range(k)]
        self._skipped = int(plan.skipped)
        return {"sides": sides, "consts": consts, "skipped": self._skipped}

    # -- results
    def _bind(self, engine, oid: int, plan: dict) -> None:
        self._installed = (engine, oid, plan)

    def edges(self) -> int:
        if self._installed is None:
            return 0
        s = self._installed[2]["sides"]
        return len(s[0]["axis"]) + len(s[1]["axis"])

    def skipped(self) -> int:
        return self._skipped

    def state(self, field: Field):
        """(j, q) per pole of one side, in pole order."""
        eng, oid, _ = self._installed
        return eng._dispersive_state(oid, int(field))

    def _sum_energy(self, field: Field) -> float:
        """Side::add_energy (dispersive.rs:1489-1517): from 0.0, pole order."""
        if self._installed is None:
            return 0.0
        _, _, plan = self._installed
        j, q = self.state(field)
        consts = plan["consts"]
        w = 0.0
        for p, ci in enumerate(plan["sides"][int(field)]["cidx"]):
            a, inv_c, inv_z, half_l, kind = consts[ci]
            w += half_l * float(j[p]) * float(j[p]) + 0.5 * inv_c * float(q[p]) * float(q[p])
        return w

    def electric_energy(self) -> float:
        return self._sum_energy(Field.Electric)

    def magnetic_energy(self) -> float:
        return self._sum_energy(Field.Magnetic)

    def energy(self) -> float:
        """Operator::energy (dispersive.rs:2789-2794)."""
        return self.electric_energy() + self.magnetic_energy()


class PoleSide(C.Structure):
    """ems_pole_side (include/ems_gpu.h)."""
    _fields_ = [("n_edges", C.c_int32), ("axis", C.POINTER(C.c_int32)),
                ("nodes", C.POINTER(C.c_int32)), ("sum_inv_z", C.POINTER(C.c_double)),
                ("pole_end", C.POINTER(C.c_uint32)), ("n_poles", C.c_uint64),
                ("cidx", C.POINTER(C.c_uint32))]


class PoleConsts(C.Structure):
    """ems_pole_consts (include/ems_gpu.h)."""
    _fields_ = [("n", C.c_int32), ("a", C.POINTER(C.c_double)), ("inv_c", C.POINTER(C.c_double)),
                ("inv_z", C.POINTER(C.c_double)), ("half_l", C.POINTER(C.c_double)),
                ("kind", C.POINTER(C.c_int32))]


def make_desc(plan: dict):
    """C descriptions of the two sides and the constants, and the arrays
    they point into."""
    keep = []

    def arr(x, t):
        a = np.ascontiguousarray(np.asarray(x, dtype=t))
        if a.size == 0:
            a = np.zeros(1, dtype=t)
        keep.append(a)
        return a

    def p(a, ct):
        return a.ctypes.data_as(C.POINTER(ct))

    sides = []
    for f in (0, 1):
        s = plan["sides"][f]
        ax = arr(s["axis"], np.int32)
        nd = arr([v for n in s["nodes"] for v in n], np.int32)
        sz = arr(s["sum_inv_z"], np.float64)
        pe = arr(s["pole_end"], np.uint32)
        ci = arr(s["cidx"], np.uint32)
        sides.append(PoleSide(len(s["axis"]), p(ax, C.c_int32), p(nd, C.c_int32),
                              p(sz, C.c_double), p(pe, C.c_uint32), len(s["cidx"]),
                              p(ci, C.c_uint32)))
    cs = plan["consts"]
    cols = [arr([c[q] for c in cs], np.float64) for q in range(4)]
    kind = arr([c[4] for c in cs], np.int32)
    consts = PoleConsts(len(cs), *(p(c, C.c_double) for c in cols), p(kind, C.c_int32))
    return sides[0], sides[1], consts, keep
