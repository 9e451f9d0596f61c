"""Lumped elements and ports (src/engine/lumped.rs, src/engine/port.rs).

The host side of `LumpedElement::install` (lumped.rs:829-1012) runs here:
per-edge values from the box or the edge list, R and C folded into the
ca/cb lanes with ``fold_parallel`` (C++ host library, lumped.rs:275-282),
branch constants, and the device description handed to
``ems_gpu_add_lumped``. Source time functions are tabulated by step at
h_time = (n + 1/2)dt (operator.rs:100-102); a callable is tabulated lazily,
ahead of every run, so any function of time is exact.
"""
from __future__ import annotations

import ctypes as C
import math
from enum import IntEnum
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError
from .engine import Axis, Waveform
from . import post


def fold_parallel(ca: float, cb: float, dt: float, g: float, c: float) -> Tuple[float, float]:
    """fold_parallel (lumped.rs:275-282) in the C++ host library."""
    a = C.c_float()
    b = C.c_float()
    rc = nat.ems_fold_parallel(float(np.float32(ca)), float(np.float32(cb)), float(dt), float(g),
                               float(c), C.byref(a), C.byref(b))
    if rc != 0:
        raise EmsError(rc, "cannot fold a lumped element into a PEC edge")
    return a.value, b.value


class LumpedBox:
    """LumpedBox (lumped.rs:138-222): nodes lo..=hi directed along axis."""

    def __init__(self, axis: int, lo, hi):
        axis = Axis(axis)
        lo = tuple(int(v) for v in lo)
        hi = tuple(int(v) for v in hi)
        a = int(axis)
        if lo[a] >= hi[a]:
            raise EmsError(nat.EMS_E_INVALID,
                           f"lumped box {lo}..={hi} is degenerate along its direction {axis.name}")
        if any(lo[t] > hi[t] for t in range(3)):
            raise EmsError(nat.EMS_E_INVALID, f"lumped box {lo}..={hi} has lo > hi")
        self.axis, self.lo, self.hi = axis, lo, hi

    def series_edges(self) -> int:
        a = int(self.axis)
        return self.hi[a] - self.lo[a]

    def columns(self) -> int:
        b, c = (int(x) for x in self.axis.others())
        return (self.hi[b] - self.lo[b] + 1) * (self.hi[c] - self.lo[c] + 1)

    def edges(self) -> List[Tuple[int, int, int]]:
        """Start nodes, column by column (lumped.rs:193-211)."""
        a = int(self.axis)
        b, c = (int(x) for x in self.axis.others())
        out = []
        for pb in range(self.lo[b], self.hi[b] + 1):
            for pc in range(self.lo[c], self.hi[c] + 1):
                for pa in range(self.lo[a], self.hi[a]):
                    p = [0, 0, 0]
                    p[a], p[b], p[c] = pa, pb, pc
                    out.append(tuple(p))
        return out

    def check_fits(self, nodes) -> None:
        if any(self.hi[t] >= nodes[t] for t in range(3)):
            raise EmsError(nat.EMS_E_INVALID,
                           f"lumped box {self.lo}..={self.hi} exceeds the grid ({tuple(nodes)} nodes)")


class Topology(IntEnum):
    """Topology (lumped.rs:225-233)."""
    Parallel = 0
    Series = 1


class EdgeValues:
    """EdgeValues (lumped.rs:286-318): one edge's own R, L, C."""

    def __init__(self, node, r: Optional[float] = None, l: Optional[float] = None,
                 c: Optional[float] = None):
        self.node = tuple(int(v) for v in node)
        self.r, self.l, self.c = r, l, c

    def present(self):
        return (self.r is not None, self.l is not None, self.c is not None)


class _Source:
    """Source (lumped.rs:236-262): a callable of t, or amplitude * waveform."""

    def __init__(self, fn: Optional[Callable[[float], float]] = None,
                 waveform: Optional[Waveform] = None, amplitude: float = 1.0):
        self.fn, self.waveform, self.amplitude = fn, waveform, float(amplitude)

    def value(self, t: float, dt: float) -> float:
        if self.fn is not None:
            return float(self.fn(t))
        return self.amplitude * self.waveform.value(t, dt)

    def support_steps(self, dt: float) -> Optional[int]:
        """Steps after which the value is 0 for good (None: unbounded)."""
        if self.fn is not None:
            return None
        return int(math.floor(self.waveform.duration() / dt)) + 2

    def validate(self) -> None:
        if self.fn is None and not math.isfinite(self.amplitude):
            raise EmsError(nat.EMS_E_INVALID, f"source amplitude {self.amplitude} is not finite")


def _check_value(name: str, v: float) -> None:
    if not (math.isfinite(v) and v > 0.0):
        raise EmsError(nat.EMS_E_INVALID, f"lumped {name} = {v} must be finite and > 0")


class LumpedElement:
    """LumpedElement (lumped.rs:407-1087): R, L, C and Thevenin sources on a
    box or an explicit edge list. Install with ``engine.add_operator``."""

    def __init__(self, region: Optional[LumpedBox] = None):
        self._box = region
        self._edge_list: Optional[List[EdgeValues]] = None
        self._axis = region.axis if region is not None else Axis.X
        self._series = region.series_edges() if region is not None else 0
        self.r_val = self.l_val = self.c_val = None
        self._topology = Topology.Parallel
        self._source: Optional[_Source] = None
        self._edge_sources: List[Tuple[_Source, List[float]]] = []
        self._drive_sources: List[Tuple[_Source, List[float]]] = []
        self._record = False
        self._explicit = False
        self._installed = None

    @staticmethod
    def new(region: LumpedBox) -> "LumpedElement":
        return LumpedElement(region)

    @staticmethod
    def resistor(region: LumpedBox, r: float) -> "LumpedElement":
        return LumpedElement(region).r(r)

    @staticmethod
    def from_edges(axis: int, topology: Topology, edges: Sequence[EdgeValues]) -> "LumpedElement":
        """lumped.rs:487-526."""
        edges = list(edges)
        if not edges:
            raise EmsError(nat.EMS_E_INVALID, "lumped element has no edges")
        pat = edges[0].present()
        if any(e.present() != pat for e in edges):
            raise EmsError(nat.EMS_E_INVALID,
                           "lumped element edges must all have the same set of R, L, C")
        nodes = sorted(e.node for e in edges)
        if any(nodes[q] == nodes[q + 1] for q in range(len(nodes) - 1)):
            raise EmsError(nat.EMS_E_INVALID, "lumped element lists an edge twice")
        a = int(axis)
        e = LumpedElement(None)
        e._axis = Axis(axis)
        e._edge_list = edges
        e._series = len({n[a] for n in nodes})
        e._topology = Topology(topology)
        return e

    # -- builder
    def r(self, r: float) -> "LumpedElement":
        self.r_val = float(r)
        return self

    def l(self, l: float) -> "LumpedElement":  # noqa: E743
        self.l_val = float(l)
        return self

    def c(self, c: float) -> "LumpedElement":
        self.c_val = float(c)
        return self

    def topology(self, t: Topology) -> "LumpedElement":
        self._topology = Topology(t)
        return self

    def source(self, fn: Callable[[float], float]) -> "LumpedElement":
        self._source = _Source(fn=fn)
        return self

    def waveform(self, w: Waveform, amplitude: float) -> "LumpedElement":
        self._source = _Source(waveform=w, amplitude=amplitude)
        return self

    def edge_source(self, weights: Sequence[float], fn: Callable[[float], float]) -> "LumpedElement":
        self._edge_sources.append((_Source(fn=fn), [float(w) for w in weights]))
        return self

    def edge_source_waveform(self, weights: Sequence[float], w: Waveform,
                             amplitude: float) -> "LumpedElement":
        self._edge_sources.append((_Source(waveform=w, amplitude=amplitude),
                                   [float(x) for x in weights]))
        return self

    def branch_drive_source(self, weights: Sequence[float],
                            fn: Callable[[float], float]) -> "LumpedElement":
        self._drive_sources.append((_Source(fn=fn), [float(w) for w in weights]))
        return self

    def record_current(self) -> "LumpedElement":
        self._record = True
        return self

    def explicit_inductor(self, on: bool) -> "LumpedElement":
        self._explicit = bool(on)
        return self

    # -- properties
    def axis(self) -> Axis:
        return self._axis

    def edge_nodes(self) -> List[Tuple[int, int, int]]:
        if self._box is not None:
            return self._box.edges()
        return [e.node for e in self._edge_list]

    def series_edges(self) -> int:
        return self._series

    def region(self) -> Optional[LumpedBox]:
        return self._box

    def _present(self):
        if self._box is not None:
            return (self.r_val is not None, self.l_val is not None, self.c_val is not None)
        return self._edge_list[0].present()

    def has_source(self) -> bool:
        return self._source is not None or bool(self._edge_sources)

    def has_branch(self) -> bool:
        _, l, c = self._present()
        return l if self._topology == Topology.Parallel else (l or c)

    def validate(self) -> None:
        """lumped.rs:753-808."""
        if self._box is not None:
            for name, v in (("R", self.r_val), ("L", self.l_val), ("C", self.c_val)):
                if v is not None:
                    _check_value(name, v)
        else:
            for e in self._edge_list:
                for name, v in (("R", e.r), ("L", e.l), ("C", e.c)):
                    if v is not None:
                        _check_value(name, v)
        r, l, c = self._present()
        if not (r or l or c):
            raise EmsError(nat.EMS_E_INVALID, "lumped element needs at least one of R, L, C")
        if self.has_source() and self._topology == Topology.Parallel and not r:
            raise EmsError(nat.EMS_E_INVALID,
                           "a source on a parallel lumped element needs a resistance")
        if self._source is not None:
            self._source.validate()
        ne = len(self.edge_nodes())
        if self._drive_sources and not self.has_branch():
            raise EmsError(nat.EMS_E_INVALID,
                           "a branch drive source needs an inductor or a series branch")
        for s, w in self._edge_sources + self._drive_sources:
            s.validate()
            if len(w) != ne:
                raise EmsError(nat.EMS_E_INVALID, f"edge source has {len(w)} weights for {ne} edges")
            bad = [x for x in w if not math.isfinite(x)]
            if bad:
                raise EmsError(nat.EMS_E_INVALID, f"edge source weight {bad[0]} is not finite")

    def _check_fits(self, nodes) -> None:
        if self._box is not None:
            self._box.check_fits(nodes)
            return
        a = int(self._axis)
        for e in self._edge_list:
            n = e.node
            if any(n[t] >= nodes[t] for t in range(3)) or n[a] + 1 >= nodes[a]:
                raise EmsError(nat.EMS_E_INVALID,
                               f"lumped edge {self._axis.name}:{n} is outside the grid")

    def _plan(self, grid, dt: float, lanes: np.ndarray) -> dict:
        """LumpedElement::install (lumped.rs:880-1012) on the host, by the
        host library (ems_lumped_plan): folds R and C into copies of the
        axis's ca/cb lanes and returns them with the device description."""
        if self._installed is not None:
            raise EmsError(nat.EMS_E_INVALID, "lumped element already installed")
        self.validate()
        nodes = grid.nodes()
        self._check_fits(nodes)
        edges = self.edge_nodes()
        a = int(self._axis)
        shape = lanes[a].shape
        ca = np.ascontiguousarray(lanes[a].reshape(-1), dtype=np.float32).copy()
        cb = np.ascontiguousarray(lanes[3 + a].reshape(-1), dtype=np.float32).copy()
        for p in edges:
            if cb[(p[0] * nodes[1] + p[1]) * nodes[2] + p[2]] == 0.0:
                raise EmsError(nat.EMS_E_INVALID,
                               f"lumped element edge {self._axis.name}:{p} is PEC "
                               "(metal or a PEC outer face)")
        ne = len(edges)
        nd = np.ascontiguousarray(np.array(edges, dtype=np.int32).reshape(-1))
        nan = math.nan
        val = lambda v: nan if v is None else float(v)  # noqa: E731
        el = nat.LumpedElementC()
        el.axis, el.n_edges, el.nodes = a, ne, nat.ptr(nd, C.c_int32)
        el.topology, el.explicit_inductor = int(self._topology), 1 if self._explicit else 0
        keep = [nd]
        if self._box is not None:
            el.series, el.columns = self._box.series_edges(), self._box.columns()
            el.r, el.l, el.c = val(self.r_val), val(self.l_val), val(self.c_val)
        else:
            el.series, el.columns = self._series, 0
            for name, attr in (("edge_r", "r"), ("edge_l", "l"), ("edge_c", "c")):
                arr = np.ascontiguousarray([val(getattr(e, attr)) for e in self._edge_list],
                                           dtype=np.float64)
                keep.append(arr)
                setattr(el, name, nat.ptr(arr, C.c_double))
        outs = {k: np.zeros(ne, dtype=np.float64)
                for k in ("cb_e", "g_src", "rec_g", "rec_c", "two_l_dt", "inv_c", "z")}
        inst = nat.LumpedInstallC()
        for k, v in outs.items():
            setattr(inst, k, nat.ptr(v, C.c_double))
        nat.check(nat.ems_lumped_plan((C.c_int32 * 3)(*nodes), float(dt), C.byref(el),
                                      nat.ptr(ca, C.c_float), nat.ptr(cb, C.c_float),
                                      C.byref(inst)))
        branch = int(inst.branch)
        listed = lambda k: [float(v) for v in outs[k]]  # noqa: E731
        terms: List[Tuple[_Source, List[float]]] = []
        if self._source is not None:
            terms.append((self._source, [1.0 / self._series] * ne))
        terms += self._edge_sources
        return {"ca": ca.reshape(shape), "cb": cb.reshape(shape), "axis": a, "edges": edges,
                "cb_e": listed("cb_e"), "g_src": listed("g_src"), "series": self._series,
                "src_in_branch": bool(inst.src_in_branch), "terms": terms,
                "drive": list(self._drive_sources), "branch": branch,
                "two_l_dt": listed("two_l_dt") if branch else None,
                "inv_c": listed("inv_c") if branch else None,
                "z": listed("z") if branch else None,
                "rec_g": listed("rec_g") if self._record else None,
                "rec_c": listed("rec_c") if self._record else None,
                "columns": 0}

    # -- results
    def _bind(self, engine, oid: int, plan: dict) -> None:
        self._installed = (engine, oid, plan)

    def current(self) -> np.ndarray:
        """Recorded terminal currents i^{n+1/2} along +axis (lumped.rs:1042-1047)."""
        if self._installed is None:
            return np.empty(0)
        eng, oid, _ = self._installed
        return eng._lumped_values(oid, 0)[0]

    def current_t0(self) -> float:
        if self._installed is None:
            return 0.0
        eng, oid, _ = self._installed
        return (eng._lumped_values(oid, 0)[1] + 0.5) * eng.dt

    def branch_state(self):
        eng, oid, _ = self._installed
        return eng._lumped_branch(oid)

    def branch_current(self) -> float:
        """Sum of j-bar over the edges / M (lumped.rs:1033-1040)."""
        if self._installed is None or not self._installed[2]["branch"]:
            return 0.0
        _, _, jav = self.branch_state()
        s = -0.0
        for x in jav:
            s += float(x)
        return s / self._series

    def branch_energy(self) -> float:
        """sum 1/2 L j^2 + 1/2 q^2 / C (lumped.rs:1075-1087)."""
        if self._installed is None or not self._installed[2]["branch"]:
            return 0.0
        eng, _, plan = self._installed
        j, q, _ = self.branch_state()
        s = -0.0
        for e in range(len(j)):
            half_l = 0.25 * plan["two_l_dt"][e] * eng.dt
            s += half_l * float(j[e]) * float(j[e]) + 0.5 * plan["inv_c"][e] * float(q[e]) * float(q[e])
        return s


class LumpedPort:
    """LumpedPort (engine/port.rs:94-229): a resistor with an optional
    Thevenin source, recording V (after E, at (n+1)dt) and I (into the
    structure, at (n+1/2)dt) every step."""

    def __init__(self, region: LumpedBox, r: float):
        self.region = region
        self.element = LumpedElement.resistor(region, r).record_current()
        self.resistance = float(r)
        self.dt = 0.0
        self._dfts = {}  # frequencies -> (V, I) running DFT ids on the device

    @staticmethod
    def new(region: LumpedBox, r: float) -> "LumpedPort":
        return LumpedPort(region, r)

    def excite(self, fn: Callable[[float], float]) -> "LumpedPort":
        self.element._source = _Source(fn=fn)
        return self

    def excite_waveform(self, w: Waveform, amplitude: float) -> "LumpedPort":
        self.element._source = _Source(waveform=w, amplitude=amplitude)
        return self

    def _plan(self, grid, dt: float, lanes: np.ndarray) -> dict:
        self.region.check_fits(grid.nodes())
        p = self.element._plan(grid, dt, lanes)
        p["columns"] = self.region.columns()
        self.dt = dt
        return p

    def _bind(self, engine, oid: int, plan: dict) -> None:
        self.element._bind(engine, oid, plan)
        self._installed = (engine, oid)

    def voltage(self) -> np.ndarray:
        eng, oid = self._installed
        return eng._lumped_values(oid, 1)[0]

    def current(self) -> np.ndarray:
        eng, oid = self._installed
        return eng._lumped_values(oid, 2)[0]

    def _first(self) -> int:
        eng, oid = self._installed
        return eng._lumped_values(oid, 1)[1]

    def voltage_t0(self) -> float:
        return (self._first() + 1.0) * self.dt

    def current_t0(self) -> float:
        return (self._first() + 0.5) * self.dt

    def voltage_times(self) -> np.ndarray:
        t0 = self.voltage_t0()
        return np.array([t0 + n * self.dt for n in range(len(self.voltage()))])

    def current_times(self) -> np.ndarray:
        t0 = self.current_t0()
        return np.array([t0 + n * self.dt for n in range(len(self.current()))])

    def branch_energy(self) -> float:
        return self.element.branch_energy()

    def analyze(self, freqs) -> post.PortResult:
        return self.analyze_with(freqs, complex(self.resistance, 0.0))

    def analyze_with(self, freqs, z_ref) -> post.PortResult:
        """port.rs:209-219."""
        v = self.voltage()
        if len(v) == 0:
            raise EmsError(nat.EMS_E_INVALID, "port has no samples yet")
        vf = post.dft(self.voltage_t0(), self.dt, v, freqs)
        i_f = post.dft(self.current_t0(), self.dt, self.current(), freqs)
        return post.PortResult(freqs, vf, i_f, z_ref)

    def track(self, freqs) -> None:
        """Starts running DFTs of the port voltage and current on the device
        at `freqs` (ems_gpu_add_dft); `analyze_device` then reads them."""
        eng, oid = self._installed
        key = tuple(float(f) for f in freqs)
        if key not in self._dfts:
            self._dfts[key] = (eng._add_dft(oid, 1, key, self.voltage_t0(), self.dt),
                               eng._add_dft(oid, 2, key, self.current_t0(), self.dt))

    def update_device(self, freqs) -> None:
        """Queues the device DFTs' accumulation of the samples recorded so
        far (no wait): queue every port's first, then read them."""
        eng, _ = self._installed
        self.track(freqs)
        for did in self._dfts[tuple(float(f) for f in freqs)]:
            eng._dft_update(did)

    def analyze_device(self, freqs, z_ref=None) -> post.PortResult:
        """analyze_with from the device's running DFTs of the port V and I:
        the same spectra as the host's transform of the recorded series, bit
        for bit, without copying the series back."""
        eng, _ = self._installed
        self.track(freqs)
        vid, iid = self._dfts[tuple(float(f) for f in freqs)]
        eng._dft_update(vid)
        eng._dft_update(iid)
        vf, nv = eng._dft_result(vid, len(freqs))
        i_f, ni = eng._dft_result(iid, len(freqs))
        if nv == 0:
            raise EmsError(nat.EMS_E_INVALID, "port has no samples yet")
        zr = complex(self.resistance, 0.0) if z_ref is None else z_ref
        return post.PortResult(list(freqs), vf, i_f, zr)


class LumpedDesc(C.Structure):
    """ems_lumped_desc (include/ems_gpu.h)."""
    _fields_ = [("axis", C.c_int32), ("n_edges", C.c_int32), ("nodes", C.POINTER(C.c_int32)),
                ("cb", C.POINTER(C.c_double)), ("g_src", C.POINTER(C.c_double)),
                ("series", C.c_int32), ("src_in_branch", C.c_int32),
                ("n_terms", C.c_int32), ("term_weight", C.POINTER(C.c_double)),
                ("n_drive", C.c_int32), ("drive_weight", C.POINTER(C.c_double)),
                ("branch", C.c_int32), ("two_l_dt", C.POINTER(C.c_double)),
                ("inv_c", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
                ("rec_g", C.POINTER(C.c_double)), ("rec_c", C.POINTER(C.c_double)),
                ("port_columns", C.c_int32)]


def make_desc(plan: dict, struct=LumpedDesc):
    """The C description of a plan, and the arrays it points into."""
    keep = []

    def arr(x, t=np.float64):
        if x is None:
            return None
        a = np.ascontiguousarray(np.asarray(x, dtype=t))
        keep.append(a)
        return a

    ctype = {np.float64: C.c_double, np.int32: C.c_int32}
    ptr = lambda a, t: None if a is None else a.ctypes.data_as(C.POINTER(ctype[t]))  # noqa: E731
    nodes = arr([v for p in plan["edges"] for v in p], np.int32)
    tw = arr([w for _, ws in plan["terms"] for w in ws]) if plan["terms"] else None
    dw = arr([w for _, ws in plan["drive"] for w in ws]) if plan["drive"] else None
    d = struct(plan["axis"], len(plan["edges"]), ptr(nodes, np.int32),
               ptr(arr(plan["cb_e"]), np.float64), ptr(arr(plan["g_src"]), np.float64),
               plan["series"], 1 if plan["src_in_branch"] else 0,
               len(plan["terms"]), ptr(tw, np.float64), len(plan["drive"]), ptr(dw, np.float64),
               plan["branch"], ptr(arr(plan["two_l_dt"]), np.float64),
               ptr(arr(plan["inv_c"]), np.float64), ptr(arr(plan["z"]), np.float64),
               ptr(arr(plan["rec_g"]), np.float64), ptr(arr(plan["rec_c"]), np.float64),
               plan["columns"])
    return d, keep


class SourceTables:
    """Step tables of a lumped element's source terms (source terms, then
    drive terms), kept ahead of the engine's step count."""

    def __init__(self, plan: dict, dt: float):
        self.sources = [s for s, _ in plan["terms"]] + [s for s, _ in plan["drive"]]
        self.dt = dt
        self.len = [0] * len(self.sources)
        self.vals: List[np.ndarray] = [np.empty(0)] * len(self.sources)

    def ensure(self, upto: int, upload) -> None:
        """Tables cover steps [0, upto); `upload(term, values)` sends one."""
        for t, s in enumerate(self.sources):
            sup = s.support_steps(self.dt)
            want = sup if sup is not None else upto
            if sup is None and self.len[t] >= upto:
                continue
            if sup is not None and self.len[t] >= sup:
                continue
            if sup is None:
                want = max(upto, 2 * self.len[t], 1024)
            old = self.vals[t]
            new = np.empty(want, dtype=np.float64)
            new[:old.size] = old
            if s.fn is None:
                # amplitude * waveform(t) with t = (n + 1/2)dt, tabulated in C++
                new[old.size:] = s.amplitude * s.waveform.values(self.dt, 0.5, want)[old.size:]
            else:
                for n in range(old.size, want):
                    new[n] = s.value((n + 0.5) * self.dt, self.dt)
            self.vals[t] = new
            self.len[t] = want
            upload(t, new)
