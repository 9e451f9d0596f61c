"""Host-side mirror of the reference engine API (src/engine/mod.rs), driving
the MI355X step engine of include/ems_gpu.h.

Names, argument meanings and error behaviour follow the reference:
``Engine.builder(grid).materials(m).faces(f).timestep(t).build()``,
``engine.add_operator(op)``, ``engine.run(n)``, ``engine.step_with_energy()``,
``engine.run_until(limits)``. Compile-side work (coefficients, CPML layers,
waveform tables) runs in the C++ host library (include/ems_host.h); stepping
runs only on the device — there is no CPU path here.

Every array handed to the device is also kept in ``engine.setup`` (the
"compiled operators"), in registration order, so the same arrays can be
replayed elsewhere.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError

# src/constants.rs
C0 = 299_792_458.0
MU0 = 1.256_637_061_27e-6
EPS0 = 1.0 / (MU0 * C0 * C0)
Z0 = MU0 * C0


class Axis(IntEnum):
    """src/axis.rs"""
    X = 0
    Y = 1
    Z = 2

    def next(self) -> "Axis":
        return Axis((self + 1) % 3)

    def prev(self) -> "Axis":
        return Axis((self + 2) % 3)

    def others(self) -> Tuple["Axis", "Axis"]:
        return (self.next(), self.next().next())


class Face(IntEnum):
    """src/fdtd.rs Face: x_min, x_max, y_min, y_max, z_min, z_max."""
    XMin = 0
    XMax = 1
    YMin = 2
    YMax = 3
    ZMin = 4
    ZMax = 5

    @property
    def axis(self) -> Axis:
        return Axis(self // 2)

    @property
    def is_max(self) -> bool:
        return self % 2 == 1

    @staticmethod
    def all() -> List["Face"]:
        return list(Face)


class FaceCondition(IntEnum):
    """src/engine/coeffs.rs:36-48"""
    Pec = 0
    Pmc = 1


# ---------------------------------------------------------------- mesh --

class Grid:
    """Rectilinear Cartesian mesh (src/engine/grid.rs)."""

    def __init__(self, lines: Sequence[Sequence[float]]):
        ls = [np.ascontiguousarray(np.asarray(l, dtype=np.float64)) for l in lines]
        if len(ls) != 3:
            raise EmsError(nat.EMS_E_INVALID, "a grid needs three axes")
        for a, l in enumerate(ls):
            if l.size < 2 or not np.all(np.isfinite(l)) or not np.all(np.diff(l) > 0):
                raise EmsError(nat.EMS_E_INVALID,
                               f"mesh lines of axis {a} must be finite, increasing, at least two")
        self.lines = ls

    @staticmethod
    def uniform(cells, cell_size) -> "Grid":
        """Grid::uniform (grid.rs:514-518): line i at i*size."""
        return Grid([[i * float(cell_size[a]) for i in range(int(cells[a]) + 1)] for a in range(3)])

    def nodes(self) -> Tuple[int, int, int]:
        return tuple(int(l.size) for l in self.lines)

    def num_nodes(self) -> int:
        n = self.nodes()
        return n[0] * n[1] * n[2]

    def cells(self) -> Tuple[int, int, int]:
        n = self.nodes()
        return (n[0] - 1, n[1] - 1, n[2] - 1)

    def num_cells(self) -> int:
        c = self.cells()
        return c[0] * c[1] * c[2]

    def index(self, node) -> int:
        n = self.nodes()
        return (node[0] * n[1] + node[1]) * n[2] + node[2]

    def delta(self, axis: int, i: int) -> float:
        l = self.lines[axis]
        return float(l[i + 1] - l[i])

    def dual_delta(self, axis: int, i: int) -> float:
        """Grid::dual_delta (grid.rs:580-597): mean of the neighbouring
        primary cells, a missing one (outside the grid) counting 0."""
        n = self.lines[axis].size
        lo = self.delta(axis, i - 1) if i > 0 else 0.0
        hi = self.delta(axis, i) if i + 1 < n else 0.0
        return 0.5 * (lo + hi)

    def _args(self):
        x, y, z = self.lines
        return (nat.ptr(x, C.c_double), x.size, nat.ptr(y, C.c_double), y.size,
                nat.ptr(z, C.c_double), z.size)


@dataclass
class Material:
    """src/engine/material.rs:22-31 (per-axis values)."""
    eps_r: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    mu_r: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    kappa: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    sigma_m: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    @staticmethod
    def vacuum() -> "Material":
        return Material()

    @staticmethod
    def dielectric(eps_r: float) -> "Material":
        return Material(eps_r=(eps_r,) * 3)

    def with_kappa(self, k: float) -> "Material":
        return Material(self.eps_r, self.mu_r, (k,) * 3, self.sigma_m)

    def with_mu_r(self, m: float) -> "Material":
        return Material(self.eps_r, (m,) * 3, self.kappa, self.sigma_m)

    def with_sigma_m(self, s: float) -> "Material":
        return Material(self.eps_r, self.mu_r, self.kappa, (s,) * 3)

    def row(self) -> List[float]:
        return list(self.eps_r) + list(self.mu_r) + list(self.kappa) + list(self.sigma_m)


class CellMaterials:
    """Per-cell materials (src/engine/material.rs:111-270): a material table,
    a u16 id per cell and a PEC mask per E edge."""

    def __init__(self, grid: Grid, background: Material = None):
        self.grid = grid
        self.table: List[Material] = [background or Material.vacuum()]
        self.ids = np.zeros(grid.cells(), dtype=np.uint16)
        self.pec: Optional[np.ndarray] = None

    @staticmethod
    def vacuum(grid: Grid) -> "CellMaterials":
        return CellMaterials(grid)

    def _id(self, m: Material) -> int:
        for q, t in enumerate(self.table):
            if t == m:
                return q
        if len(self.table) >= 65535:
            raise EmsError(nat.EMS_E_INVALID, "too many materials")
        self.table.append(m)
        return len(self.table) - 1

    def fill_box(self, ranges, m: Material) -> None:
        """CellMaterials::fill_box: cells in the three index ranges."""
        q = self._id(m)
        self.ids[ranges[0][0]:ranges[0][1], ranges[1][0]:ranges[1][1], ranges[2][0]:ranges[2][1]] = q

    def set(self, cell, m: Material) -> None:
        self.ids[cell[0], cell[1], cell[2]] = self._id(m)

    def fill_solid(self, solid, m: Material) -> int:
        """The cells of a box, sphere or cylinder (ems_amd.raster.Solid) by
        their centres, rasterised by the host library; returns how many."""
        from .raster import raster_cells
        q = self._id(m)
        return raster_cells(self.grid, solid, self.ids, q)

    def _pec(self) -> np.ndarray:
        if self.pec is None:
            n = self.grid.nodes()
            self.pec = np.zeros((3,) + n, dtype=np.uint8)
        return self.pec

    def set_pec_edge(self, axis: int, node) -> None:
        self._pec()[axis, node[0], node[1], node[2]] = 1

    def fill_pec_box(self, ranges) -> None:
        """CellMaterials::fill_pec_box (material.rs:262-281): every edge of
        the node box whose end stays inside it."""
        pm = self._pec()
        for a in range(3):
            r = [list(x) for x in ranges]
            r[a][1] = r[a][1] - 1
            if r[a][1] > r[a][0]:
                pm[a, r[0][0]:r[0][1], r[1][0]:r[1][1], r[2][0]:r[2][1]] = 1


class TimeStep:
    """src/engine/coeffs.rs:125-239 (Courant and Fixed)."""

    def __init__(self, kind: str, value: float):
        self.kind, self.value = kind, float(value)

    @staticmethod
    def courant(s: float = 0.99) -> "TimeStep":
        return TimeStep("courant", s)

    @staticmethod
    def fixed(dt: float) -> "TimeStep":
        return TimeStep("fixed", dt)

    def resolve(self, grid: Grid) -> float:
        mx = courant_limit(grid)
        v = self.value
        if self.kind == "courant":
            if math.isfinite(v) and 0.0 < v <= 1.0:
                return v * mx
            raise EmsError(nat.EMS_E_INVALID, f"Courant factor {v} must be in (0, 1]")
        if math.isfinite(v) and 0.0 < v <= mx:
            return v
        raise EmsError(nat.EMS_E_INVALID,
                       f"time step {v:e} s must be positive and at most the stability limit {mx:e} s")


def courant_limit(grid: Grid) -> float:
    return float(nat.ems_courant_limit(*grid._args()))


def coefficients(grid: Grid, mats: Optional[CellMaterials], faces: Sequence[int], dt: float,
                 rounding: int = 0, threads: int = 0) -> np.ndarray:
    """Coefficients (12 dense lanes, shape (12, nx, ny, nz)) from cell
    materials: EdgeMaterials::from_cells + Coefficients::from_edges_rounded."""
    n = grid.nodes()
    lanes = np.empty((12,) + n, dtype=np.float32)
    if mats is None:
        mats = CellMaterials.vacuum(grid)
    table = np.ascontiguousarray(np.array([m.row() for m in mats.table], dtype=np.float64))
    ids = np.ascontiguousarray(mats.ids)
    pec = None if mats.pec is None else np.ascontiguousarray(mats.pec)
    fp = (C.c_int32 * 6)(*[1 if f == FaceCondition.Pec else 0 for f in faces])
    if threads <= 0:
        import os
        threads = min(32, os.cpu_count() or 1)
    rc = nat.ems_coefficients_from_cells(*grid._args(), nat.ptr(ids, C.c_uint16),
                                         nat.ptr(table, C.c_double), len(mats.table),
                                         nat.ptr(pec, C.c_uint8), fp, dt, rounding, threads,
                                         nat.ptr(lanes, C.c_float))
    nat.check(rc)
    return lanes


def edge_materials(grid: Grid, mats: Optional[CellMaterials], threads: int = 0):
    """EdgeMaterials::from_cells (edge_material.rs:376-417): (eps_r, mu_r),
    each shape (3, nx, ny, nz): the relative permittivity of every E edge
    and permeability of every H dual edge by axis (0 on padding)."""
    n = grid.nodes()
    if mats is None:
        mats = CellMaterials.vacuum(grid)
    table = np.ascontiguousarray(np.array([m.row() for m in mats.table], dtype=np.float64))
    ids = np.ascontiguousarray(mats.ids)
    eps = np.empty((3,) + n, dtype=np.float64)
    mu = np.empty((3,) + n, dtype=np.float64)
    if threads <= 0:
        import os
        threads = min(32, os.cpu_count() or 1)
    nat.check(nat.ems_edge_materials_from_cells(*grid._args(), nat.ptr(ids, C.c_uint16),
                                                nat.ptr(table, C.c_double), len(mats.table),
                                                threads, nat.ptr(eps, C.c_double),
                                                nat.ptr(mu, C.c_double)))
    return eps, mu


def pool_lane(dense: np.ndarray):
    """Lane::pool: (rows, index) if pooled, else None."""
    nx, ny, nz = dense.shape
    d = np.ascontiguousarray(dense)
    rows = np.empty((max(nx * ny // 4, 1), nz), dtype=np.float32)
    index = np.empty(nx * ny, dtype=np.uint32)
    nrows = C.c_uint64(0)
    r = nat.ems_pool_lane(nat.ptr(d, C.c_float), nx, ny, nz, nat.ptr(rows, C.c_float),
                          nat.ptr(index, C.c_uint32), C.byref(nrows))
    if r < 0:
        nat.check(r)
    if r == 0:
        return None
    return np.ascontiguousarray(rows[: nrows.value]), index


# ------------------------------------------------------------ waveform --

class Waveform:
    """src/waveform.rs: the ems Gaussian (cos * exp on [0, 2T]) or openEMS's
    Gaussian pulse (``openems_gaussian``: tau = 3/(2 pi fc), T = 3 tau, its
    samples k = 1 .. floor(2T/dt))."""

    def __init__(self, f0: float, fc: float, openems: bool = False):
        self.f0, self.fc = float(f0), float(fc)
        self.openems = bool(openems)

    @staticmethod
    def gaussian(f0: float, fc: float) -> "Waveform":
        return Waveform(f0, fc)

    @staticmethod
    def openems_gaussian(f0: float, fc: float) -> "Waveform":
        return Waveform(f0, fc, openems=True)

    def value(self, t: float, dt: float = 0.0) -> float:
        if self.openems:
            return float(nat.ems_openems_gaussian(self.f0, self.fc, t, dt))
        return float(nat.ems_gaussian(self.f0, self.fc, t))

    def duration(self) -> float:
        if self.openems:
            return float(nat.ems_openems_gaussian_duration(self.fc))
        return float(nat.ems_gaussian_duration(self.fc))

    def values(self, dt: float, offset: float, count: int) -> np.ndarray:
        """value((n + offset) * dt) for n < count, in f64."""
        out = np.empty(count, dtype=np.float64)
        if self.openems:
            for n in range(count):
                out[n] = self.value((n + offset) * dt, dt)
            return out
        nat.ems_gaussian_values(self.f0, self.fc, dt, offset, count, nat.ptr(out, C.c_double))
        return out

    def table(self, dt: float, offset: float, count: int) -> np.ndarray:
        if self.openems:
            return self.values(dt, offset, count).astype(np.float32)
        out = np.empty(count, dtype=np.float32)
        nat.ems_gaussian_table(self.f0, self.fc, dt, offset, count, nat.ptr(out, C.c_float))
        return out


class OpenEmsSignal:
    """The amplitude of a model excitation with openEMS's timing
    (``openems_signal`` / ``openems_lag``, src/sim/map.rs:278-305): a field
    at time T receives the openEMS Gaussian's sample at T - dt (soft
    sources sample mid-update, so their lag is dt/2; hard ones dt), delayed
    by floor(delay/dt) whole steps."""

    def __init__(self, waveform: "Waveform", delay: float = 0.0):
        if not waveform.openems:
            raise EmsError(nat.EMS_E_INVALID, "OpenEmsSignal takes an openEMS Gaussian")
        self.waveform, self.delay = waveform, float(delay)

    def table(self, kind: int, dt: float) -> np.ndarray:
        count = int(math.ceil((self.waveform.duration() + self.delay) / dt)) + 3
        out = np.empty(count, dtype=np.float32)
        nat.check(nat.ems_openems_source_table(self.waveform.f0, self.waveform.fc, self.delay,
                                               int(kind), dt, count, nat.ptr(out, C.c_float)))
        return out


# ----------------------------------------------------------- operators --

class SourceKind(IntEnum):
    """src/engine/excite.rs:10-28"""
    SoftE = 0
    HardE = 1
    SoftH = 2
    HardH = 3


# Sample time offset of each kind, in steps (excite.rs:156-172).
_SOURCE_OFFSET = {SourceKind.SoftE: 0.5, SourceKind.HardE: 1.0,
                  SourceKind.SoftH: 0.0, SourceKind.HardH: 0.5}


class EdgeSource:
    """EdgeSource (src/engine/excite.rs:55-116). The amplitude is a
    :class:`Waveform` (tabulated over its duration) or a callable with an
    explicit ``horizon`` in steps after which it is taken as zero."""

    def __init__(self, amplitude, horizon: Optional[int] = None):
        self.amplitude = amplitude
        self.horizon = horizon
        self._kind = SourceKind.SoftE
        self.edges: List[Tuple[int, Tuple[int, int, int], float]] = []

    def kind(self, k: SourceKind) -> "EdgeSource":
        self._kind = SourceKind(k)
        return self

    def edge(self, axis: int, node, weight: float) -> "EdgeSource":
        self.edges.append((int(axis), tuple(int(v) for v in node), float(weight)))
        return self

    def _table(self, dt: float) -> np.ndarray:
        off = _SOURCE_OFFSET[self._kind]
        if isinstance(self.amplitude, OpenEmsSignal):
            return self.amplitude.table(self._kind, dt)
        if isinstance(self.amplitude, Waveform):
            count = int(math.ceil(self.amplitude.duration() / dt)) + 2
            return self.amplitude.table(dt, off, count)
        if self.horizon is None:
            raise EmsError(nat.EMS_E_INVALID, "a callable amplitude needs a horizon (steps)")
        return np.array([np.float32(self.amplitude((n + off) * dt)) for n in range(self.horizon)],
                        dtype=np.float32)


class ProbeQuantity(IntEnum):
    Voltage = 0
    Current = 1


class Probe:
    """Probe (src/engine/probe.rs:240-513)."""

    def __init__(self, quantity: ProbeQuantity):
        self.quantity = ProbeQuantity(quantity)
        self.entries: List[Tuple[int, Tuple[int, int, int], float]] = []
        self._interval = 1
        self._f32_sum = False
        self._engine = None
        self._id = -1

    @staticmethod
    def voltage(axis: int, node) -> "Probe":
        return Probe(ProbeQuantity.Voltage).entry(axis, node, 1.0)

    @staticmethod
    def current(axis: int, node) -> "Probe":
        return Probe(ProbeQuantity.Current).entry(axis, node, 1.0)

    def entry(self, axis: int, node, weight: float) -> "Probe":
        self.entries.append((int(axis), tuple(int(v) for v in node), float(weight)))
        return self

    def interval(self, n: int) -> "Probe":
        self._interval = max(int(n), 1)
        return self

    def f32_sum(self, on: bool = True) -> "Probe":
        """Probe::f32_sum (probe.rs:398-404): the sum of the weighted
        entries in f32 (openEMS's current probes)."""
        self._f32_sum = bool(on)
        return self

    def _c_quantity(self) -> int:
        return int(self.quantity) | (nat.EMS_PROBE_F32_SUM if self._f32_sum else 0)

    def values(self) -> np.ndarray:
        return self._engine._probe_values(self._id)

    def times(self) -> np.ndarray:
        return self._engine._probe_times(self._id)


@dataclass
class Cpml:
    """Cpml settings (src/fdtd.rs:177-296)."""
    cells: int = 8
    grading_order: float = 3.0
    sigma_factor: float = 1.0
    kappa_max: float = 1.0
    alpha_max: float = 0.0
    convolution: int = 1  # 0 Constant, 1 Linear (default)

    @staticmethod
    def with_cells(n: int) -> "Cpml":
        return Cpml(cells=n)

    def c_params(self) -> nat.CpmlParams:
        return nat.CpmlParams(self.cells, self.grading_order, self.sigma_factor, self.kappa_max,
                              self.alpha_max, self.convolution)

    def validate(self) -> None:
        if self.cells < 1:
            raise EmsError(nat.EMS_E_INVALID, "CPML: needs at least one cell")
        if not (math.isfinite(self.grading_order) and self.grading_order >= 0):
            raise EmsError(nat.EMS_E_INVALID, "CPML: grading order must be finite and >= 0")
        if not (math.isfinite(self.kappa_max) and self.kappa_max >= 1):
            raise EmsError(nat.EMS_E_INVALID, "CPML: kappa_max must be finite and >= 1")
        if not (math.isfinite(self.alpha_max) and self.alpha_max >= 0):
            raise EmsError(nat.EMS_E_INVALID, "CPML: alpha_max must be finite and >= 0")
        if not (math.isfinite(self.sigma_factor) and self.sigma_factor >= 0):
            raise EmsError(nat.EMS_E_INVALID, "CPML: sigma factor must be finite and >= 0")


def cpml_coefficients(p: Cpml, rho: float, delta: float, dt: float, n: float = 1.0) -> dict:
    out = (C.c_double * 7)()
    prm = p.c_params()
    nat.ems_cpml_coefficients(C.byref(prm), rho, delta, dt, n, out)
    keys = ("rho", "sigma", "kappa", "alpha", "b", "a", "direct")
    return dict(zip(keys, list(out)))


class CpmlBoundary:
    """CpmlBoundary (src/engine/boundary/cpml.rs:752-1047): one layer per
    (face, Cpml), sigma_max tuned to each layer's mean refractive index
    (layer_index) measured from the engine's coefficients."""

    def __init__(self, slabs: List[dict], faces: List[Tuple[Face, Cpml, float]], grid=None,
                 dt=None):
        self.slabs = slabs
        self.faces = faces
        self._grid, self._dt = grid, dt

    @staticmethod
    def new(engine: "Engine", layers) -> "CpmlBoundary":
        return CpmlBoundary._build(engine.grid, engine.dt, layers, engine._dense_lanes)

    @staticmethod
    def with_grid(grid: Grid, dt: float, layers) -> "CpmlBoundary":
        return CpmlBoundary._build(grid, dt, layers, None)

    @staticmethod
    def _build(grid: Grid, dt: float, layers, lanes) -> "CpmlBoundary":
        if not (math.isfinite(dt) and dt > 0):
            raise EmsError(nat.EMS_E_INVALID, f"CPML: invalid time step {dt}")
        n = grid.nodes()
        cells = [0] * 6
        slabs, faces = [], []
        for face, params in layers:
            face = Face(face)
            params.validate()
            if cells[face] != 0:
                raise EmsError(nat.EMS_E_INVALID, f"CPML: face {face.name} given twice")
            total = n[face.axis] - 1
            if params.cells > total:
                raise EmsError(nat.EMS_E_INVALID,
                               f"CPML on {face.name}: {params.cells} cells but the mesh has only {total}")
            cells[face] = params.cells
            if lanes is None:
                nf = 1.0
            else:
                lf = np.ascontiguousarray(lanes)
                nf = float(nat.ems_cpml_layer_index(*grid._args(), nat.ptr(lf, C.c_float), dt,
                                                    int(face), params.cells))
            np_ = n[face.axis]
            out = (nat.CpmlSlab * 4)()
            coeff = np.zeros(4 * 3 * np_, dtype=np.float32)
            prm = params.c_params()
            rc = nat.ems_cpml_layer(*grid._args(), int(face), C.byref(prm), dt, nf, out,
                                    nat.ptr(coeff, C.c_float))
            nat.check(rc)
            for q in range(4):
                s = out[q]
                depth = s.hi[s.normal] - s.lo[s.normal]
                base = q * 3 * np_
                slabs.append(dict(e_side=s.e_side, comp=s.comp, normal=s.normal,
                                  lo=tuple(s.lo), hi=tuple(s.hi),
                                  b=coeff[base:base + depth].copy(),
                                  a=coeff[base + np_:base + np_ + depth].copy(),
                                  k=coeff[base + 2 * np_:base + 2 * np_ + depth].copy()))
            faces.append((face, params, nf))
        for p in range(3):
            if cells[2 * p] + cells[2 * p + 1] > n[p] - 1:
                raise EmsError(nat.EMS_E_INVALID, f"CPML: opposite layers along axis {p} overlap")
        return CpmlBoundary(slabs, faces, grid, dt)

    def refractive_index(self, face: Face) -> Optional[float]:
        for f, _, nf in self.faces:
            if f == face:
                return nf
        return None

    def _profile(self, face: Face, e_side: int):
        for f, params, nf in self.faces:
            if f == face:
                g = self._grid
                np_ = g.nodes()[Face(face).axis]
                out = np.zeros(7 * np_, dtype=np.float64)
                cnt = C.c_int()
                prm = params.c_params()
                nat.check(nat.ems_cpml_profile(*g._args(), int(face), C.byref(prm), self._dt, nf,
                                               e_side, nat.ptr(out, C.c_double), C.byref(cnt)))
                keys = ("rho", "sigma", "kappa", "alpha", "b", "a", "direct")
                return [dict(zip(keys, out[7 * q:7 * q + 7])) for q in range(cnt.value)]
        return None

    def e_profile(self, face: Face):
        """Grading at the E positions of `face` (cpml.rs:1005-1011)."""
        return self._profile(face, 1)

    def h_profile(self, face: Face):
        """Grading at the H positions of `face` (cpml.rs:1013-1018)."""
        return self._profile(face, 0)


# ------------------------------------------------------------ run loop --

class StopReason(IntEnum):
    """src/engine/run.rs:141-162"""
    EnergyDecay = 0
    MaxTimesteps = 1
    Aborted = 2
    Diverged = 3


@dataclass
class Settle:
    """src/fdtd.rs Settle (defaults tolerance 1e-4, window 1/8, max_extension 1/2)."""
    frequency: float
    tolerance: float = 1e-4
    window: float = 0.125
    max_extension: float = 0.5


@dataclass
class EndCriteria:
    """src/fdtd.rs:428-605"""
    energy_ratio: float
    check_interval: int = 50
    settle: Optional[Settle] = None

    @staticmethod
    def energy_decay_db(db: float) -> "EndCriteria":
        return EndCriteria(10.0 ** (db / 10.0))


@dataclass
class RunLimits:
    """src/engine/run.rs:22-139. energy_measure: 0 EnergyMeasure::Invariant,
    1 EnergyMeasure::OpenEms."""
    max_timesteps: int
    end_criteria: Optional[EndCriteria] = None
    excitation_end: Optional[float] = None
    check_after_interval: bool = False
    extra_check: Optional[int] = None
    energy_measure: int = 0

    def c(self) -> nat.RunLimitsC:
        ec = self.end_criteria
        st = ec.settle if ec else None
        return nat.RunLimitsC(
            self.max_timesteps, 1 if ec else 0, ec.energy_ratio if ec else 0.0,
            ec.check_interval if ec else 50, 0 if self.excitation_end is None else 1,
            self.excitation_end if self.excitation_end is not None else 0.0,
            1 if st else 0, st.frequency if st else 0.0, st.tolerance if st else 0.0,
            st.window if st else 0.0, st.max_extension if st else 0.0,
            1 if self.check_after_interval else 0, 0 if self.extra_check is None else 1,
            0 if self.extra_check is None else int(self.extra_check), int(self.energy_measure))


@dataclass
class RunOutcome:
    """src/engine/run.rs:192-214"""
    reason: StopReason
    ratio_db: float
    steps: int
    timestep: int
    time: float
    energy: float
    max_energy: float
    elapsed: float
    criterion_steps: Optional[int]

    @staticmethod
    def from_c(o: nat.RunOutcomeC) -> "RunOutcome":
        return RunOutcome(StopReason(o.reason), o.ratio_db, o.steps, o.timestep, o.time, o.energy,
                          o.max_energy, o.elapsed_s,
                          o.criterion_steps if o.has_criterion_steps else None)


def drive_script(w: Sequence[float], limits: RunLimits, signal: Optional[Sequence[float]] = None):
    """run::drive over a scripted stepper (run.rs:548-582), for testing."""
    wa = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    sa = None if signal is None else np.ascontiguousarray(np.asarray(signal, dtype=np.float64))
    out = nat.RunOutcomeC()
    lim = limits.c()
    rc = nat.ems_drive_script(nat.ptr(wa, C.c_double), wa.size, nat.ptr(sa, C.c_double),
                              0 if sa is None else sa.size, C.byref(lim), C.byref(out))
    nat.check(rc)
    return RunOutcome.from_c(out)


# -------------------------------------------------------------- engine --

class EngineBuilder:
    """EngineBuilder (src/engine/mod.rs:225-382)."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self._mats: Optional[CellMaterials] = None
        self._faces = [FaceCondition.Pec] * 6
        self._ts = TimeStep.courant(0.99)
        self._curl = 0
        self._rounding = 0

    def materials(self, m: CellMaterials) -> "EngineBuilder":
        self._mats = m
        return self

    def faces(self, f: Sequence[int]) -> "EngineBuilder":
        self._faces = [FaceCondition(x) for x in f]
        return self

    def timestep(self, t: TimeStep) -> "EngineBuilder":
        self._ts = t
        return self

    def curl_order(self, oe: bool) -> "EngineBuilder":
        self._curl = 1 if oe else 0
        return self

    def coeff_rounding(self, openems: bool) -> "EngineBuilder":
        self._rounding = 1 if openems else 0
        return self

    def build(self) -> "Engine":
        dt = self._ts.resolve(self.grid)
        lanes = coefficients(self.grid, self._mats, self._faces, dt, self._rounding)
        return Engine(self.grid, dt, lanes, curl_order=self._curl)


class Engine:
    """A device FDTD engine (Engine, src/engine/mod.rs:460-992)."""

    @staticmethod
    def builder(grid: Grid) -> EngineBuilder:
        return EngineBuilder(grid)

    def __init__(self, grid: Grid, dt: float, lanes: np.ndarray, curl_order: int = 0,
                 pool: bool = True):
        self.grid = grid
        self.dt = float(dt)
        self.curl_order = curl_order
        n = grid.nodes()
        desc = nat.GridDesc((C.c_int32 * 3)(*n), self.dt, curl_order)
        h = C.c_void_p()
        rc = nat.ems_gpu_create(C.byref(desc), C.byref(h))
        if rc != 0:
            raise EmsError(rc, "ems_gpu_create failed (no usable GPU?)")
        self._h = h
        self._dense_lanes = lanes
        self._owns_lanes = False
        self._pool = pool
        self.unfused_prefetch = 11
        self.setup: List[tuple] = []
        self.ops: list = []
        self._tables: list = []
        self._lumped_n: dict = {}
        self._poles_n: dict = {}
        self._pw_tables: list = []
        for l in range(12):
            self._upload_lane(l, lanes[l], pool)

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise EmsError(rc, nat.ems_gpu_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            nat.ems_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def patch_lanes(self, changed) -> None:
        """Replaces whole coefficient lanes ({lane index: dense lane}) on the
        host copy and the device (coefficient patches such as conformal
        metal, before operators that read the lanes are added)."""
        if self._dense_lanes is None:
            raise EmsError(nat.EMS_E_INVALID, "the host lanes were released")
        if not self._owns_lanes:  # the caller's array stays as it was
            self._dense_lanes = self._dense_lanes.copy()
            self._owns_lanes = True
        for l, lane in sorted(changed.items()):
            self._dense_lanes[l] = lane
            self._upload_lane(l, self._dense_lanes[l], self._pool)

    def release_dense_lanes(self) -> None:
        """Drops the host copy of the dense coefficients (after setup)."""
        self._dense_lanes = None

    # -- properties
    @property
    def timestep_count(self) -> int:
        return int(nat.ems_gpu_timestep(self._h))

    def timestep(self) -> int:
        return int(nat.ems_gpu_timestep(self._h))

    def time(self) -> float:
        return self.timestep() * self.dt

    def coefficients(self) -> np.ndarray:
        return self._dense_lanes

    # -- operators
    def add_operator(self, op):
        """Engine::add_operator (mod.rs:549-572); returns the operator."""
        h = self._h
        if isinstance(op, CpmlBoundary):
            arr = (nat.CpmlSlab * len(op.slabs))()
            keep = []
            for q, s in enumerate(op.slabs):
                b, a, k = (np.ascontiguousarray(s[x]) for x in ("b", "a", "k"))
                keep += [b, a, k]
                arr[q] = nat.CpmlSlab(s["e_side"], s["comp"], s["normal"],
                                      (C.c_int32 * 3)(*s["lo"]), (C.c_int32 * 3)(*s["hi"]),
                                      nat.ptr(b, C.c_float), nat.ptr(a, C.c_float),
                                      nat.ptr(k, C.c_float))
            self._check(nat.ems_gpu_add_cpml(h, arr, len(op.slabs)))
            self.setup.append(("cpml", op.slabs))
        elif isinstance(op, EdgeSource):
            amp = np.ascontiguousarray(op._table(self.dt))
            axis, nodes, w = self._edges(op.edges, np.float32)
            oid = C.c_int()
            self._check(nat.ems_gpu_add_source(h, int(op._kind), len(op.edges),
                                               nat.ptr(axis, C.c_int32), nat.ptr(nodes, C.c_int32),
                                               nat.ptr(w, C.c_float), nat.ptr(amp, C.c_float),
                                               amp.size, C.byref(oid)))
            self.setup.append(("source", int(op._kind), axis, nodes, w, amp))
            op._id = oid.value
        elif isinstance(op, Probe):
            axis, nodes, w = self._edges(op.entries, np.float64)
            oid = C.c_int()
            q = op._c_quantity()
            self._check(nat.ems_gpu_add_probe(h, q, len(op.entries),
                                              nat.ptr(axis, C.c_int32), nat.ptr(nodes, C.c_int32),
                                              nat.ptr(w, C.c_double), op._interval, C.byref(oid)))
            self.setup.append(("probe", q, axis, nodes, w, op._interval))
            op._engine = self
            op._id = oid.value
        elif type(op).__name__ == "PlaneWave":
            from . import planewave as PW
            d, keep = PW.make_desc(op)
            oid = C.c_int()
            self._check(nat.ems_gpu_add_planewave(h, C.byref(d), C.byref(oid)))
            table = PW.SourceTable(op)
            self._pw_tables.append((oid.value, table))
            self.setup.append(("planewave", op))
            op._bind(self, oid.value)
            self._ensure_tables(self.timestep() + 1)
        elif type(op).__name__ == "MergedFaces":
            # the reduced currents of a conformal geometry (ConformalPec::apply)
            from . import merged as MF
            d, keep = MF.make_desc(op)
            oid = C.c_int()
            self._check(nat.ems_gpu_add_merged(h, C.byref(d), C.byref(oid)))
            self.setup.append(("merged", op))
            op._bind(self, oid.value)
        elif type(op).__name__ == "Dispersive":
            # Dispersive::add_to (dispersive.rs:2479-2520): host install,
            # then the device medium.
            from . import dispersive as DP
            if self._dense_lanes is None:
                raise EmsError(nat.EMS_E_INVALID, "dispersive media need the host lanes")
            plan = op._plan(self.grid, self.dt, self._dense_lanes)
            se, sh, cs, keep = DP.make_desc(plan)
            oid = C.c_int()
            self._check(nat.ems_gpu_add_dispersive(h, C.byref(se), C.byref(sh), C.byref(cs),
                                                   C.byref(oid)))
            self._poles_n[oid.value] = (len(plan["sides"][0]["cidx"]), len(plan["sides"][1]["cidx"]))
            self.setup.append(("dispersive", plan))
            op._bind(self, oid.value, plan)
        elif type(op).__name__ == "MurBoundary":
            d, keep = op.desc()
            oid = C.c_int()
            self._check(nat.ems_gpu_add_mur(h, C.byref(d), C.byref(oid)))
            self.setup.append(("mur", op))
        elif type(op).__name__ == "OpenEmsPml":
            d, keep = op.desc()
            oid = C.c_int()
            self._check(nat.ems_gpu_add_upml(h, C.byref(d), C.byref(oid)))
            op._oid = getattr(op, "_oid", {})
            op._oid[id(self)] = oid.value
            self.setup.append(("upml", op))
        elif hasattr(op, "_plan"):
            # LumpedElement / LumpedPort (lumped.rs:880-1012): fold into the
            # lanes, re-upload them, register the device operator.
            from . import lumped as LM
            if self._dense_lanes is None:
                raise EmsError(nat.EMS_E_INVALID, "lumped elements need the host lanes")
            plan = op._plan(self.grid, self.dt, self._dense_lanes)
            a = plan["axis"]
            if not self._owns_lanes:  # the caller's array stays as it was
                self._dense_lanes = self._dense_lanes.copy()
                self._owns_lanes = True
            for l, lane in ((a, plan["ca"]), (3 + a, plan["cb"])):
                self._dense_lanes[l] = lane
                self._upload_lane(l, lane)
            d, keep = LM.make_desc(plan)
            oid = C.c_int()
            self._check(nat.ems_gpu_add_lumped(h, C.byref(d), C.byref(oid)))
            tables = LM.SourceTables(plan, self.dt)
            self._tables.append((oid.value, tables))
            self._lumped_n[oid.value] = len(plan["edges"])
            self.setup.append(("lumped", plan, tables))
            op._bind(self, oid.value, plan)
            self._ensure_tables(self.timestep() + 1)
        else:
            raise EmsError(nat.EMS_E_UNSUPPORTED, f"operator {type(op).__name__}")
        self.ops.append(op)
        return op

    def _upload_lane(self, l: int, lane: np.ndarray, pool: bool = True) -> None:
        n = self.grid.nodes()
        lane = np.ascontiguousarray(lane)
        pooled = pool_lane(lane) if pool else None
        if pooled is None:
            rows = lane.reshape(-1, n[2])
            self._check(nat.ems_gpu_set_lane(self._h, l, nat.ptr(rows, C.c_float), rows.shape[0], None))
            self.setup.append(("lane", l, rows, None))
        else:
            rows, index = pooled
            self._check(nat.ems_gpu_set_lane(self._h, l, nat.ptr(rows, C.c_float), rows.shape[0],
                                             nat.ptr(index, C.c_uint32)))
            self.setup.append(("lane", l, rows, index))

    def _ensure_tables(self, upto: int) -> None:
        for oid, t in self._pw_tables:
            def up_pw(vals, oid=oid):
                v = np.ascontiguousarray(vals)
                self._check(nat.ems_gpu_planewave_table(self._h, oid, nat.ptr(v, C.c_double),
                                                        v.size))
            t.ensure(upto, up_pw)
        for oid, t in self._tables:
            def up(term, vals, oid=oid):
                v = np.ascontiguousarray(vals)
                self._check(nat.ems_gpu_lumped_table(self._h, oid, term, nat.ptr(v, C.c_double),
                                                     v.size))
            t.ensure(upto, up)

    def _lumped_values(self, oid: int, which: int):
        cnt = C.c_uint64()
        first = C.c_uint64()
        self._check(nat.ems_gpu_lumped_values(self._h, oid, which, None, 0, C.byref(cnt),
                                              C.byref(first)))
        out = np.empty(cnt.value, dtype=np.float64)
        self._check(nat.ems_gpu_lumped_values(self._h, oid, which, nat.ptr(out, C.c_double),
                                              out.size, C.byref(cnt), C.byref(first)))
        return out, first.value

    def _add_dft(self, oid: int, which: int, freqs, t0: float, dt: float) -> int:
        f = np.ascontiguousarray(np.asarray(freqs, dtype=np.float64))
        did = C.c_int()
        self._check(nat.ems_gpu_add_dft(self._h, oid, which, nat.ptr(f, C.c_double), f.size,
                                        float(t0), float(dt), C.byref(did)))
        return did.value

    def _dft_update(self, did: int) -> None:
        self._check(nat.ems_gpu_dft_update(self._h, did))

    def _dft_result(self, did: int, nf: int):
        out = np.empty(2 * nf, dtype=np.float64)
        n = C.c_uint64()
        self._check(nat.ems_gpu_dft_result(self._h, did, nat.ptr(out, C.c_double), C.byref(n)))
        return [complex(out[2 * k], out[2 * k + 1]) for k in range(nf)], n.value

    def _planewave_aux(self, oid: int, naux: int):
        e, hh = np.zeros(naux), np.zeros(naux - 1)
        self._check(nat.ems_gpu_planewave_aux(self._h, oid, nat.ptr(e, C.c_double),
                                              nat.ptr(hh, C.c_double)))
        return e, hh

    def _merged_state(self, oid: int, ng: int):
        cur, circ = np.zeros(ng), np.zeros(ng)
        self._check(nat.ems_gpu_merged_state(self._h, oid, nat.ptr(cur, C.c_double),
                                             nat.ptr(circ, C.c_double)))
        return cur, circ

    def _dispersive_state(self, oid: int, field: int):
        n = self._poles_n[oid][field]
        j, q = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        self._check(nat.ems_gpu_dispersive_state(self._h, oid, field, nat.ptr(j, C.c_double),
                                                 nat.ptr(q, C.c_double)))
        return j[:n], q[:n]

    def _lumped_branch(self, oid: int):
        n = self._lumped_n[oid]
        j, q, jav = (np.zeros(n) for _ in range(3))
        self._check(nat.ems_gpu_lumped_branch(self._h, oid, nat.ptr(j, C.c_double),
                                              nat.ptr(q, C.c_double), nat.ptr(jav, C.c_double)))
        return j, q, jav

    @staticmethod
    def _edges(edges, wtype):
        axis = np.ascontiguousarray(np.array([e[0] for e in edges], dtype=np.int32))
        nodes = np.ascontiguousarray(np.array([e[1] for e in edges], dtype=np.int32).reshape(-1))
        w = np.ascontiguousarray(np.array([e[2] for e in edges], dtype=wtype))
        return axis, nodes, w

    # -- stepping
    def run(self, n: int) -> None:
        self._ensure_tables(self.timestep() + int(n))
        self._check(nat.ems_gpu_run(self._h, int(n)))

    def step(self) -> None:
        self.run(1)

    def step_with_energy(self) -> float:
        self._ensure_tables(self.timestep() + 1)
        w = C.c_double()
        self._check(nat.ems_gpu_step_with_energy(self._h, C.byref(w)))
        return w.value

    def energy(self) -> float:
        w = C.c_double()
        self._check(nat.ems_gpu_energy(self._h, C.byref(w)))
        return w.value

    def energy_in(self, measure: int) -> float:
        """Engine::energy_in (mod.rs:747-752): 0 the invariant, 1 openEMS's
        measure."""
        w = C.c_double()
        self._check(nat.ems_gpu_energy_in(self._h, int(measure), C.byref(w)))
        return w.value

    def energy_openems(self) -> float:
        """Engine::energy_openems (mod.rs:727-741)."""
        return self.energy_in(1)

    def run_until(self, limits: RunLimits, watch: Sequence[Probe] = ()) -> RunOutcome:
        self._ensure_tables(self.timestep() + int(limits.max_timesteps) + 1)
        out = nat.RunOutcomeC()
        lim = limits.c()
        ids = np.ascontiguousarray(np.array([p._id for p in watch], dtype=np.int32))
        self._check(nat.ems_run_until(self._h, C.byref(lim), nat.ptr(ids, C.c_int32), ids.size,
                                      C.byref(out)))
        return RunOutcome.from_c(out)

    def synchronize(self) -> None:
        self._check(nat.ems_gpu_synchronize(self._h))

    # -- fields
    def _read(self, field: int, axis: int) -> np.ndarray:
        out = np.empty(self.grid.nodes(), dtype=np.float32)
        self._check(nat.ems_gpu_read_component(self._h, field, int(axis), nat.ptr(out, C.c_float)))
        return out

    def u_component(self, axis: int) -> np.ndarray:
        return self._read(0, axis)

    def i_component(self, axis: int) -> np.ndarray:
        return self._read(1, axis)

    def set_component(self, field: int, axis: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(self.grid.nodes()))
        self._check(nat.ems_gpu_write_component(self._h, field, int(axis), nat.ptr(v, C.c_float)))
        self.setup.append(("write", field, int(axis), v.copy()))

    def _probe_values(self, pid: int) -> np.ndarray:
        cnt = C.c_uint64()
        self._check(nat.ems_gpu_probe_values(self._h, pid, None, 0, C.byref(cnt)))
        out = np.empty(cnt.value, dtype=np.float64)
        self._check(nat.ems_gpu_probe_values(self._h, pid, nat.ptr(out, C.c_double), out.size,
                                             C.byref(cnt)))
        return out

    def _probe_times(self, pid: int) -> np.ndarray:
        cnt = C.c_uint64()
        self._check(nat.ems_gpu_probe_times(self._h, pid, None, 0, C.byref(cnt)))
        out = np.empty(cnt.value, dtype=np.float64)
        self._check(nat.ems_gpu_probe_times(self._h, pid, nat.ptr(out, C.c_double), out.size,
                                            C.byref(cnt)))
        return out

    def upml_state(self, op, side: int) -> Tuple[np.ndarray, np.ndarray]:
        """The state (flux, value) of a registered OpenEmsPml's E edges
        (side 0) or H dual edges (side 1), in its edge order."""
        n = op.edge_counts()[side]
        f, v = np.empty(n, np.float32), np.empty(n, np.float32)
        self._check(nat.ems_gpu_upml_state(self._h, op._oid[id(self)], int(side),
                                           nat.ptr(f, C.c_float), nat.ptr(v, C.c_float)))
        return f, v

    # -- measurement
    def time_kernels(self, steps: int) -> Tuple[float, float, float]:
        out = (C.c_double * 3)()
        self._check(nat.ems_gpu_time_kernels(self._h, int(steps), out))
        return tuple(out)

    def set_schedule(self, fused=True, tile_rows: int = 0, x_chunks: int = 0) -> None:
        """Fused (one pass per step, double-buffered fields; 2 = without the
        next-plane prefetch; 3 full rows; 4 packed short rows, nz <= 64,
        5 the same without the prefetch) or unfused (H, CPML-H, E, CPML-E
        launches) steps; the fused tile height (3, 7 or 15 rows) and the
        number of x chunks (0 = automatic)."""
        mode = int(fused) if not isinstance(fused, bool) else (1 if fused else 0)
        self._check(nat.ems_gpu_set_schedule(self._h, mode, int(tile_rows), int(x_chunks)))

    def set_cpml_split(self, on: bool) -> None:
        """Whether the tiled fused step corrects its z-normal CPML slabs in
        passes of their own (before the kernel for H, after it for E) rather
        than inside it; the same values either way."""
        self._check(nat.ems_gpu_set_cpml_split(self._h, 1 if on else 0))

    def set_tile_classes(self, on: bool) -> None:
        """Whether the tiled fused step runs the tiles a CPML slab reaches
        and the others in two launches (the latter without the slab code);
        the same values either way."""
        self._check(nat.ems_gpu_set_tile_classes(self._h, 1 if on else 0))

    def set_unfused_prefetch(self, which: int = 11) -> None:
        """The unfused kernels: bits 0 (H) and 1 (E) march longer x chunks
        with the next plane's inputs loaded ahead; bits 2 (H) and 3 (E),
        with the side's loading-ahead kernel, correct the side's CPML slabs
        (but a z-normal suffix) inside it. Default 11; the same values
        either way."""
        self._check(nat.ems_gpu_set_unfused_prefetch(self._h, int(which)))
        self.unfused_prefetch = int(which)

    def set_xcd_group(self, group: int) -> None:
        """Tile rows per XCD visit in the fused kernels' block order (0:
        plain order)."""
        self._check(nat.ems_gpu_set_xcd_group(self._h, int(group)))

    def autotune(self, candidates=None, steps_each: int = 4, xcd_groups=(1, 2, 4, 8)):
        """Times a few real steps of each fused tiling on this engine (they
        advance the simulation like any other steps) and keeps the fastest.
        Candidates are (mode, tile_rows, x_chunks), mode 1 with and 2
        without the next-plane prefetch (mode 3, the full-row kernel, only
        when listed). Returns the choice with its mean kernel time in ms,
        and all timings."""
        if candidates is None:
            candidates = [(m, t, c) for m in (1, 2) for t in (3, 7, 15)
                          for c in (0, 16, 32, 64, 96, 128)]
            if self.grid.nodes()[2] <= 64:
                candidates += [(m, 7, c) for m in (4, 5) for c in (0, 16, 32, 64)]
        results = []
        for m, t, c in candidates:
            self.set_schedule(m, t, c)
            self.time_kernels(1)  # first launch of a tiling: not timed
            ms = self.time_kernels(steps_each)[0]
            results.append((m, t, c, ms))
        best = min(results, key=lambda r: r[3])
        self.set_schedule(best[0], best[1], best[2])
        # then the XCD grouping of the tile order for that tiling
        groups = []
        for g in xcd_groups:
            self.set_xcd_group(g)
            self.time_kernels(1)
            groups.append((g, self.time_kernels(steps_each)[0]))
        self.xcd_trials = groups
        if groups:
            g_best = min(groups, key=lambda r: r[1])
            self.set_xcd_group(g_best[0])
            self.xcd_group = g_best[0]
            if g_best[1] < best[3]:
                best = (best[0], best[1], best[2], g_best[1])
        return best, results

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """Device-to-device copy bandwidth (GB/s of read + write)."""
        out = C.c_double()
        self._check(nat.ems_gpu_copy_bandwidth(self._h, int(nbytes), int(reps), C.byref(out)))
        return out.value

    def traffic_model(self) -> Tuple[float, float, float]:
        out = (C.c_double * 3)()
        self._check(nat.ems_gpu_traffic_model(self._h, out))
        return tuple(out)


def energy_many(engines: Sequence["Engine"], active: Optional[Sequence[bool]] = None) -> List[float]:
    """ems_gpu_energy_many: the energy of every engine (of the active ones;
    the others' entries are 0.0), each the value its own energy() gives.
    Engines without operator energies (no poles, cut-face currents or
    lumped branches) of one curl order and row width are evaluated
    together, one launch per kernel."""
    n = len(engines)
    arr = (C.c_void_p * n)(*[e._h for e in engines])
    act = None
    if active is not None:
        act = np.ascontiguousarray(np.array([1 if a else 0 for a in active], dtype=np.uint8))
    out = np.zeros(n, dtype=np.float64)
    rc = nat.ems_gpu_energy_many(arr, n, nat.ptr(act, C.c_uint8), out.ctypes.data_as(nat._dp))
    if rc != 0:
        raise EmsError(rc, "ems_gpu_energy_many failed")
    return [float(v) for v in out]


class EngineBatch:
    """Engines stepped together (include/ems_gpu.h batches): one fused launch
    over the tiles of every active model per step, each model bit-identical
    to stepping it alone. No reference counterpart (the reference runs one
    model per engine); the run loop per model is run::drive
    (src/engine/run.rs:424-540)."""

    def __init__(self, engines: Sequence[Engine]):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e._h for e in self.engines])
        h = C.c_void_p()
        rc = nat.ems_gpu_batch_create(arr, len(self.engines), C.byref(h))
        if rc != 0:
            raise EmsError(rc, "ems_gpu_batch_create failed (mixed curl orders, a dead engine "
                                "or an engine listed twice)")
        self._h = h
        self._arr = arr

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise EmsError(rc, nat.ems_gpu_batch_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            nat.ems_gpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schedule(self, tile_rows: int = 7, x_chunks: int = 0, prefetch: bool = False,
                     packed: bool = False) -> None:
        """The fused kernel of every model: tiled (tile height 3, 7 or 15) or
        packed short rows (64 // nz whole rows per wave, every nz <= 64);
        x chunks per model (0 automatic); the next-plane prefetch (off by
        default: without it the kernels run eight waves per SIMD, and the
        256-model patch batch steps in 1.201 ms against 1.336 with it)."""
        mode = (4 if prefetch else 5) if packed else (1 if prefetch else 2)
        self._check(nat.ems_gpu_batch_set_schedule(self._h, mode, int(tile_rows), int(x_chunks)))

    def set_graphs(self, on: bool) -> None:
        """Whether runs replay captured HIP graphs (default) or queue the
        same launches one by one; the same results either way."""
        self._check(nat.ems_gpu_batch_set_graphs(self._h, 1 if on else 0))

    def run(self, n: int, active: Optional[Sequence[bool]] = None) -> None:
        act = None
        if active is not None:
            act = np.ascontiguousarray(np.array([1 if a else 0 for a in active], dtype=np.uint8))
        for q, e in enumerate(self.engines):
            if active is None or active[q]:
                e._ensure_tables(e.timestep() + int(n))
        self._check(nat.ems_gpu_batch_run(self._h, nat.ptr(act, C.c_uint8), int(n)))

    def run_until(self, limits: "RunLimits") -> List["RunOutcome"]:
        for e in self.engines:
            e._ensure_tables(e.timestep() + int(limits.max_timesteps) + 1)
        outs = (nat.RunOutcomeC * len(self.engines))()
        lim = limits.c()
        self._check(nat.ems_run_until_batch(self._h, len(self.engines), self._arr, C.byref(lim),
                                            outs))
        return [RunOutcome.from_c(o) for o in outs]

    def autotune(self, candidates=None, steps_each: int = 8):
        """Times a few real steps of each (tile_rows, x_chunks) on the batch
        (they advance every model) and keeps the fastest."""
        import time as _t
        if candidates is None:
            candidates = [(t, c) for t in (3, 7) for c in (0, 1, 2, 4)]
            if all(e.grid.nodes()[2] <= 64 for e in self.engines):
                candidates += [("packed", c) for c in (0, 1, 2, 4)]
        results = []
        for t, c in candidates:
            if t == "packed":
                self.set_schedule(7, c, packed=True)
            else:
                self.set_schedule(t, c)
            self.run(2)
            t0 = _t.perf_counter()
            self.run(steps_each)
            results.append((t, c, (_t.perf_counter() - t0) / steps_each * 1e3))
        best = min(results, key=lambda r: r[2])
        if best[0] == "packed":
            self.set_schedule(7, best[1], packed=True)
        else:
            self.set_schedule(best[0], best[1])
        return best, results
