"""Cells of solids on a rectilinear grid, rasterised by the host library
(ems_raster_cells): a cell belongs to a box, a sphere or a cylinder when its
centre (the midpoints of its node lines) lies inside. The cells' material
ids then go to the quarter-cell averaging of ems_coefficients_from_cells
(coeffs.rs:876-966), which weighs each cell touching an edge by its quarter
of the edge's dual face. The reference takes its solids from its external
csx crate; these primitives cover the solids of the benchmark workloads.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError


@dataclass(frozen=True)
class Solid:
    """ems_solid: kind 0 box, 1 sphere, 2 cylinder; p as the header states."""
    kind: int
    axis: int
    p: Tuple[float, ...]

    @staticmethod
    def box(lo: Sequence[float], hi: Sequence[float]) -> "Solid":
        """[lo, hi) per axis, metres."""
        return Solid(0, 0, tuple(float(v) for v in list(lo) + list(hi)))

    @staticmethod
    def sphere(centre: Sequence[float], radius: float) -> "Solid":
        return Solid(1, 0, tuple(float(v) for v in centre) + (float(radius), 0.0, 0.0))

    @staticmethod
    def cylinder(axis: int, centre: Sequence[float], radius: float, lo: float,
                 hi: float) -> "Solid":
        """Along `axis`, centred at `centre` in the other two axes (in
        increasing axis order), between lo and hi along the axis."""
        return Solid(2, int(axis), (float(centre[0]), float(centre[1]), float(radius), float(lo),
                                    float(hi), 0.0))

    def c(self) -> nat.SolidC:
        return nat.SolidC(self.kind, self.axis, (C.c_double * 6)(*self.p))


def raster_cells(grid, solid: Solid, ids: np.ndarray, value: int) -> int:
    """Sets ids[cell] = value (ids: uint16, the grid's cell shape) for the
    cells of `solid`; returns how many."""
    if ids.dtype != np.uint16 or ids.shape != tuple(grid.cells()) or not ids.flags.c_contiguous:
        raise EmsError(nat.EMS_E_INVALID, "ids: a C-contiguous uint16 array of the cell shape")
    s = solid.c()
    n = nat.ems_raster_cells(*grid._args(), C.byref(s), int(value),
                             ids.ctypes.data_as(C.POINTER(C.c_uint16)))
    if n < 0:
        raise EmsError(int(n), "ems_raster_cells failed")
    return int(n)


def cell_mask(grid, solid: Solid) -> np.ndarray:
    """The solid's cells as a boolean array of the grid's cell shape."""
    ids = np.zeros(grid.cells(), dtype=np.uint16)
    raster_cells(grid, solid, ids, 1)
    return ids.astype(bool)


__all__ = ["Solid", "cell_mask", "raster_cells"]
