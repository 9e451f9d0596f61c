"""Benchmark and parity workloads, built through the engine API.

* ``vacuum_box``: a uniform 1 mm PEC box of n^3 cells, a soft Gaussian Ez
  source at the centre and an Ez probe (as the reference's
  examples/bench_vacuum.rs, with the source and probe of the 50^3 case).
* ``cpml_box``: an n^3-cell graded mesh (spacing 0.5-1.5 mm from a seeded
  walk, neighbour ratio <= 1.3), 8-cell CPML on all six faces, seeded FR-4
  blocks (eps_r 4.3, kappa 0.02 S/m, 10-60 cells) in the interior, a soft
  Gaussian Ez source (5 GHz / 4 GHz) at the centre and Ex/Ey/Ez voltage
  probes at 16 seeded points.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import engine as E


@dataclass
class Workload:
    name: str
    grid: E.Grid
    dt: float
    lanes: np.ndarray
    build_ops: object  # callable(engine) -> list of probes
    steps: int
    cells: int


def graded_lines(n_cells: int, rng: np.random.Generator, lo=0.5e-3, hi=1.5e-3, ratio=1.3):
    """Seeded spacings, each drawn uniform in [lo, hi] and clamped to within
    `ratio` of its neighbour."""
    d = np.empty(n_cells)
    d[0] = rng.uniform(lo, hi)
    t = rng.uniform(lo, hi, size=n_cells)
    for q in range(1, n_cells):
        d[q] = min(max(t[q], d[q - 1] / ratio), d[q - 1] * ratio)
    return np.concatenate([[0.0], np.cumsum(d)])


def vacuum_box(n: int = 50, steps: int = 2000) -> Workload:
    g = E.Grid.uniform([n, n, n], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    c = n // 2

    def ops(e):
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(2e9, 1e9)).edge(E.Axis.Z, [c, c, c], 1.0))
        return [e.add_operator(E.Probe.voltage(E.Axis.Z, [n // 5, c, c]))]

    return Workload(f"vacuum_box_{n}", g, dt, lanes, ops, steps, g.num_cells())


def cpml_box(n: int = 512, steps: int = 5000, seed: int = 2026, blocks: int = 64,
             pml: int = 8, threads: int = 0) -> Workload:
    rng = np.random.default_rng(seed)
    g = E.Grid([graded_lines(n, rng) for _ in range(3)])
    m = E.CellMaterials.vacuum(g)
    fr4 = E.Material.dielectric(4.3).with_kappa(0.02)
    lo_c, hi_c = pml + 2, n - pml - 2
    for _ in range(blocks):
        size = rng.integers(10, 61, size=3)
        size = np.minimum(size, hi_c - lo_c)
        start = [int(rng.integers(lo_c, hi_c - size[a] + 1)) for a in range(3)]
        m.fill_box([(start[a], start[a] + int(size[a])) for a in range(3)], fr4)
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, m, [E.FaceCondition.Pec] * 6, dt, threads=threads)
    c = n // 2
    pts = [[int(rng.integers(lo_c, hi_c)) for _ in range(3)] for _ in range(16)]

    def ops(e):
        e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(pml)) for f in E.Face]))
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(5e9, 4e9)).edge(E.Axis.Z, [c, c, c], 1.0))
        return [e.add_operator(E.Probe.voltage(a, p)) for p in pts for a in range(3)]

    return Workload(f"cpml_box_{n}", g, dt, lanes, ops, steps, g.num_cells())


@dataclass
class PatchModel:
    """One model of the patch-antenna sweep."""
    name: str
    grid: E.Grid
    dt: float
    lanes: np.ndarray
    build_ops: object  # callable(engine) -> the LumpedPort
    cells: int
    params: dict


def patch_model(patch_mm: float, sub_mm: float, feed_mm: float, air_cells: int,
                cell: float = 0.5e-3, pml: int = 8) -> PatchModel:
    """A square patch on 2 mm FR-4 (eps_r 4.3) over a PEC ground (the z_min
    face), 8-cell CPML on the other faces, a 50-ohm lumped port from the
    ground to the patch `feed_mm` off its centre along x, Gaussian
    excitation 1-3 GHz (gaussian(2e9, 1e9)). The substrate spans the whole
    lateral domain (sub_mm square); air_cells of air above it."""
    from .lumped import LumpedBox, LumpedPort
    n_xy = int(round(sub_mm * 1e-3 / cell))
    t_sub = int(round(2e-3 / cell))
    nz = t_sub + air_cells
    g = E.Grid.uniform([n_xy, n_xy, nz], [cell] * 3)
    m = E.CellMaterials.vacuum(g)
    m.fill_box([(0, n_xy), (0, n_xy), (0, t_sub)], E.Material.dielectric(4.3))
    half = int(round(0.5 * patch_mm * 1e-3 / cell))
    c = n_xy // 2
    m.fill_pec_box([(c - half, c + half + 1), (c - half, c + half + 1), (t_sub, t_sub + 1)])
    faces = [E.FaceCondition.Pec] * 6
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, m, faces, dt)
    fx = c + int(round(feed_mm * 1e-3 / cell))

    def ops(e):
        layers = [(f, E.Cpml.with_cells(pml)) for f in E.Face if f != E.Face.ZMin]
        e.add_operator(E.CpmlBoundary.new(e, layers))
        return e.add_operator(LumpedPort(LumpedBox(E.Axis.Z, [fx, c, 0], [fx, c, t_sub]), 50.0)
                              .excite_waveform(E.Waveform.gaussian(2e9, 1e9), 1.0))

    return PatchModel(f"patch_{patch_mm:.1f}_{sub_mm:.0f}_{feed_mm:.1f}_{nz}", g, dt, lanes, ops,
                      g.num_cells(), {"patch_mm": patch_mm, "sub_mm": sub_mm, "feed_mm": feed_mm,
                                      "cells": [n_xy, n_xy, nz]})


def patch_sweep(count: int = 256, seed: int = 4242) -> List[PatchModel]:
    """The seeded sweep: patch side 8-20 mm, substrate 20-60 mm, feed offset
    0-6 mm (clamped inside the patch), 0.5 mm mesh, 8-20 air cells: grids
    from 40x40x12 to 120x120x24 cells."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        sub = float(rng.uniform(20.0, 60.0))
        patch = float(rng.uniform(8.0, min(20.0, sub - 4.0)))
        feed = float(rng.uniform(0.0, min(6.0, 0.5 * patch - 1.0)))
        air = int(rng.integers(8, 21))
        out.append(patch_model(patch, sub, feed, air))
    return out
