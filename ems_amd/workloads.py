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
