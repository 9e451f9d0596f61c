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
import math
from typing import List

import numpy as np

from . import engine as E
from .raster import Solid


@dataclass
class Workload:
    name: str
    grid: E.Grid
    dt: float
    lanes: np.ndarray
    build_ops: object  # callable(engine) -> list of probes
    steps: int
    cells: int
    # conformal metal: whether the stability estimate ran (None: no metal)
    stability_checked: object = None


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
        # the block between node lines start and start + size: its cells by
        # their centres (the host library's rasteriser)
        m.fill_solid(Solid.box([g.lines[a][start[a]] for a in range(3)],
                               [g.lines[a][start[a] + int(size[a])] for a in range(3)]), fr4)
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, m, [E.FaceCondition.Pec] * 6, dt, threads=threads)
    c = n // 2
    pts = [[int(rng.integers(lo_c, hi_c)) for _ in range(3)] for _ in range(16)]

    def ops(e):
        e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(pml)) for f in E.Face]))
        e.add_operator(E.EdgeSource(E.Waveform.gaussian(5e9, 4e9)).edge(E.Axis.Z, [c, c, c], 1.0))
        return [e.add_operator(E.Probe.voltage(a, p)) for p in pts for a in range(3)]

    return Workload(f"cpml_box_{n}", g, dt, lanes, ops, steps, g.num_cells())


def dispersive_box(n=(384, 384, 256), steps: int = 8000, seed: int = 2027, pml: int = 10,
                   cell: float = 1e-3, lorentz: int = 32, drude: int = 16,
                   cylinder_cells: float = 40.0, radii=(6.0, 14.0),
                   skip_stability=None) -> Workload:
    """Dispersive media and conformal metal: an n-cell uniform grid
    (cell mm), `pml`-cell CPML on all faces, `lorentz` seeded spheres of a
    single-pole Lorentz medium (eps_inf 2, delta-eps 3, f0 8 GHz, damping
    0.5 GHz: omega_p = sqrt(3) omega_0, gamma = 2 * 2 pi 0.5 GHz) and
    `drude` of a Drude medium (f_p 20 GHz, gamma 2 pi 1 GHz), radii in
    `radii` cells, outside a conformal PEC cylinder of radius `cylinder_cells`
    along z at the centre (flat caps inside the TF/SF box, whose surface
    stays in vacuum); an x-polarised TF/SF plane wave travelling +y,
    Gaussian 1-10 GHz (gaussian(5.5e9, 4.5e9)); energy every 50 steps."""
    from . import conformal as CF
    from . import dispersive as DP
    from . import planewave as PW
    rng = np.random.default_rng(seed)
    g = E.Grid.uniform(list(n), [cell] * 3)
    m = E.CellMaterials.vacuum(g)
    nc = g.cells()
    c = [q / 2.0 for q in nc]
    centres = np.indices(nc, dtype=np.float32)
    centres += 0.5
    lor_mask = np.zeros(nc, dtype=bool)
    dru_mask = np.zeros(nc, dtype=bool)
    lo_c = pml + 6
    placed = tries = 0
    while placed < lorentz + drude:
        tries += 1
        if tries > 100000:
            raise ValueError(f"dispersive_box: cannot place {lorentz + drude} spheres of radius "
                             f"{radii} in {n} cells off a {cylinder_cells}-cell cylinder")
        r = float(rng.uniform(radii[0], radii[1]))
        p = [float(rng.uniform(lo_c + r, nc[a] - lo_c - r)) for a in range(3)]
        if math.hypot(p[0] - c[0], p[1] - c[1]) < cylinder_cells + r + 2.0:
            continue  # keep the spheres off the cylinder
        if any(p[a] - r < pml + 6 or p[a] + r > nc[a] - pml - 6 for a in range(3)):
            continue
        lo = [max(int(p[a] - r) - 1, 0) for a in range(3)]
        hi = [min(int(p[a] + r) + 2, nc[a]) for a in range(3)]
        sub = tuple(slice(lo[a], hi[a]) for a in range(3))
        d2 = sum((centres[a][sub] - p[a]) ** 2 for a in range(3))
        target = lor_mask if placed < lorentz else dru_mask
        target[sub] |= d2 < r * r
        placed += 1
    lor_mask &= ~dru_mask
    m_lor = E.Material.dielectric(2.0)
    idx = np.nonzero(lor_mask)
    for q in zip(*idx):
        m.set(q, m_lor)
    pw_lo, pw_hi = pml + 4, [q - pml - 4 for q in g.nodes()]
    # The stability estimate (300 power iterations over every edge, one
    # host core) takes about nine minutes at 385 x 385 x 257 nodes. The
    # default geometry passes it in the first round (s = 0.9907, measured
    # once), so the engine is the same without it: only that exact model
    # skips it (skip_stability=None), and the workload records whether it
    # was checked. Any other geometry runs it unless the caller says so.
    default = (tuple(n), seed, pml, cell, lorentz, drude, cylinder_cells, tuple(radii)) == (
        (384, 384, 256), 2027, 10, 1e-3, 32, 16, 40.0, (6.0, 14.0))
    skip = default if skip_stability is None else bool(skip_stability)
    opts = CF.ConformalOptions(iterations=0) if skip else None
    pec = CF.cylinder_z(g, (c[0] * cell, c[1] * cell), cylinder_cells * cell, mats=m,
                        options=opts, metal="inside", z_nodes=(pw_lo + 4, pw_hi[2] - 4))
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, m, [E.FaceCondition.Pec] * 6, dt)
    w0 = 2.0 * math.pi * 8e9
    lor = DP.Pole.lorentz(math.sqrt(3.0) * w0, w0, 2.0 * 2.0 * math.pi * 0.5e9)
    dru = DP.Pole.drude(2.0 * math.pi * 20e9, 2.0 * math.pi * 1e9)
    wave = E.Waveform.gaussian(5.5e9, 4.5e9)

    def ops(e):
        pec.apply(e)
        d = DP.Dispersive.new()
        d.fill_mask(g, lor_mask, DP.Field.Electric, [lor])
        d.fill_mask(g, dru_mask, DP.Field.Electric, [dru])
        e.add_operator(E.CpmlBoundary.new(e, [(f, E.Cpml.with_cells(pml)) for f in E.Face]))
        e.add_operator(d)
        spec = PW.PlaneWaveSpec.new([pw_lo] * 3, pw_hi, [0.0, 1.0, 0.0], [1.0, 0.0, 0.0])
        e.add_operator(PW.PlaneWave.new(e, spec, lambda t: wave.value(t)))
        # between the cylinder and the far face of the TF/SF box
        y = int((c[1] + cylinder_cells + pw_hi[1]) / 2)
        return [e.add_operator(E.Probe.voltage(E.Axis.X, [int(c[0]), y, int(c[2])]))]

    return Workload(f"dispersive_box_{n[0]}x{n[1]}x{n[2]}", g, dt, lanes, ops, steps,
                    g.num_cells(), stability_checked=not skip)


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
