"""The openEMS fixture ``pml_end`` (tests/fixtures/openems/pml_end of the
reference: openEMS v0.37 output of a patch antenna on a 4 mm FR-4-like
substrate (eps_r 4, kappa 0.01) in a 40 x 32 x 24 mm box of 2 mm cells,
PML_6 on every face but z_min (PEC), a soft Gaussian E_z feed; openEMS
stopped it by its energy criterion (-50 dB) after 1956 steps at
-50.5179 dB; data copied under tests/golden/openems_pml_end), built as the
reference's file path builds it for any engine (device or oracle), with
openEMS's run limits (sim/mod.rs:915-929): the energy checked every
Nyquist interval after the steps (check_after_interval), once more after
half the Gaussian's length, in openEMS's measure."""
from __future__ import annotations

import math
import os

import numpy as np

import ems_amd as E
from ems_amd import engine as EE

import openems_case as OC
import openems_pml_case as PC

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "openems_pml_end")

F = E.Face
PEC = E.FaceCondition.Pec
FACES = [PEC] * 6            # the PML faces are built as PEC
PML = [(F.XMin, 6), (F.XMax, 6), (F.YMin, 6), (F.YMax, 6), (F.ZMax, 6)]
MAX_STEPS = 20000            # NumberOfTimesteps
END_RATIO = 1e-5             # endCriteria
F0, FC = 5e9, 5e9
FILES = ["ut", "et", "ht"]
STOP_STEPS = 1956            # meta.txt: "after 1956 timesteps (-50.5179dB)"
STOP_DB = -50.5179


def grid():
    return E.Grid([np.array([2e-3 * q for q in range(21)]),
                   np.array([2e-3 * q for q in range(17)]),
                   np.array([2e-3 * q for q in range(13)])])


def dt():
    return float(OC.rennings_dt_uniform([2e-3, 2e-3, 2e-3]))


def materials(g):
    """The substrate (z 0..4 mm: cells k = 0, 1) and the metal patch (the
    sheet z = 4 mm over x 12..28, y 10..22 mm: nodes 6..14, 5..11, k = 2)."""
    m = E.CellMaterials.vacuum(g)
    m.fill_box([(0, 20), (0, 16), (0, 2)], E.Material(eps_r=(4.0,) * 3, kappa=(0.01,) * 3))
    m.fill_pec_box([(6, 15), (5, 12), (2, 3)])
    return m


def lanes(g, step):
    ln = E.coefficients(g, materials(g), FACES, step, rounding=1)
    PC.upper_face_normal_h(ln, g.nodes(), [F.XMax, F.YMax, F.ZMax])
    return ln


def build(engine, g):
    """The layers (with the edge materials), the feed (E_z at x 14, y 16 mm
    over z 0..4 mm) and the probes ut (z 0..4 mm at x 18, y 16 mm), et at
    (20, 16, 10) mm, ht at (30, 16, 10) mm (dual node one below)."""
    eps, mu = E.edge_materials(g, materials(g))
    op = E.OpenEmsPml.with_materials(engine, PML, eps, mu)
    engine.add_operator(op)
    wave = E.Waveform.openems_gaussian(F0, FC)
    src = E.EdgeSource(EE.OpenEmsSignal(wave)).kind(E.SourceKind.SoftE)
    for k in (0, 1):
        src.edge(E.Axis.Z, [7, 8, k], 1.0 * g.delta(2, k))
    engine.add_operator(src)
    probes = {}
    ut = E.Probe(E.ProbeQuantity.Voltage)
    for k in (0, 1):
        ut.entry(E.Axis.Z, [9, 8, k], 1.0)
    probes["ut"] = [engine.add_operator(ut)]
    node = (10, 8, 5)
    probes["et"] = [engine.add_operator(E.Probe(E.ProbeQuantity.Voltage).entry(
        a, node, 1.0 / g.delta(a, node[a]))) for a in range(3)]
    dual = (14, 7, 4)
    probes["ht"] = [engine.add_operator(E.Probe(E.ProbeQuantity.Current).entry(
        a, dual, 1.0 / g.dual_delta(a, dual[a]))) for a in range(3)]
    return op, probes


def limits(step):
    """openEMS's run limits: (check interval, extra check, probe interval)."""
    wave = E.Waveform.openems_gaussian(F0, FC)
    k = int(math.floor(1.0 / (2.0 * (F0 + FC) * step)))
    span = int(math.floor(wave.duration() / step)) + 1
    return k, span // 2, max(1, k // 4)


def drive(engine, energy, step):
    """run::drive with check_after_interval (run.rs:424-540) over the
    engine's `run` and an energy function of the state; returns (steps,
    level in dB)."""
    k, extra, _ = limits(step)
    steps, w_max = 0, 0.0
    while steps < MAX_STEPS:
        regular = (steps // k + 1) * k
        nxt = extra if steps < extra < regular else regular
        nxt = min(nxt, MAX_STEPS)
        engine.run(nxt - steps)
        steps = nxt
        if steps % k != 0 and steps != extra:
            continue
        w = energy()
        w_max = max(w_max, w)
        ratio = w / w_max if w_max > 0.0 else 1.0
        if not math.isfinite(w) or w < -1e-12 * w_max:
            raise AssertionError(f"diverged at {steps}")
        if w_max > 0.0 and ratio < END_RATIO:
            return steps, 10.0 * math.log10(ratio)
    return steps, None


def read_rows(name):
    rows = []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            if line.startswith("%") or not line.strip():
                continue
            rows.append([float(v) for v in line.split()])
    return np.array(rows)


def sampled(name, values, interval, rows):
    if name in ("ut", "et"):
        return np.array([0.0] + [values[interval * m - 1] for m in range(1, rows)])
    return np.array([values[interval * m] for m in range(rows)])
