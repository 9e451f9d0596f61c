"""The openEMS fixture ``pml_face`` (tests/fixtures/openems/pml_face of the
reference: openEMS v0.37 output of a 30 x 10 x 10 mm vacuum box, 1 mm
cells, PML_8 on x_max, PMC on both y faces, PEC elsewhere, a soft Gaussian
E excitation over a box with the oblique direction (0.3, 0.5, 1), 1000
steps; probes outside the layer and 2 and 5 cells deep in it; data copied
under tests/golden/openems_pml_face), built as the reference's file path
builds it for any engine (device or oracle): openEMS's time step,
CoeffRounding::OpenEms coefficients, the openEMS curl order, openEMS's PMC
half cell (compile/mod.rs:1572-1630) and upper-face normal H
(:1709-1730), then the layer operator OpenEmsPml::with_materials
(compile/mod.rs:1224-1231) and the excitation and probes of the model."""
from __future__ import annotations

import os

import numpy as np

import ems_amd as E
from ems_amd import engine as EE

import openems_case as OC

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "openems_pml_face")

STEPS = 1000
INTERVAL = 6  # the probes' sampling interval (1.1555e-11 s / dt in the files)
FILES = ["ut", "et_out", "ht_out", "et_d2", "ht_d2", "et_d5", "ht_d5"]
# ProbeBox points of model.xml (mm = node index on this 1 mm grid)
POINTS = {"out": (18, 7, 6), "d2": (24, 4, 3), "d5": (27, 6, 7)}
PML_CELLS = 8

F = E.Face
PMC, PEC = E.FaceCondition.Pmc, E.FaceCondition.Pec
FACES = [PEC, PEC, PMC, PMC, PEC, PEC]  # x_max is the PML face, built as PEC


def grid():
    return E.Grid([np.array([q * 1e-3 for q in range(31)]),
                   np.array([q * 1e-3 for q in range(11)]),
                   np.array([q * 1e-3 for q in range(11)])])


def dt():
    return float(OC.rennings_dt_uniform([1e-3, 1e-3, 1e-3]))


def pmc_half_cell(lanes, n, faces):
    """openEMS's PMC (pmc_half_cell, compile/mod.rs:1572-1630): on each PMC
    face zero the tangential H of the dual layer next to it, the tangential
    E and the normal H on the face (both coefficients of each edge)."""
    for f, cond in enumerate(faces):
        if cond != PMC:
            continue
        t = f // 2
        h_layer, e_layer = (n[t] - 2, n[t] - 1) if f % 2 else (0, 0)
        zero = []
        for a in range(3):
            if a == t:
                zero.append((False, a, e_layer))
            else:
                zero += [(False, a, h_layer), (True, a, e_layer)]
        for e_side, a, layer in zero:
            l0 = a if e_side else 6 + a
            for l in (l0, l0 + 3):
                sl = [slice(None)] * 3
                sl[t] = layer
                lanes[l][tuple(sl)] = 0.0


def upper_face_normal_h(lanes, n, faces):
    """upper_face_normal_h (compile/mod.rs:1709-1730): da = db = 0 for the
    normal H on the last node plane of each listed face."""
    for f in faces:
        t = int(F(f).axis)
        sl = [slice(None)] * 3
        sl[t] = n[t] - 1
        lanes[6 + t][tuple(sl)] = 0.0
        lanes[9 + t][tuple(sl)] = 0.0


def lanes(g, step):
    """The coefficients after the compile step's openEMS adjustments (the
    upper faces: the PML face and the PEC z_max, compile/mod.rs:1199-1218)."""
    ln = E.coefficients(g, None, FACES, step, rounding=1)
    n = g.nodes()
    pmc_half_cell(ln, n, FACES)
    upper_face_normal_h(ln, n, [F.XMax, F.ZMax])
    return ln


def excitation_edges(g):
    """The soft E excitation over the box (6, 2, 3)-(7, 3, 5) mm with the
    direction (0.3, 0.5, 1): every edge along a whose span lies in the box
    (closed), weight direction[a] times the edge length."""
    direction = (0.3, 0.5, 1.0)
    lo, hi = (6, 2, 3), (7, 3, 5)
    out = []
    for a in range(3):
        rng = [range(lo[b], hi[b] + 1) if b != a else range(lo[b], hi[b]) for b in range(3)]
        for i in rng[0]:
            for j in rng[1]:
                for k in rng[2]:
                    node = (i, j, k)
                    out.append((a, node, direction[a] * g.delta(a, node[a])))
    return out


def build(engine, g, pml=True):
    """Registers the layers (first), the excitation and the probes (ut,
    then the E and H probes in file order); returns (pml, probes)."""
    op = None
    if pml:
        op = E.OpenEmsPml.with_materials(engine, [(F.XMax, PML_CELLS)], 1.0, 1.0)
        engine.add_operator(op)
    wave = E.Waveform.openems_gaussian(10e9, 9e9)
    src = E.EdgeSource(EE.OpenEmsSignal(wave)).kind(E.SourceKind.SoftE)
    for a, node, w in excitation_edges(g):
        src.edge(a, node, w)
    engine.add_operator(src)
    probes = {}
    ut = E.Probe(E.ProbeQuantity.Voltage)
    for k in range(10):
        ut.entry(E.Axis.Z, [14, 3, k], 1.0)
    probes["ut"] = [engine.add_operator(ut)]
    for name in FILES[1:]:
        node = POINTS[name.split("_")[1]]
        if name.startswith("et"):
            probes[name] = [engine.add_operator(E.Probe(E.ProbeQuantity.Voltage).entry(
                a, node, 1.0 / g.delta(a, node[a]))) for a in range(3)]
        else:
            dual = tuple(v - 1 for v in node)
            probes[name] = [engine.add_operator(E.Probe(E.ProbeQuantity.Current).entry(
                a, dual, 1.0 / g.dual_delta(a, dual[a]))) for a in range(3)]
    return op, probes


def read_rows(name):
    rows = []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            if line.startswith("%") or not line.strip():
                continue
            rows.append([float(v) for v in line.split()])
    return np.array(rows)


def sampled(name, values):
    """The file's rows from per-step samples: E-time files at t = 6 m dt
    (row 0 the initial zero, row m > 0 the sample of step 6 m - 1), H-time
    files at (6 m + 1/2) dt (the sample of step 6 m)."""
    rows = STEPS // INTERVAL + 1
    if name.startswith("u") or name.startswith("et"):
        return np.array([0.0] + [values[INTERVAL * m - 1] for m in range(1, rows)])
    return np.array([values[INTERVAL * m] for m in range(rows)])


def columns(probes, engine_values):
    """{file: (rows, columns)} from the registered probes."""
    out = {}
    for name, ps in probes.items():
        out[name] = np.stack([sampled(name, engine_values(p)) for p in ps], axis=1)
    return out


def rel_l2_worst(ref, got):
    """tests/openems_pml.rs:105-138: every value column against the norm of
    the largest one; the worst column."""
    n = min(ref.shape[0], got.shape[0])
    a, b = ref[:n], got[:n]
    scale = max(float(np.linalg.norm(a[:, c])) for c in range(a.shape[1]))
    return max(float(np.linalg.norm(a[:, c] - b[:, c])) / max(scale, 1e-300)
               for c in range(a.shape[1]))
