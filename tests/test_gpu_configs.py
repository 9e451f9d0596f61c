"""The benchmark workloads themselves against the oracle, through the C ABI.

* The 50^3-cell vacuum PEC box (the reference's examples/bench_vacuum.rs
  case): soft Gaussian Ez source (2 GHz / 1 GHz) at the centre, Ez probe at
  node (10, 25, 25), exactly 2000 steps; the probe trace and every field
  bit for bit, on each kernel schedule.
* The 512^3-cell graded CPML box at full size (513^3 nodes, 8-cell CPML on
  all faces, 64 seeded FR-4 blocks, 48 voltage probes): every field seeded
  with random values (padding entries kept at zero, as the reference keeps
  them) so that every probe, every CPML slab and every FR-4 block is active
  from the first step; 24 steps on the device's default schedule and on the
  unfused one; all 48 probe traces, the f64 energy and all six fields bit
  for bit.
* The 384x384x256-cell dispersive / conformal / plane-wave box at full size
  (385x385x257 nodes: 10-cell CPML, 32 Lorentz and 16 Drude spheres, the
  conformal PEC cylinder with its reduced currents, the TF/SF plane wave):
  every field seeded the same way, so that every operator works from the
  first step; 10 steps on the device's default unfused kernels (loading
  ahead, the CPML-E slabs inside the E kernel) with the energy, operators'
  terms included, after 6 and 10, the probe trace and all six fields bit
  for bit.
"""
import os

import numpy as np
import pytest

import ems_amd as E
from ems_amd import workloads as W

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def oracle_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def same_fields(g, o, label):
    for a in range(3):
        for name, fg, fo in (("u", g.u_component(a), o.u_component(a)),
                             ("i", g.i_component(a), o.i_component(a))):
            bg, bo = bits(fg), bits(fo)
            if not np.array_equal(bg, bo):
                bad = np.argwhere(bg != bo)
                raise AssertionError(f"{label}: {name}[{a}] differs at {bad[:5].tolist()} "
                                     f"({bad.shape[0]} entries)")


VACUUM_SCHEDULES = [(0, 0, 0), (1, 7, 0), (2, 7, 16), (3, 7, 0), (4, 7, 0), (5, 7, 4), (6, 8, 0),
                    (6, 16, 7)]


@pytest.fixture(scope="module")
def vacuum_oracle():
    from oracle.bridge import OracleEngine
    wl = W.vacuum_box(50, 2000)
    o = OracleEngine(wl.grid, wl.dt, wl.lanes, threads=oracle_threads())
    (p,) = wl.build_ops(o)
    o.run(wl.steps)
    return wl, o, o.probe_values(p)


@pytest.mark.parametrize("sched", VACUUM_SCHEDULES, ids=lambda s: "mode%d_tj%d_xc%d" % s)
def test_vacuum_box_50_2000_steps(vacuum_oracle, sched):
    wl, o, trace = vacuum_oracle
    assert wl.grid.nodes() == (51, 51, 51) and wl.steps == 2000
    e = E.Engine(wl.grid, wl.dt, wl.lanes)
    e.set_schedule(*sched)
    (p,) = wl.build_ops(e)
    e.run(wl.steps)
    got = p.values()
    assert got.size == trace.size == 2000
    assert np.array_equal(bits(got), bits(trace)), "Ez probe at (10,25,25) differs"
    assert np.max(np.abs(trace)) > 1e-4  # the pulse reached the probe (peak 6.8e-4)
    same_fields(e, o, "vacuum_box_50")
    e.close()


def seeded_fields(nodes, seed):
    """Random f32 fields in [-1, 1) with every padding entry zero: u_a at the
    last node along a, i_a at the last node along either other axis."""
    rng = np.random.default_rng(seed)
    out = []
    for field in range(2):
        for a in range(3):
            v = rng.uniform(-1.0, 1.0, size=nodes).astype(np.float32)
            axes = [a] if field == 0 else [b for b in range(3) if b != a]
            for b in axes:
                sl = [slice(None)] * 3
                sl[b] = nodes[b] - 1
                v[tuple(sl)] = 0.0
            out.append((field, a, v))
    return out


@pytest.fixture(scope="module")
def cpml_box_512():
    wl = W.cpml_box(512)
    assert wl.grid.nodes() == (513, 513, 513)
    return wl, seeded_fields(wl.grid.nodes(), 512)


@pytest.mark.parametrize("sched", [None, (0, 0, 0), (6, 16, 0), "split", "classes"],
                         ids=["default", "unfused", "band16", "cpml_split", "tile_classes"])
def test_cpml_box_512_full_size(cpml_box_512, sched):
    from oracle.bridge import OracleEngine
    wl, init = cpml_box_512
    steps = 24
    e = E.Engine(wl.grid, wl.dt, wl.lanes)
    if sched == "split":
        e.set_cpml_split(True)  # the z-normal slabs in passes of their own
    elif sched == "classes":
        e.set_tile_classes(True)  # slab tiles and the others in two launches
    elif sched is not None:
        e.set_schedule(*sched)
    pe = wl.build_ops(e)
    for field, a, v in init:
        e.set_component(field, a, v)
    e.setup.clear()  # (the 3 GB of copies are not needed again)
    e.run(steps)
    we = e.energy()
    ge = [p.values() for p in pe]
    o = OracleEngine(wl.grid, wl.dt, wl.lanes, threads=oracle_threads())
    po = wl.build_ops(o)
    for field, a, v in init:
        o.set_component(field, a, v)
    o.run(steps)
    wo = o.energy()
    assert len(pe) == len(po) == 48
    for q, (a, p) in enumerate(zip(ge, po)):
        b = o.probe_values(p)
        assert a.size == b.size == steps
        assert np.array_equal(bits(a), bits(b)), f"probe {q} differs"
    assert np.float64(we).tobytes() == np.float64(wo).tobytes(), (we, wo)
    same_fields(e, o, "cpml_box_512")
    e.close()
    o.close()


def test_coefficient_pool_over_4_gib():
    """A 601^3-node graded box: six dense coefficient lanes make a 5.2 GB
    pool, addressed by the fused kernel through 64-bit row offsets; 3 steps
    from seeded fields bit for bit against the oracle (every field and the
    f64 energy)."""
    from oracle.bridge import OracleEngine
    n = 600
    rng = np.random.default_rng(600)
    lines = [np.concatenate([[0.0], np.cumsum(1e-3 * (0.8 + 0.4 * rng.random(n)))])
             for _ in range(3)]
    g = E.Grid(lines)
    dt = 0.99 * E.courant_limit(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    dense = sum(int(np.any(lanes[l] != lanes[l].reshape(-1)[0])) for l in range(12))
    assert dense >= 6
    assert 4 * dense * g.num_nodes() > (1 << 32)
    init = seeded_fields(g.nodes(), 601)
    e = E.Engine(g, dt, lanes)
    for field, a, v in init:
        e.set_component(field, a, v)
    e.setup.clear()
    e.run(3)
    we = e.energy()
    o = OracleEngine(g, dt, lanes, threads=oracle_threads())
    for field, a, v in init:
        o.set_component(field, a, v)
    del init
    o.run(3)
    wo = o.energy()
    assert np.float64(we).tobytes() == np.float64(wo).tobytes(), (we, wo)
    same_fields(e, o, "pool_over_4_gib")
    e.close()
    o.close()


def test_dispersive_box_full_size():
    from oracle.bridge import OracleEngine
    wl = W.dispersive_box()
    assert wl.grid.nodes() == (385, 385, 257)
    init = seeded_fields(wl.grid.nodes(), 384)
    e = E.Engine(wl.grid, wl.dt, wl.lanes)
    (pe,) = wl.build_ops(e)
    for field, a, v in init:
        e.set_component(field, a, v)
    o = OracleEngine(wl.grid, wl.dt, wl.lanes, threads=oracle_threads())
    (po,) = wl.build_ops(o)
    for field, a, v in init:
        o.set_component(field, a, v)
    for n in (5, 3):
        e.run(n)
        o.run(n)
        we, wo = e.step_with_energy(), o.step_with_energy()
        assert np.float64(we).tobytes() == np.float64(wo).tobytes(), (we, wo)
    a, b = pe.values(), o.probe_values(po)
    assert a.size == b.size == 10 and np.abs(b).max() > 0.0
    assert np.array_equal(bits(a), bits(b)), "probe differs"
    same_fields(e, o, "dispersive_box_384")
    e.close()
    o.close()
