"""The oracle's fused schedule (Pool::fused_a / fused_b, src/engine/pool.rs:
405-594: per thread slab the last plane's H first, then plane by plane H
with its CPML H corrections and E with its CPML E corrections) gives the
unfused schedule's fields, probe traces, port records and energies bit for
bit, at one and several threads (the reference's tests/fused.rs)."""
import numpy as np
import pytest

import ems_amd as E
from ems_amd import workloads as W
from oracle.bridge import OracleEngine


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def run(wl, fused, threads, steps):
    o = OracleEngine(wl.grid, wl.dt, wl.lanes, threads=threads, fused=fused)
    probes = wl.build_ops(o)
    o.run(steps)
    w = o.energy()
    vals = [o.probe_values(p) for p in probes]
    f = [o.u_component(a) for a in range(3)] + [o.i_component(a) for a in range(3)]
    o.close()
    return vals, w, f


@pytest.mark.parametrize("threads", [1, 3, 4])
def test_fused_equals_unfused_cpml_box(threads):
    wl = W.cpml_box(n=40, blocks=8, pml=8, seed=11)
    a = run(wl, 1, threads, 90)
    b = run(wl, 2, threads, 90)
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(bits(x), bits(y))
    assert np.float64(a[1]).tobytes() == np.float64(b[1]).tobytes()
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(bits(x), bits(y))
    assert np.max(np.abs(a[2][2])) > 0.0


def test_fused_equals_unfused_patch_port():
    """A lumped port (pre_e moved before the fused phases, post_e after) and
    CPML on five faces."""
    m = W.patch_model(8.0, 20.0, 1.0, 8)
    out = []
    for fused in (1, 2):
        o = OracleEngine(m.grid, m.dt, m.lanes, threads=3, fused=fused)
        p = m.build_ops(o)
        o.run(400)
        out.append((p.voltage().copy(), p.current().copy(), o.energy(),
                    [o.u_component(a) for a in range(3)]))
        o.close()
    (va, ia, wa, fa), (vb, ib, wb, fb) = out
    assert np.array_equal(bits(va), bits(vb)) and np.array_equal(bits(ia), bits(ib))
    assert np.float64(wa).tobytes() == np.float64(wb).tobytes()
    for x, y in zip(fa, fb):
        assert np.array_equal(bits(x), bits(y))
