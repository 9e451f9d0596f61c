"""The openEMS fixture gauss_soft_e on the device: openEMS's curl order,
coefficient rounding, Gaussian timing and f32 current sum, bit for bit
against the oracle (every probe sample and field) on each kernel schedule,
and the probe files to their printed digits."""
import numpy as np
import pytest

import ems_amd as E
import openems_case as OC

pytestmark = pytest.mark.gpu

SCHEDULES = [(0, 0, 0), (1, 3, 0), (2, 7, 4), (3, 7, 0), (4, 7, 0), (5, 7, 2), (6, 4, 2), (6, 8, 0)]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.fixture(scope="module")
def oracle_run():
    from oracle.bridge import OracleEngine
    g, dt = OC.grid(), OC.dt()
    o = OracleEngine(g, dt, OC.lanes(g, dt), curl_order=1, threads=2)
    probes = OC.build(o, g)
    o.run(OC.STEPS)
    return o, {k: [o.probe_values(p) for p in ps] for k, ps in probes.items()}


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "mode%d_tj%d_xc%d" % s)
def test_gauss_soft_e_device(oracle_run, sched):
    o, want = oracle_run
    g, dt = OC.grid(), OC.dt()
    e = E.Engine(g, dt, OC.lanes(g, dt), curl_order=1)
    e.set_schedule(*sched)
    probes = OC.build(e, g)
    e.run(OC.STEPS)
    for name, ps in probes.items():
        got = [p.values() for p in ps]
        for a, b in zip(got, want[name]):
            assert a.size == b.size == OC.STEPS
            assert np.array_equal(bits(a), bits(b)), name
        ref = OC.read_rows(name)
        rows = np.stack([OC.sampled(name, v) for v in got], axis=1)
        assert OC.rel_l2(ref[:, 1:].ravel(), rows.ravel()) < 1e-10, name
    for a in range(3):
        assert np.array_equal(bits(e.u_component(a)), bits(o.u_component(a)))
        assert np.array_equal(bits(e.i_component(a)), bits(o.i_component(a)))
    e.close()
