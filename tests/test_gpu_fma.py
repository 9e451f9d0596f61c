"""The build with FMA contraction (EMS_AMD_FP_CONTRACT=fast,
libems_gpu_fma.so): probe traces within 1e-6 of the reference engine's
(the oracle's), relative over all of a model's probe samples (L2, and the
largest deviation against the largest magnitude), on the 50^3
vacuum box (2000 steps), a 64^3 graded CPML box with FR-4 blocks (400
steps) and a patch antenna's lumped port (3000 steps). The device runs in a
process of its own, since the build is chosen when the library loads."""
import os
import subprocess
import sys

import numpy as np
import pytest

import ems_amd as E
from ems_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
REL = 1e-6


def device_traces(name, tmp_path):
    out = str(tmp_path / f"{name}.npz")
    env = dict(os.environ, EMS_AMD_FP_CONTRACT="fast")
    r = subprocess.run([sys.executable, os.path.join(HERE, "fma_case.py"), name, out], env=env,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    z = np.load(out)
    assert str(z["lib"]).endswith("libems_gpu_fma.so")
    return [z[f"arr_{q}"] for q in range(len(z.files) - 1)]


def oracle_traces(name):
    from oracle.bridge import OracleEngine
    if name == "patch":
        m = W.patch_model(10.0, 24.0, 2.5, 10)
        o = OracleEngine(m.grid, m.dt, m.lanes, threads=4)
        p = m.build_ops(o)
        o.run(3000)
        out = [p.voltage().copy(), p.current().copy()]
    else:
        import fma_case
        wl, steps = fma_case.workload(name)
        o = OracleEngine(wl.grid, wl.dt, wl.lanes, threads=4)
        probes = wl.build_ops(o)
        o.run(steps)
        out = [o.probe_values(p) for p in probes]
    o.close()
    return out


@pytest.mark.parametrize("name", ["vacuum", "cpml", "patch"])
def test_fma_build_within_1e6(name, tmp_path):
    dev = device_traces(name, tmp_path)
    ref = oracle_traces(name)
    assert len(dev) == len(ref) > 0
    a = np.concatenate([np.asarray(x, dtype=np.float64) for x in dev])
    b = np.concatenate([np.asarray(x, dtype=np.float64) for x in ref])
    assert a.shape == b.shape
    rel_l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    rel_max = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    per = [float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300)) for x, y in zip(dev, ref)]
    print(f"{name}: rel L2 {rel_l2:.3e}, largest deviation / largest magnitude {rel_max:.3e}, "
          f"worst single trace {max(per):.3e}")
    assert rel_l2 <= REL and rel_max <= REL, (rel_l2, rel_max)
    # The build contracts ca*u + cb*curl into fma(ca, u, cb*curl): where
    # ca = da = 1 (the lossless vacuum box) that product is exact and the
    # bits are the exact build's; lossy FR-4 and the CPML recursions make
    # the contraction show.
    if name != "vacuum":
        assert not np.array_equal(a, b)
