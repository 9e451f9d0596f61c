"""Runs a workload on the device in a process of its own (the library build
is chosen at import: EMS_AMD_FP_CONTRACT) and saves its probe traces:
python fma_case.py <vacuum|cpml|patch> <out.npz>."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np

import ems_amd as E
from ems_amd import _native as nat
from ems_amd import workloads as W


def workload(name):
    if name == "vacuum":
        return W.vacuum_box(50, 2000), 2000
    if name == "cpml":
        return W.cpml_box(n=64, blocks=8, seed=5), 400
    raise ValueError(name)


def main(name, out):
    if name == "patch":
        m = W.patch_model(10.0, 24.0, 2.5, 10)
        e = E.Engine(m.grid, m.dt, m.lanes)
        p = m.build_ops(e)
        e.run(3000)
        traces = [p.voltage(), p.current()]
    else:
        wl, steps = workload(name)
        e = E.Engine(wl.grid, wl.dt, wl.lanes)
        probes = wl.build_ops(e)
        e.run(steps)
        traces = [p.values() for p in probes]
    np.savez(out, lib=np.array(nat.LIB_PATH), *traces)
    e.close()


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
