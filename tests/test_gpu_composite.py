"""Parity of a composite model with the oracle: CPML, Lorentz and Drude
spheres (dispersive poles), a conformal PEC cylinder with flat caps and an
x-polarised TF/SF plane wave, all at once (the dispersive_box workload at
a reduced size); fields, probe trace and energies bit for bit."""
import numpy as np
import pytest

import ems_amd as E
from ems_amd import workloads as W

from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu


def test_dispersive_conformal_planewave_box():
    wl = W.dispersive_box((48, 48, 32), pml=6, lorentz=2, drude=1, cylinder_cells=4.0,
                          radii=(2.5, 3.5))
    from oracle.bridge import OracleEngine
    g = E.Engine(wl.grid, wl.dt, wl.lanes)
    o = OracleEngine(wl.grid, wl.dt, wl.lanes, threads=4)
    pg = wl.build_ops(g)[0]
    po = wl.build_ops(o)[0]
    for n in (1, 120, 180):
        g.run(n)
        o.run(n)
        wg, wo = g.step_with_energy(), o.step_with_energy()
        assert np.float64(wg).tobytes() == np.float64(wo).tobytes(), (wg, wo)
    assert np.array_equal(bits(pg.values()), bits(o.probe_values(po)))
    assert np.abs(pg.values()).max() > 0.0
    same_fields(g, o)
