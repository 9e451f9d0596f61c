"""ems_gpu_energy_noalloc (the energy evaluation whose scratch is a static
device array) against ems_gpu_energy, through the C ABI: bit for bit, on
two calls in a row, the second on a smaller grid; then the grid limit and
the operators it refuses."""
import ctypes as C

import numpy as np
import pytest

import ems_amd as E
from ems_amd import _native as nat
from ems_amd import workloads as W

pytestmark = pytest.mark.gpu


def energies(e):
    a, b = C.c_double(), C.c_double()
    assert nat.ems_gpu_energy(e._h, C.byref(a)) == nat.EMS_OK
    rc = nat.ems_gpu_energy_noalloc(e._h, C.byref(b))
    assert rc == nat.EMS_OK, nat.ems_gpu_last_error(e._h).decode()
    return a.value, b.value


def same(a, b):
    return np.float64(a).tobytes() == np.float64(b).tobytes()


def test_noalloc_energy_matches_twice_then_smaller():
    big = W.cpml_box(n=48, blocks=6, pml=8)
    small = W.cpml_box(n=20, blocks=2, pml=4, seed=7)
    e1 = E.Engine(big.grid, big.dt, big.lanes)
    big.build_ops(e1)
    e1.run(40)
    w_ref, w_na = energies(e1)
    assert w_ref > 0.0 and same(w_ref, w_na), (w_ref, w_na)
    e2 = E.Engine(small.grid, small.dt, small.lanes)
    small.build_ops(e2)
    e2.run(25)
    w_ref2, w_na2 = energies(e2)
    assert w_ref2 > 0.0 and same(w_ref2, w_na2), (w_ref2, w_na2)
    # the first engine again, after the smaller call left its contents
    e1.run(3)
    w_ref3, w_na3 = energies(e1)
    assert same(w_ref3, w_na3), (w_ref3, w_na3)
    e1.close()
    e2.close()


def test_noalloc_energy_refuses_grids_above_its_scratch():
    # nx * (3*ny - 2) + 13*nx + 1 > 2^21 doubles
    g = E.Grid.uniform([999, 699, 2], [1e-3] * 3)
    dt = E.TimeStep.courant(0.99).resolve(g)
    lanes = E.coefficients(g, None, [E.FaceCondition.Pec] * 6, dt)
    nx, ny = g.nodes()[:2]
    assert nx * (3 * ny - 2) + 13 * nx + 1 > nat_limit()
    e = E.Engine(g, dt, lanes)
    w = C.c_double(123.0)
    assert nat.ems_gpu_energy_noalloc(e._h, C.byref(w)) == nat.EMS_E_INVALID
    assert "static array" in nat.ems_gpu_last_error(e._h).decode()
    assert w.value == 123.0
    e.close()


def nat_limit():
    return 2 * 1024 * 1024
