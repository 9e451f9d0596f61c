"""Parity of the device dispersive media with the oracle: fields, branch
state and energies bit for bit, electric poles on the fused schedule
(u^n read from the previous buffer) and the unfused one, magnetic poles
(always unfused), both integration schemes, and a medium inside a CPML."""
import numpy as np
import pytest

import ems_amd as E
from ems_amd.dispersive import Dispersive, Field, Pole, Scheme

import cases as K
from test_gpu_parity import bits, same_fields

pytestmark = pytest.mark.gpu

SCHEDULES = [(0, 0, 0), (1, 7, 0), (1, 3, 4), (2, 15, 0)]


@pytest.fixture(autouse=True, params=SCHEDULES, ids=lambda s: "mode%d_tj%d_xc%d" % s)
def schedule(request):
    f, t, c = request.param
    old = dict(K.SCHEDULE)
    K.SCHEDULE.update(fused=f, tile_rows=t, x_chunks=c)
    yield
    K.SCHEDULE.update(old)


def same_state(hg, ho):
    for f in (Field.Electric, Field.Magnetic):
        for a, b in zip(hg.state(f), ho.state(f)):
            assert np.array_equal(bits(np.asarray(a)), bits(np.asarray(b)))


@pytest.mark.parametrize("magnetic", [False, True])
@pytest.mark.parametrize("lossless", [True, False])
@pytest.mark.parametrize("scheme", [Scheme.Trapezoidal, Scheme.Leapfrog])
def test_cavity_poles(magnetic, lossless, scheme):
    (g, hg), (o, ho) = (K.disp_cavity("gpu", lossless=lossless, magnetic=magnetic, scheme=scheme),
                        K.disp_cavity("oracle", threads=2, lossless=lossless, magnetic=magnetic,
                                      scheme=scheme))
    for n in (1, 30, 121):
        g.run(n)
        o.run(n)
        assert np.float64(g.energy()).tobytes() == np.float64(o.energy()).tobytes()
    for _ in range(3):
        assert np.float64(g.step_with_energy()).tobytes() == \
            np.float64(o.step_with_energy()).tobytes()
    same_fields(g, o)
    same_state(hg, ho)


@pytest.mark.parametrize("scheme", [Scheme.Trapezoidal, Scheme.Leapfrog])
def test_line_medium_into_cpml(scheme):
    """Drude half-space running into the far CPML, soft source, probe."""
    pole = Pole.drude(2.0 * np.pi * 4e9 * np.sqrt(2.0), 1.0 / 0.2e-9)
    med = (2.0, 1.0, [pole], [Pole.lorentz(1e10, 3e10, 1e9)], 100)
    (g, pg), (o, po) = (K.line(kind, 172, 0.5e-3, 24, 48, [88], medium=med, steps=600,
                               scheme=scheme) for kind in ("gpu", "oracle"))
    g.run(600)
    o.run(600)
    assert np.array_equal(bits(pg[0].values()), bits(o.probe_values(po[0])))
    same_fields(g, o)


def test_poles_and_lumped_port_on_one_grid():
    """A dispersive block and a lumped port: the port's u^n and u* see the
    pole-corrected field (slab hook before the serial post_e)."""
    from ems_amd.lumped import LumpedBox, LumpedPort

    def build(kind):
        e, h = K.disp_cavity(kind, threads=2, lossless=False)
        p = e.add_operator(LumpedPort(LumpedBox(E.Axis.Z, [3, 3, 1], [3, 3, 4]), 50.0)
                           .excite_waveform(E.Waveform.gaussian(5e9, 4e9), 1.0))
        return e, h, p
    (g, hg, pg), (o, ho, po) = build("gpu"), build("oracle")
    g.run(400)
    o.run(400)
    assert np.array_equal(bits(pg.voltage()), bits(po.voltage()))
    assert np.array_equal(bits(pg.current()), bits(po.current()))
    same_fields(g, o)
    same_state(hg, ho)
