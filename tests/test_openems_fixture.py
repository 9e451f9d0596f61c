"""The oracle's openEMS curl order, openEMS Gaussian and f32 current sum
pinned by the reference's own openEMS fixture ``gauss_soft_e`` (openEMS v0.37
output; checked the way tests/openems_core.rs:96-216 checks ems): the time
step against the printed one, the excitation samples et/ht, and every probe
file's values by relative L2. The reference accepts 1e-4; in openEMS's curl
order the oracle reproduces the files to their 12 printed digits (~1e-12),
and each of the pinned choices is shown to matter: the Paired order misses
by ~1e-7..5e-6, an f64 loop sum by ~7e-8."""
import numpy as np
import pytest

import ems_amd as E
import openems_case as OC


def console_dt():
    with open(OC.GOLDEN + "/console.txt") as f:
        for line in f:
            if "FDTD timestep is:" in line:
                return float(line.split("FDTD timestep is:")[1].split()[0])
    raise AssertionError("no time step in console.txt")


def test_time_step_is_openems():
    dt = OC.dt()
    assert abs(dt / console_dt() - 1.0) < 6e-6
    # Uniform cube: the Rennings step reduces to 1/sqrt(3 X) (the exact
    # Courant limit up to the f32 couplings).
    c0 = 1.0 / np.sqrt(OC.OE_EPS0 * OC.OE_MU0)
    courant = 1e-3 / (c0 * np.sqrt(3.0))
    assert abs(dt / courant - 1.0) < 1e-6


@pytest.mark.parametrize("name,offset", [("et", 0.0), ("ht", 0.5)])
def test_excitation_samples(name, offset):
    """tests/openems_core.rs:121-150: sample k of the openEMS Gaussian at
    (k + offset) dt against the signal file."""
    dt = OC.dt()
    wf = E.Waveform.openems_gaussian(1.5e9, 1.5e9)
    rows = OC.read_rows(name)
    assert rows.shape == (400, 2)
    peak = np.max(np.abs(rows[:, 1]))
    for k, (t, v) in enumerate(rows):
        tk = (k + offset) * dt
        assert abs(t - tk) <= 1e-5 * max(tk, dt)
        got = wf.value(tk, dt)
        assert abs(got - v) <= 2e-5 * peak + 1e-6 * abs(v), (k, got, v)
    assert wf.value(0.0, dt) == 0.0  # sample 0 forced to zero


def run_oracle(curl_order, f32_loop=True):
    from oracle.bridge import OracleEngine
    g, dt = OC.grid(), OC.dt()
    o = OracleEngine(g, dt, OC.lanes(g, dt), curl_order=curl_order, threads=2)
    probes = OC.build(o, g)
    if not f32_loop:  # the same loop summed in f64
        it = probes["it1"][0]
        p = E.Probe(E.ProbeQuantity.Current)
        for a, n, w in it.entries:
            p.entry(a, n, w)
        probes["it1_f64"] = [o.add_operator(p)]
    o.run(OC.STEPS)
    out = {}
    for name, ps in probes.items():
        out[name] = np.stack([OC.sampled(name.split("_")[0], o.probe_values(p)) for p in ps],
                             axis=1)
    return out


@pytest.fixture(scope="module")
def openems_order():
    return run_oracle(1, f32_loop=False)


@pytest.mark.parametrize("name", ["ut1", "it1", "et1", "ht1"])
def test_probe_files_to_printed_digits(openems_order, name):
    ref = OC.read_rows(name)
    got = openems_order[name]
    assert got.shape == (ref.shape[0], ref.shape[1] - 1) == (20, got.shape[1])
    e = OC.rel_l2(ref[:, 1:].ravel(), got.ravel())
    assert e < 1e-10, f"{name}: rel. L2 {e:.3e}"


def test_choices_are_discriminated(openems_order):
    """The fixture tells the pinned choices from their alternatives."""
    paired = run_oracle(0)
    for name in ("ut1", "it1", "et1", "ht1"):
        ref = OC.read_rows(name)[:, 1:].ravel()
        e_oe = OC.rel_l2(ref, openems_order[name].ravel())
        e_pa = OC.rel_l2(ref, paired[name].ravel())
        assert e_pa > 1e-8 and e_pa > 1e3 * e_oe, (name, e_pa, e_oe)
    ref = OC.read_rows("it1")[:, 1]
    assert OC.rel_l2(ref, openems_order["it1_f64"][:, 0]) > 1e-8
