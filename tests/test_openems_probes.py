"""The openEMS probe file format (src/io/openems/probes.rs) against
openEMS's own files in the reference's fixtures: every file read and
written back is the same text (the number format is C's %.12g / %.6g, the
header lines and the column layout), and the files made from the oracle's
run of ``gauss_soft_e`` (its probes sampled every 21 steps, the spectra
from the samples every Nyquist interval of 86 steps) are openEMS's files
but for the program line: the time column to the character, the values to
their printed digits, and the frequency-domain files (6 digits) to the
character."""
import os

import numpy as np
import pytest

from ems_amd import openems_probes as OP

import openems_case as OC

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = {
    "openems_gauss_soft_e": ["ut1", "it1", "et1", "ht1",
                             "ut1_FD", "it1_FD", "et1_FD", "ht1_FD"],
    "openems_pml_face": ["ut", "et_out", "ht_out", "et_d2", "ht_d2", "et_d5", "ht_d5"],
    "openems_pml_end": ["ut", "et", "ht"],
}


def text(case, name):
    with open(os.path.join(HERE, "golden", case, name)) as f:
        return f.read()


def test_format_g():
    """probes.rs:560-566 (the doc examples) and C's corner cases."""
    assert OP.format_g(1.92583319e-12, 6) == "1.92583e-12"
    assert OP.format_g(2.5e9, 6) == "2.5e+09"
    assert OP.format_g(0.012, 6) == "0.012"
    assert OP.format_g(0.0, 12) == "0" and OP.format_g(-0.0, 12) == "-0"
    assert OP.format_g(1e-5, 12) == "1e-05" and OP.format_g(123456789012.0, 12) == "123456789012"
    assert OP.format_g(1234567890123.0, 12) == "1.23456789012e+12"
    assert OP.format_g(float("nan"), 6) == "nan" and OP.format_g(-float("inf"), 6) == "-inf"


@pytest.mark.parametrize("case,name", [(c, n) for c, ns in FIXTURES.items() for n in ns])
def test_openems_files_round_trip(case, name):
    t = text(case, name)
    f = OP.ProbeFile.parse(t)
    assert f.generator.startswith("openEMS v0.37")
    assert f.to_text() == t


def oracle_run():
    from oracle.bridge import OracleEngine
    import ems_amd as E
    g, dt = OC.grid(), OC.dt()
    o = OracleEngine(g, dt, OC.lanes(g, dt), curl_order=1, threads=2)
    probes = OC.build(o, g)
    o.run(OC.STEPS)
    series = {}
    for name, ps in probes.items():
        vals = np.stack([o.probe_values(p) for p in ps], axis=1)
        n = vals.shape[0]
        off = 0.5 if name in ("it1", "ht1") else 1.0
        series[name] = (np.array([(s + off) * dt for s in range(n)]), vals)
    o.close()
    return dt, series


KINDS = {"ut1": OP.ProbeFileKind.Voltage, "it1": OP.ProbeFileKind.Current,
         "et1": OP.ProbeFileKind.EField, "ht1": OP.ProbeFileKind.HField}
FREQS = {"ut1": [1e9, 2e9, 2.5e9], "it1": [1e9, 2e9], "et1": [1e9, 2e9], "ht1": [1e9, 2e9]}


@pytest.fixture(scope="module")
def made():
    dt, series = oracle_run()
    out = {}
    for name, kind in KINDS.items():
        ref = OP.ProbeFile.parse(text("openems_gauss_soft_e", name))
        times, vals = series[name]
        out[name] = OP.time_domain(kind, dt, times, vals, OC.INTERVAL, ref.location, ref.generator)
        out[name + "_FD"] = OP.frequency_domain(kind, dt, times, vals, FREQS[name], 86,
                                                ref.location, ref.generator)
    return out


@pytest.mark.parametrize("name", list(KINDS))
def test_time_domain_files(made, name):
    want = text("openems_gauss_soft_e", name).splitlines()
    got = made[name].to_text().splitlines()
    assert got[:4] == want[:4]  # header: program line, corners, columns
    assert len(got) == len(want)
    w = np.array([[float(v) for v in l.split()] for l in want[4:]])
    g = np.array([[float(v) for v in l.split()] for l in got[4:]])
    assert [l.split("\t")[0] for l in got[4:]] == [l.split("\t")[0] for l in want[4:]]
    assert OC.rel_l2(w[:, 1:].ravel(), g[:, 1:].ravel()) < 1e-10


@pytest.mark.parametrize("name", [n + "_FD" for n in KINDS])
def test_frequency_domain_files(made, name):
    assert made[name].to_text() == text("openems_gauss_soft_e", name)


def test_nyquist_interval_of_the_fixture():
    """floor(1/(2 (f0 + fc) dt)) with openEMS's Gaussian f0 = fc = 1.5 GHz:
    86 steps, and the time-domain files every 86 // 4 = 21."""
    dt = OC.dt()
    n = int(np.floor(1.0 / (2.0 * 3e9 * dt)))
    assert n == 86 and n // 4 == OC.INTERVAL
