"""openEMS probe files (src/io/openems/probes.rs): the text format openEMS
writes for its ProbeBox results, read and written, and the time- and
frequency-domain files made from a probe's samples.

A file is a header of '%' lines (the domain, quantity and program; the
snapped start and stop corners; the column names) and one row per sample:
time-domain files print every value with C's ``%.12g``, frequency-domain
files ``%.6g`` with a real and an imaginary column per component.
Time-domain rows are the probe's samples every `interval` steps (E-type
quantities at t = k dt with the initial zero state as row 0, H-type ones at
(k + 1/2) dt); frequency-domain rows are, per FD_Samples frequency f,
sum over the samples every `nyquist` steps of 2 N dt exp(-2 pi i f t) x.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field
from enum import Enum
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat
from ._native import EmsError

VERSION = "0.2"


def format_g(v: float, prec: int) -> str:
    """C printf("%.{prec}g") (probes.rs:554-599): prec significant digits,
    trailing zeros removed, exponent form (sign, two digits) below 1e-4 or
    at 10^prec and above; "-0", "nan", "inf" as C prints them."""
    v = float(v)
    if math.isnan(v):
        return "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return "%.*g" % (max(int(prec), 1), v)


class ProbeDomain(Enum):
    Time = "time-domain"
    Frequency = "frequency-domain"


class ProbeFileKind(Enum):
    """ProbeFileKind (probes.rs:95-175): title, columns, H-type."""
    Voltage = ("voltage integration", ("voltage",), False)
    Current = ("current integration", ("current",), True)
    EField = ("electric field probe", ("Ex/(V/m)", "Ey/(V/m)", "Ez/(V/m)"), False)
    HField = ("magnetic field probe", ("Hx/(A/m)", "Hy/(A/m)", "Hz/(A/m)"), True)
    VoltageModeMatching = ("voltage mode matching", ("voltage", "mode_purity"), False)
    CurrentModeMatching = ("current mode matching", ("current", "mode_purity"), True)

    @property
    def title(self) -> str:
        return self.value[0]

    @property
    def columns(self) -> Tuple[str, ...]:
        return self.value[1]

    @property
    def components(self) -> int:
        return len(self.value[1])

    @property
    def is_magnetic(self) -> bool:
        return self.value[2]


@dataclass
class ProbeLocation:
    """The snapped corners: coordinates [m] and node indices."""
    start_coords: Tuple[float, float, float]
    start: Tuple[int, int, int]
    stop_coords: Tuple[float, float, float]
    stop: Tuple[int, int, int]


def asctime_utc(t: Optional[float] = None) -> str:
    """C asctime of `t` (seconds since the epoch, now if None) in UTC."""
    return time.strftime("%a %b %e %H:%M:%S %Y", time.gmtime(time.time() if t is None else t))


def generator() -> str:
    return f"ems_amd v{VERSION} @{asctime_utc()}"


@dataclass
class ProbeFile:
    """ProbeFile (probes.rs:178-330): domain, quantity, corners, the
    program line and the numeric rows (t or f first)."""
    domain: ProbeDomain
    kind: ProbeFileKind
    location: Optional[ProbeLocation]
    generator: str
    rows: List[List[float]] = field(default_factory=list)

    def to_text(self) -> str:
        """ProbeFile::to_text (probes.rs:340-397)."""
        out = [f"% {self.domain.value} {self.kind.title} by {self.generator}\n"]
        if self.location is not None:
            l = self.location
            for what, c, i in (("start", l.start_coords, l.start), ("stop", l.stop_coords, l.stop)):
                out.append("%% %s-coordinates: (%s,%s,%s) m -> [%d,%d,%d]\n"
                           % (what, format_g(c[0], 6), format_g(c[1], 6), format_g(c[2], 6),
                              i[0], i[1], i[2]))
        freq = self.domain == ProbeDomain.Frequency
        prec, absc = (6, "f/Hz") if freq else (12, "t/s")
        line = "% " + absc
        for c in self.kind.columns:
            line += "\t" + c + ("\t" if freq else "")
        out.append(line + "\n")
        if freq:
            out.append("%" + "\treal\timag" * self.kind.components + "\n")
        for r in self.rows:
            out.append("\t".join(format_g(v, prec) for v in r) + "\n")
        return "".join(out)

    @staticmethod
    def parse(text: str) -> "ProbeFile":
        """ProbeFile::parse (probes.rs:206-286)."""
        lines = text.splitlines()
        if not lines or not lines[0].startswith("% "):
            raise EmsError(nat.EMS_E_INVALID, "not a probe file")
        first = lines[0][2:]
        dom = next((d for d in ProbeDomain if first.startswith(d.value + " ")), None)
        if dom is None:
            raise EmsError(nat.EMS_E_INVALID, f"unknown probe domain in {first!r}")
        rest = first[len(dom.value) + 1:]
        kind = next((k for k in ProbeFileKind if rest.startswith(k.title + " by ")), None)
        if kind is None:
            raise EmsError(nat.EMS_E_INVALID, f"unknown probe kind in {first!r}")
        gen = rest[len(kind.title) + len(" by "):]
        start = stop = None
        rows = []
        width = 1 + kind.components * (2 if dom == ProbeDomain.Frequency else 1)
        for no, line in enumerate(lines[1:]):
            if line.startswith("%"):
                body = line[1:].strip()
                for what in ("start", "stop"):
                    if body.startswith(what + "-coordinates:"):
                        c, i = body.split(":", 1)[1].split("->")
                        cs = tuple(float(x) for x in c.strip().strip("m").strip()[1:-1].split(","))
                        ix = tuple(int(x) for x in i.strip()[1:-1].split(","))
                        if what == "start":
                            start = (cs, ix)
                        else:
                            stop = (cs, ix)
                continue
            if not line.strip():
                continue
            vals = [float(v) for v in line.split()]
            if len(vals) != width:
                raise EmsError(nat.EMS_E_INVALID,
                               f"probe file line {no + 2}: {len(vals)} columns, expected {width}")
            rows.append(vals)
        loc = None
        if start is not None and stop is not None:
            loc = ProbeLocation(start[0], start[1], stop[0], stop[1])
        return ProbeFile(dom, kind, loc, gen, rows)


def samples(kind: ProbeFileKind, dt: float, times: Sequence[float],
            values: np.ndarray) -> List[Tuple[int, float, List[float]]]:
    """samples (probes.rs:470-495): (step k, time, values) of every sample,
    k = round(t/dt - offset) with offset 1/2 for H-type quantities, the
    E-type initial state (k = 0, zero) added when the recording starts at
    k = 1. `values`: one row per sample, one column per component."""
    off = 0.5 if kind.is_magnetic else 0.0
    vals = np.asarray(values, dtype=np.float64).reshape(len(times), -1)

    def step(t):
        return max(int(round(t / dt - off)), 0)
    out = []
    if not kind.is_magnetic and len(times) and step(times[0]) == 1:
        out.append((0, 0.0, [0.0] * kind.components))
    for t, v in zip(times, vals):
        out.append((step(t), float(t), [float(x) for x in v]))
    return out


def time_domain(kind: ProbeFileKind, dt: float, times, values, interval: int,
                location: Optional[ProbeLocation] = None, gen: Optional[str] = None) -> ProbeFile:
    """ProbeFile::time_domain (probes.rs:399-423): the samples every
    `interval` steps."""
    iv = max(int(interval), 1)
    rows = [[t] + v for k, t, v in samples(kind, dt, times, values) if k % iv == 0]
    return ProbeFile(ProbeDomain.Time, kind, location, gen or generator(), rows)


def frequency_domain(kind: ProbeFileKind, dt: float, times, values, freqs: Sequence[float],
                     nyquist: int, location: Optional[ProbeLocation] = None,
                     gen: Optional[str] = None) -> Optional[ProbeFile]:
    """ProbeFile::frequency_domain (probes.rs:425-468): per frequency the
    sum of 2 N dt exp(-2 pi i f t) x over the samples every N = `nyquist`
    steps; None without frequencies."""
    if not len(freqs):
        return None
    n = max(int(nyquist), 1)
    picked = [(t, v) for k, t, v in samples(kind, dt, times, values) if k % n == 0]
    w = 2.0 * n * dt
    rows = []
    for f in freqs:
        acc = [0j] * kind.components
        for t, v in picked:
            ph = -2.0 * math.pi * f * t
            e = complex(w * math.cos(ph), w * math.sin(ph))
            acc = [a + e * x for a, x in zip(acc, v)]
        r = [float(f)]
        for a in acc:
            r += [a.real, a.imag]
        rows.append(r)
    return ProbeFile(ProbeDomain.Frequency, kind, location, gen or generator(), rows)


def write_probe(directory: str, name: str, kind: ProbeFileKind, dt: float, times, values,
                interval: int, freqs: Sequence[float] = (), nyquist: int = 1,
                location: Optional[ProbeLocation] = None) -> List[str]:
    """write_probe (probes.rs:620-661): `<name>` and, with frequencies,
    `<name>_FD` in `directory`; returns the paths written."""
    os.makedirs(directory, exist_ok=True)
    gen = generator()
    paths = []
    td = time_domain(kind, dt, times, values, interval, location, gen)
    p = os.path.join(directory, name)
    with open(p, "w") as f:
        f.write(td.to_text())
    paths.append(p)
    fd = frequency_domain(kind, dt, times, values, freqs, nyquist, location, gen)
    if fd is not None:
        p = os.path.join(directory, name + "_FD")
        with open(p, "w") as f:
            f.write(fd.to_text())
        paths.append(p)
    return paths


__all__ = ["ProbeDomain", "ProbeFile", "ProbeFileKind", "ProbeLocation", "format_g",
           "frequency_domain", "samples", "time_domain", "write_probe"]
