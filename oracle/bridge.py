"""ctypes bridge to the CPU oracle (liboracle.so in this folder): TEST INFRASTRUCTURE.

The oracle is the checker: tests build an :class:`OracleEngine` from the same
grid, dt, coefficient lanes and operators as the device engine and compare.
This module is test infrastructure; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from ems_amd import engine as E
from ems_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")


def _load():
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    return C.CDLL(ORACLE_LIB)


lib = _load()
_d, _i, _u64, _p = C.c_double, C.c_int, C.c_uint64, C.c_void_p
_fp, _dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
_ip, _u32p, _u8p = C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype, f.argtypes = res, list(args)
    return f


orc_create = _sig("orc_create", _p, _i, _i, _i, _d, _i, _i)
orc_destroy = _sig("orc_destroy", None, _p)
orc_set_lane = _sig("orc_set_lane", _i, _p, _i, _fp, _u64, _u32p)
orc_add_cpml = _sig("orc_add_cpml", _i, _p, _i, _i, _i, _ip, _ip, _fp, _fp, _fp)
orc_add_source = _sig("orc_add_source", _i, _p, _i, _i, _ip, _ip, _fp, _fp, _u64)
orc_add_probe = _sig("orc_add_probe", _i, _p, _i, _i, _ip, _ip, _dp, _u64)
orc_add_lumped = _sig("orc_add_lumped", _i, _p, _p)
orc_lumped_table = _sig("orc_lumped_table", _i, _p, _i, _i, _dp, _u64)
orc_lumped_values = _sig("orc_lumped_values", _u64, _p, _i, _i, _dp, _u64)
orc_lumped_branch = _sig("orc_lumped_branch", _i, _p, _i, _dp, _dp, _dp)
orc_add_dispersive = _sig("orc_add_dispersive", _i, _p, _p, _p, _p)
orc_dispersive_state = _sig("orc_dispersive_state", _i, _p, _i, _i, _dp, _dp)
orc_add_planewave = _sig("orc_add_planewave", _i, _p, _p)
orc_planewave_table = _sig("orc_planewave_table", _i, _p, _i, _dp, _u64)
orc_planewave_aux = _sig("orc_planewave_aux", _i, _p, _i, _dp, _dp)
orc_add_merged = _sig("orc_add_merged", _i, _p, _p)
orc_merged_state = _sig("orc_merged_state", _i, _p, _i, _dp, _dp)
orc_run = _sig("orc_run", None, _p, _u64)
orc_set_fused = _sig("orc_set_fused", None, _p, _i)
_ip2 = C.POINTER(C.c_int)
orc_add_mur = _sig("orc_add_mur", _i, _p, _i, _ip2, _ip2, _ip2, C.POINTER(C.c_float),
                   C.POINTER(C.c_uint64), _i, _ip2, _ip2, _ip2, _ip2, C.POINTER(C.c_float),
                   C.POINTER(C.c_float), C.POINTER(C.c_uint64), _i, _i)
orc_add_upml = _sig("orc_add_upml", _i, _p, _i, _ip2, _ip2, C.POINTER(C.c_float), _i, _ip2, _ip2,
                    C.POINTER(C.c_float))
orc_upml_state = _sig("orc_upml_state", _i, _p, _i, _i, C.POINTER(C.c_float), C.POINTER(C.c_float))
orc_step_with_energy = _sig("orc_step_with_energy", _d, _p)
orc_energy = _sig("orc_energy", _d, _p)
orc_energy_openems = _sig("orc_energy_openems", _d, _p)
orc_energy_openems_ex = _sig("orc_energy_openems_ex", _d, _p, _i)
orc_timestep = _sig("orc_timestep", _u64, _p)
orc_read = _sig("orc_read", _i, _p, _i, _i, _fp)
orc_write = _sig("orc_write", _i, _p, _i, _i, _fp)
orc_probe_count = _sig("orc_probe_count", _u64, _p, _i)
orc_probe_values = _sig("orc_probe_values", _u64, _p, _i, _dp, _u64)
orc_courant_limit = _sig("orc_courant_limit", _d, _dp, _i, _dp, _i, _dp, _i)
orc_coefficients = _sig("orc_coefficients", _i, _dp, _i, _dp, _i, _dp, _i, _dp, _dp, _dp, _dp, _u8p,
                        _ip, _d, _fp)
orc_cpml_coefficients = _sig("orc_cpml_coefficients", None, _d, _d, _d, _d, _i, _d, _d, _d, _d, _dp)
orc_drive_script = _sig("orc_drive_script", _i, _dp, _u64, _u64, _i, _d, _u64, _i, _d,
                        C.POINTER(_u64))


def P(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class OracleEngine:
    """The CPU restatement, fed the same arrays as :class:`ems_amd.Engine`."""

    def __init__(self, grid: E.Grid, dt: float, lanes: np.ndarray, curl_order: int = 0,
                 threads: int = 1, pool: bool = True, fused: int = 0):
        self.grid, self.dt = grid, float(dt)
        self._dense_lanes = lanes
        n = grid.nodes()
        self._h = orc_create(n[0], n[1], n[2], self.dt, curl_order, threads)
        assert self._h, "orc_create failed"
        # FusedSteps: 0 Auto (the reference's default), 1 Always, 2 Never
        orc_set_fused(self._h, fused)
        for l in range(12):
            lane = np.ascontiguousarray(lanes[l])
            pooled = E.pool_lane(lane) if pool else None
            if pooled is None:
                rows = lane.reshape(-1, n[2])
                assert orc_set_lane(self._h, l, P(rows, C.c_float), rows.shape[0], None) == 0
            else:
                rows, index = pooled
                assert orc_set_lane(self._h, l, P(rows, C.c_float), rows.shape[0],
                                    P(index, C.c_uint32)) == 0
        self.ops = []
        self._tables = []
        self._lumped = {}
        self._poles_n = {}
        self._pw_tables = []

    def coefficients(self) -> np.ndarray:
        return self._dense_lanes

    def _set_lane(self, l, lane):
        n = self.grid.nodes()
        lane = np.ascontiguousarray(lane)
        pooled = E.pool_lane(lane)
        if pooled is None:
            rows = lane.reshape(-1, n[2])
            assert orc_set_lane(self._h, l, P(rows, C.c_float), rows.shape[0], None) == 0
        else:
            rows, index = pooled
            assert orc_set_lane(self._h, l, P(rows, C.c_float), rows.shape[0],
                                P(index, C.c_uint32)) == 0

    def patch_lanes(self, changed):
        lanes = self._dense_lanes.copy()
        for l, lane in sorted(changed.items()):
            lanes[l] = lane
            self._set_lane(l, lanes[l])
        self._dense_lanes = lanes

    def _ensure_tables(self, upto):
        for oid, t in self._pw_tables:
            def up_pw(vals, oid=oid):
                v = np.ascontiguousarray(vals)
                assert orc_planewave_table(self._h, oid, P(v, C.c_double), v.size) == 0
            t.ensure(upto, up_pw)
        for oid, t in self._tables:
            def up(term, vals, oid=oid):
                v = np.ascontiguousarray(vals)
                assert orc_lumped_table(self._h, oid, term, P(v, C.c_double), v.size) == 0
            t.ensure(upto, up)

    def _lumped_values(self, oid, which):
        n = orc_lumped_values(self._h, oid, which, None, 0)
        out = np.empty(n, dtype=np.float64)
        orc_lumped_values(self._h, oid, which, P(out, C.c_double), n)
        return out, self._lumped[oid][1]

    def _planewave_aux(self, oid, naux):
        ev, hv = np.zeros(naux), np.zeros(naux - 1)
        assert orc_planewave_aux(self._h, oid, P(ev, C.c_double), P(hv, C.c_double)) == 0
        return ev, hv

    def _merged_state(self, oid, ng):
        cur, circ = np.zeros(ng), np.zeros(ng)
        assert orc_merged_state(self._h, oid, P(cur, C.c_double), P(circ, C.c_double)) == 0
        return cur, circ

    def _dispersive_state(self, oid, field):
        n = self._poles_n[oid][field]
        j, q = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        assert orc_dispersive_state(self._h, oid, field, P(j, C.c_double), P(q, C.c_double)) == 0
        return j[:n], q[:n]

    def _lumped_branch(self, oid):
        n = self._lumped[oid][0]
        j, q, jav = (np.zeros(n) for _ in range(3))
        assert orc_lumped_branch(self._h, oid, P(j, C.c_double), P(q, C.c_double),
                                 P(jav, C.c_double)) == 0
        return j, q, jav

    def close(self):
        if self._h:
            orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_operator(self, op):
        h = self._h
        if type(op).__name__ == "MurBoundary":
            a = op.arrays()
            I = C.c_int
            r = orc_add_mur(h, len(op.edges), P(a["eax"], I), P(a["enode"], I), P(a["einner"], I),
                            P(a["ek"], C.c_float), P(a["ehold"], C.c_uint64), len(op.corners),
                            P(a["cax"], I), P(a["cnode"], I), P(a["cin0"], I), P(a["cin1"], I),
                            P(a["ck0"], C.c_float), P(a["ck1"], C.c_float),
                            P(a["chold"], C.c_uint64), 1 if op._openems_order else 0,
                            1 if op._openems_corners else 0)
            assert r >= 0, "orc_add_mur failed"
            self.ops.append(op)
            return op
        if type(op).__name__ == "OpenEmsPml":
            (ea, en, ec), (ha, hn, hc) = op.e, op.h
            I, F = C.c_int, C.c_float
            r = orc_add_upml(h, ea.size, P(ea, I), P(en, I), P(ec, F), ha.size, P(ha, I), P(hn, I),
                             P(hc, F))
            assert r >= 0, "orc_add_upml refused the layers"
            op._oid = getattr(op, "_oid", {})
            op._oid[id(self)] = r
            self.ops.append(op)
            return op
        if isinstance(op, E.CpmlBoundary):
            for s in op.slabs:
                lo = (C.c_int * 3)(*s["lo"])
                hi = (C.c_int * 3)(*s["hi"])
                b, a, k = (np.ascontiguousarray(s[x]) for x in ("b", "a", "k"))
                assert orc_add_cpml(h, s["e_side"], s["comp"], s["normal"], lo, hi,
                                    P(b, C.c_float), P(a, C.c_float), P(k, C.c_float)) >= 0
        elif isinstance(op, E.EdgeSource):
            amp = np.ascontiguousarray(op._table(self.dt))
            axis, nodes, w = E.Engine._edges(op.edges, np.float32)
            axis = axis.astype(np.intc)
            nodes = nodes.astype(np.intc)
            op._oid = orc_add_source(h, int(op._kind), len(op.edges), P(axis, C.c_int),
                                     P(nodes, C.c_int), P(w, C.c_float), P(amp, C.c_float),
                                     amp.size)
            assert op._oid >= 0
        elif isinstance(op, E.Probe):
            axis, nodes, w = E.Engine._edges(op.entries, np.float64)
            axis = axis.astype(np.intc)
            nodes = nodes.astype(np.intc)
            op._oid = orc_add_probe(h, op._c_quantity(), len(op.entries), P(axis, C.c_int),
                                    P(nodes, C.c_int), P(w, C.c_double), op._interval)
            assert op._oid >= 0
        elif type(op).__name__ == "PlaneWave":
            from ems_amd import planewave as PW
            d, keep = PW.make_desc(op)
            oid = orc_add_planewave(h, C.addressof(d))
            assert oid >= 0
            self._pw_tables.append((oid, PW.SourceTable(op)))
            op._bind(self, oid)
        elif type(op).__name__ == "MergedFaces":
            from ems_amd import merged as MF
            d, keep = MF.make_desc(op)
            oid = orc_add_merged(h, C.addressof(d))
            assert oid >= 0
            op._bind(self, oid)
        elif type(op).__name__ == "Dispersive":
            from ems_amd import dispersive as DP
            plan = op._plan(self.grid, self.dt, self._dense_lanes)
            se, sh, cs, keep = DP.make_desc(plan)
            oid = orc_add_dispersive(h, C.addressof(se), C.addressof(sh), C.addressof(cs))
            assert oid >= 0
            self._poles_n[oid] = (len(plan["sides"][0]["cidx"]), len(plan["sides"][1]["cidx"]))
            op._bind(self, oid, plan)
        elif hasattr(op, "_plan"):
            from ems_amd import lumped as LM
            plan = op._plan(self.grid, self.dt, self._dense_lanes)
            a = plan["axis"]
            self._dense_lanes = self._dense_lanes.copy()
            for l, lane in ((a, plan["ca"]), (3 + a, plan["cb"])):
                self._dense_lanes[l] = lane
                self._set_lane(l, lane)
            d, keep = LM.make_desc(plan)
            oid = orc_add_lumped(h, C.addressof(d))
            assert oid >= 0
            self._lumped[oid] = (len(plan["edges"]), self.timestep())
            self._tables.append((oid, LM.SourceTables(plan, self.dt)))
            op._bind(self, oid, plan)
        else:
            raise TypeError(type(op))
        self.ops.append(op)
        return op

    def run(self, n):
        self._ensure_tables(self.timestep() + int(n))
        orc_run(self._h, int(n))

    def step_with_energy(self):
        self._ensure_tables(self.timestep() + 1)
        return orc_step_with_energy(self._h)

    def energy_openems(self, layers=True):
        return orc_energy_openems_ex(self._h, 1 if layers else 0)

    def energy(self):
        return orc_energy(self._h)

    def timestep(self):
        return orc_timestep(self._h)

    def _read(self, field, axis):
        out = np.empty(self.grid.nodes(), dtype=np.float32)
        assert orc_read(self._h, field, int(axis), P(out, C.c_float)) == 0
        return out

    def u_component(self, axis):
        return self._read(0, axis)

    def i_component(self, axis):
        return self._read(1, axis)

    def set_component(self, field, axis, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(self.grid.nodes()))
        assert orc_write(self._h, field, int(axis), P(v, C.c_float)) == 0

    def upml_state(self, op, side):
        n = op.edge_counts()[side]
        f, v = np.empty(n, np.float32), np.empty(n, np.float32)
        r = orc_upml_state(self._h, op._oid[id(self)], int(side), P(f, C.c_float),
                           P(v, C.c_float))
        assert r == 0
        return f, v

    def probe_values(self, op):
        n = orc_probe_count(self._h, op._oid)
        out = np.empty(n, dtype=np.float64)
        orc_probe_values(self._h, op._oid, P(out, C.c_double), n)
        return out


def oracle_coefficients(grid: E.Grid, eps_r, kappa, mu_r, sigma_m, pec, pec_face, dt):
    """Coefficients::from_edges restated in the oracle (per-edge inputs)."""
    n = grid.nodes()
    N = n[0] * n[1] * n[2]
    out = np.empty((12,) + n, dtype=np.float32)
    arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(3 * N))
            for a in (eps_r, kappa, mu_r, sigma_m)]
    pm = None if pec is None else np.ascontiguousarray(np.asarray(pec, dtype=np.uint8).reshape(3 * N))
    pf = (C.c_int * 6)(*pec_face)
    x, y, z = grid.lines
    rc = orc_coefficients(P(x, C.c_double), x.size, P(y, C.c_double), y.size, P(z, C.c_double),
                          z.size, *[P(a, C.c_double) for a in arrs], P(pm, C.c_uint8), pf, dt,
                          P(out, C.c_float))
    assert rc == 0
    return out


def oracle_cpml_coefficients(p: E.Cpml, rho, delta, dt, n=1.0):
    out = (C.c_double * 7)()
    orc_cpml_coefficients(p.grading_order, p.sigma_factor, p.kappa_max, p.alpha_max,
                          p.convolution, rho, delta, dt, n, out)
    return dict(zip(("rho", "sigma", "kappa", "alpha", "b", "a", "direct"), list(out)))


def oracle_drive_script(w, max_steps, energy_ratio=None, interval=50, excitation_end=None):
    wa = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    steps = C.c_uint64()
    r = orc_drive_script(P(wa, C.c_double), wa.size, max_steps, energy_ratio is not None,
                         energy_ratio or 0.0, interval, excitation_end is not None,
                         excitation_end or 0.0, C.byref(steps))
    return r, steps.value
