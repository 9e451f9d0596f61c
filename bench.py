#!/usr/bin/env python3
"""Throughput of the MI355X FDTD step engine on one GPU.

Default workload: a 512^3-cell graded CPML box (8-cell CPML on all faces, 64
seeded FR-4 blocks, soft Gaussian Ez source, 48 voltage probes), 5000 steps.
One "step" is one full FDTD time step (H and E updates, CPML, source,
probes) over the whole grid; value = cells x steps / wall time (Gcell/s),
with all inputs resident in HBM before the timed region.

Prints one JSON line with a roofline figure for the dominant kernel (HIP
events on the engine's stream) and a CPU baseline: the C oracle (a port of
the reference engine, x-slab threads) timed on a bounded sample of the same
workload. The default run also carries, under "dispersive_box", the line of
the 384^3-cell dispersive / conformal / plane-wave box (unfused schedule,
8000 steps) and, under "patch_sweep", the batched 256-model patch-antenna
sweep's S11 solves/hr, each run in a child process after the CPML box's
engine is closed (--no-dispersive skips both).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def physical_cores() -> int:
    avail = len(os.sched_getaffinity(0))
    sib = set()
    cpu = 0
    while True:
        p = f"/sys/devices/system/cpu/cpu{cpu}/topology/thread_siblings_list"
        if not os.path.exists(p):
            break
        with open(p) as f:
            sib.add(f.read().strip())
        cpu += 1
    smt = max(cpu // max(len(sib), 1), 1) if sib else 1
    return max(avail // smt, 1)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, seconds: float, log) -> dict:
    """The oracle (C port of the reference engine, with its fused per-plane
    schedule, src/engine/pool.rs:405-594, under FusedSteps::Auto) on the
    same compiled arrays. Threads by the reference's rule
    (default_threads_for, src/engine/mod.rs:404-454: one per physical core,
    at most one per 4096 cells and per 5 x-planes), within the 16 host
    threads one GPU's share of the box allows. Best of three samples."""
    from oracle.bridge import OracleEngine
    nodes = wl.grid.nodes()
    cells = 1
    for n in nodes:
        cells *= max(n - 1, 1)
    threads = max(1, min(physical_cores(), cells // 4096, nodes[0] // 5))
    share = min(threads, 16)
    o = OracleEngine(wl.grid, wl.dt, wl.lanes, threads=share)
    wl.build_ops(o)
    o.run(1)
    t0 = time.perf_counter()
    o.run(1)
    one = max(time.perf_counter() - t0, 1e-6)
    n = max(1, min(int(seconds / 3.0 / one), 100000))
    log(f"cpu baseline: {share} threads (reference rule: {threads}), {one*1e3:.1f} ms/step, "
        f"best of 3 x {n} steps")
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        o.run(n)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    o.close()
    return {"value": wl.cells * n / best / 1e9, "unit": "Gcell/s", "cores": share,
            "kind": "port",
            "sample": f"best of 3 runs of {n} steps of the same workload after 2 warm-up steps "
                      f"({best:.1f} s each; fused per-plane schedule; {share} threads, the "
                      f"reference's rule gives {threads} on this host; host CPU: {cpu_model()})"}


def sweep(args, log) -> int:
    """The patch-antenna sweep: every model solved to -40 dB energy decay
    and analysed (S11 at 201 points over 1-3 GHz, the port spectra from
    running DFTs on the device), first one after the other
    on the GPU (each its own engine), then all together in one batch.
    value = batched S11 solves per hour; both include set-up and analysis."""
    import ems_amd as E
    from ems_amd import learn as L
    from ems_amd import workloads as W
    freqs = [1e9 + 2e9 * q / 200.0 for q in range(201)]
    lim = E.RunLimits(args.max_steps, E.EndCriteria.energy_decay_db(-40.0),
                      excitation_end=E.Waveform.gaussian(2e9, 1e9).duration())
    t0 = time.perf_counter()
    models = W.patch_sweep(args.models, seed=args.seed)
    log(f"{len(models)} patch models ({sum(m.cells for m in models)/len(models):.0f} cells on "
        f"average) in {time.perf_counter()-t0:.1f} s")

    phase = {"setup": 0.0, "run": 0.0, "analysis": 0.0}

    packed = args.kernel == "packed" and all(m.grid.nodes()[2] <= 64 for m in models)
    tj = args.tile_rows or 7

    def solve_one(m):
        ta = time.perf_counter()
        e = E.Engine(m.grid, m.dt, m.lanes)
        e.set_schedule(5 if packed else 2, tj, 0)  # the batch's kernels, one model
        p = m.build_ops(e)
        e.release_dense_lanes()
        tb = time.perf_counter()
        o = e.run_until(lim)
        tc = time.perf_counter()
        s11 = p.analyze_device(freqs).s11()
        e.close()
        phase["setup"] += tb - ta
        phase["run"] += tc - tb
        phase["analysis"] += time.perf_counter() - tc
        return o, s11

    # warm-up: first engine, first graphs
    solve_one(models[0])
    for k in phase:
        phase[k] = 0.0
    seq, t_seq, seq_phase = None, None, None
    if not args.skip_sequential:
        t0 = time.perf_counter()
        seq = [solve_one(m) for m in models]
        t_seq = time.perf_counter() - t0
        cell_steps = sum(m.cells * o.steps for m, (o, _) in zip(models, seq))
        log(f"sequential: {t_seq:.2f} s, {cell_steps / t_seq / 1e9:.2f} Gcell/s "
            f"(setup {phase['setup']:.2f} s, run {phase['run']:.2f} s, "
            f"analysis {phase['analysis']:.2f} s)")
        seq_phase = dict(phase)

    t0 = time.perf_counter()
    engines, ports = L._batch_engines(models)
    b = E.EngineBatch(engines)
    b.set_schedule(tj, args.x_chunks, packed=packed)
    if args.no_graphs:
        b.set_graphs(False)
        log(f"device copy bandwidth: {engines[0].copy_bandwidth(1 << 30, 5):.0f} GB/s "
            "(calibration for counter passes)")
    t1 = time.perf_counter()
    outs = b.run_until(lim)
    t2 = time.perf_counter()
    # every port's V and I spectra from the device's running DFTs (the
    # series stay on the device)
    for p in ports:
        p.update_device(freqs)
    s11b = [p.analyze_device(freqs).s11() for p in ports]
    t_bat = time.perf_counter() - t0
    bat_phase = {"setup": t1 - t0, "run": t2 - t1, "analysis": time.perf_counter() - t2}
    cell_steps = sum(m.cells * o.steps for m, o in zip(models, outs))
    log(f"batched: {t_bat:.2f} s, {cell_steps / t_bat / 1e9:.2f} Gcell/s "
        f"(setup {bat_phase['setup']:.2f} s, run {bat_phase['run']:.2f} s, "
        f"analysis {bat_phase['analysis']:.2f} s)")
    # The batched kernel's roofline: the models that stopped on the more
    # common buffer parity (a batch steps in lockstep) stepped 64 more
    # times in one batch, the time per step from the synchronised runs'
    # wall clock (graph replays of the fused and gap kernels), algorithmic
    # bytes the models' traffic models summed, HBM traffic from the counter
    # passes of a run without graphs (a kernel trace of the batch's graph
    # launches crashed inside the runtime).
    par = [e.timestep() % 2 for e in engines]
    keep = 1 if 2 * sum(par) >= len(par) else 0
    act = [p_ == keep for p_ in par]
    b.run(2, act)  # (the graphs of this set captured)
    tk = time.perf_counter()
    b.run(64, act)
    per_step = (time.perf_counter() - tk) / 64
    byt = sum(e.traffic_model()[0] for e, a in zip(engines, act) if a)
    achieved = byt / per_step / 1e9
    kname = "step_packed_batch" if packed else "step_fused_batch"
    traffic, traffic_cfg = None, None
    tpath = os.path.join(ROOT, "profiles", f"{kname}_patch_sweep_{len(models)}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        traffic, traffic_cfg = tr["traffic_bytes"], tr["config"]
    roof = {"bound": "hbm", "kernel": kname,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_measured_at": traffic_cfg,
            "bytes_per_launch": byt, "launch_ms": round(per_step * 1e3, 4),
            "models": int(sum(act)),
            "timing": "wall clock of synchronised batch runs (64 steps, fused and gap kernels); "
                      "counters from a run without graphs"}
    log(f"batched kernel: {per_step*1e3:.3f} ms per step over {sum(act)} models, "
        f"{achieved:.0f} GB/s algorithmic")
    d, same_steps = None, None
    if seq is not None:
        d = 0.0
        same_steps = all(o.steps == so.steps for o, (so, _) in zip(outs, seq))
        for a, (_, bb) in zip(s11b, seq):
            d = max(d, max(abs(x - y) for x, y in zip(a, bb)))
    b.close()
    for e in engines:
        e.close()
    n = len(models)
    line = {
        "metric": "S11 solves/hr", "value": round(n / t_bat * 3600.0, 1), "unit": "solves/hr",
        "n_gpus": 1, "steps": int(sum(o.steps for o in outs)), "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded patch antennas)",
        "config": {"workload": f"patch_sweep_{n}", "models": n, "seed": args.seed,
                   "frequencies": 201, "end": "-40 dB energy decay",
                   "schedule": {"kernel": "packed rows" if packed else f"tiled, {tj} rows",
                                "x_chunks": args.x_chunks}},
        "sequential_solves_per_hr": round(n / t_seq * 3600.0, 1) if t_seq else None,
        "speedup_vs_sequential": round(t_seq / t_bat, 3) if t_seq else None,
        "max_abs_dS11": d, "same_step_counts": same_steps,
        "phases_s": {"sequential": ({k: round(v, 3) for k, v in seq_phase.items()}
                                    if seq_phase else None),
                     "batched": {k: round(v, 3) for k, v in bat_phase.items()}},
        "cell_updates": cell_steps, "batched_s": round(t_bat, 3),
        "sequential_s": round(t_seq, 3) if t_seq else None,
        "roofline": roof,
    }
    print(json.dumps(line), flush=True)
    return 0


def earlystop(args, log) -> int:
    """The learned early stop on a fresh patch sweep: every model solved in
    full to -40 dB energy decay (one batch), then early-stopped (one batch
    of `observed` steps, the trained network predicting the rest of each
    port trace, full runs for the models it flags); S11 at 201 points over
    1-3 GHz compared with the full runs, and with the observed samples
    alone (the method without the network). value = early-stopped S11
    solves per hour, set-up and analysis included."""
    import numpy as np
    import ems_amd as E
    from ems_amd import learn as L
    from ems_amd import workloads as W
    path = args.model or L.PATCH_MODEL
    es = L.EarlyStop.load(path)
    freqs = L.FREQS_1_3_GHZ
    lim = E.RunLimits(args.max_steps, E.EndCriteria.energy_decay_db(-40.0),
                      excitation_end=E.Waveform.gaussian(2e9, 1e9).duration())
    models = W.patch_sweep(args.models, seed=args.seed)
    log(f"{len(models)} patch models, network {path} (horizon {es.fmt.horizon}, observed "
        f"{es.fmt.observed}, decimation {es.fmt.decimation})")
    # warm-up: one small batch through both paths
    L.early_stop_sweep(es, models[:4], lim)

    t0 = time.perf_counter()
    engines, ports = L._batch_engines(models)
    b = E.EngineBatch(engines)
    L._schedule(b, models, args.tile_rows or 7, args.x_chunks)
    outs = b.run_until(lim)
    for p in ports:
        p.update_device(freqs)
    full = [p.analyze_device(freqs).s11() for p in ports]
    b.close()
    for e in engines:
        e.close()
    t_full = time.perf_counter() - t0
    full_steps = [o.steps for o in outs]
    log(f"full runs: {t_full:.2f} s, {np.mean(full_steps):.0f} steps on average")

    t0 = time.perf_counter()
    res = L.early_stop_sweep(es, models, lim, freqs, tile_rows=args.tile_rows or 7)
    t_es = time.perf_counter() - t0
    log(f"early-stopped: {t_es:.2f} s, {sum(res.fallback)} fallbacks")
    err = [max(abs(x - y) for x, y in zip(a, f)) for a, f in zip(res.s11, full)]
    tr = L.early_stop_sweep(es, models, lim, freqs, tile_rows=args.tile_rows or 7, truncated=True)
    terr = [max(abs(x - y) for x, y in zip(a, f)) for a, f in zip(tr.s11, full)]
    n = len(models)
    line = {
        "metric": "S11 solves/hr (learned early stop)", "value": round(n / t_es * 3600.0, 1),
        "unit": "solves/hr", "n_gpus": 1, "steps": int(sum(res.steps)), "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (network: bf16 weights, f32 accumulate)",
        "data": "synthetic (seeded patch antennas)",
        "config": {"workload": f"patch_earlystop_{n}", "models": n, "seed": args.seed,
                   "frequencies": 201, "observed_steps": es.fmt.observed,
                   "horizon": es.fmt.horizon, "decimation": es.fmt.decimation,
                   "fallback": f"predicted port power over the last 16 samples > "
                               f"{es.threshold_db:.0f} dB of its peak"},
        "full_solves_per_hr": round(n / t_full * 3600.0, 1),
        "speedup_vs_full_batched": round(t_full / t_es, 3),
        "mean_full_steps": float(np.mean(full_steps)),
        "fallbacks": int(sum(res.fallback)),
        "max_abs_dS11": float(np.max(err)), "median_abs_dS11": float(np.median(err)),
        "p90_abs_dS11": float(np.percentile(err, 90)),
        "models_within_0.01": int(sum(e <= 0.01 for e in err)),
        "truncated_max_abs_dS11": float(np.max(terr)),
        "truncated_median_abs_dS11": float(np.median(terr)),
        "full_s": round(t_full, 3), "early_s": round(t_es, 3),
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cpml_box",
                    choices=["cpml_box", "vacuum_box", "patch_sweep", "patch_earlystop",
                             "dispersive_box"])
    ap.add_argument("--models", type=int, default=256, help="patch_sweep: models")
    ap.add_argument("--seed", type=int, default=4242, help="patch_sweep: seed")
    ap.add_argument("--max-steps", type=int, default=40000, help="patch_sweep: step limit")
    ap.add_argument("--model", default="", help="patch_earlystop: trained network (npz)")
    ap.add_argument("--skip-sequential", action="store_true",
                    help="patch_sweep: the batched leg only (for profiling)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="patch_sweep: the batch queues its launches one by one instead of "
                         "replaying HIP graphs (same results; for launch-tracing profilers), "
                         "and a copy-kernel calibration runs first")
    ap.add_argument("--kernel", default="packed", choices=["packed", "tiled"],
                    help="patch workloads: fused kernel (packed short rows where nz <= 64)")
    ap.add_argument("--n", type=int, default=None, help="cells per axis")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-steps", type=int, default=20)
    ap.add_argument("--unfused", action="store_true", help="unfused schedule (H, CPML, E launches)")
    ap.add_argument("--mode", type=int, default=1,
                    help="fused kernel with --tile-rows: 1 tiled, 2 tiled without prefetch, 3 full rows")
    ap.add_argument("--tile-rows", type=int, default=0, help="fused tile height (3, 7, 15)")
    ap.add_argument("--x-chunks", type=int, default=None,
                    help="x chunks per model (0 automatic; patch workloads default 16)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--xcd-group", type=int, default=None,
                    help="with --tile-rows: tile rows per XCD visit (0 launch order)")
    ap.add_argument("--no-dispersive", action="store_true",
                    help="cpml_box: skip the dispersive box and sweep runs the line carries")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path "
              "(every measured figure is per GPU); run with --gpus 1", file=sys.stderr)
        return 2

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.x_chunks is None:
        args.x_chunks = 16 if args.workload.startswith("patch") else 0
    if args.workload == "patch_sweep":
        return sweep(args, log)
    if args.workload == "patch_earlystop":
        return earlystop(args, log)

    import ems_amd as E
    from ems_amd import workloads as W

    t0 = time.perf_counter()
    energy_every = 0
    if args.workload == "cpml_box":
        wl = W.cpml_box(n=args.n or 512)
    elif args.workload == "dispersive_box":
        # Plane waves and the TF/SF gaps run the unfused schedule; the run
        # records the energy every 50 steps.
        wl = W.dispersive_box()
        args.unfused, energy_every = True, 50
    else:
        wl = W.vacuum_box(n=args.n or 50)
    steps = args.steps if args.steps is not None else wl.steps
    log(f"workload {wl.name}: {wl.grid.nodes()} nodes, built in {time.perf_counter()-t0:.1f} s")

    eng = E.Engine(wl.grid, wl.dt, wl.lanes)
    t0 = time.perf_counter()
    probes = wl.build_ops(eng)
    log(f"operators in {time.perf_counter()-t0:.1f} s")
    stream_gbs = eng.copy_bandwidth(1 << 30, 10)
    log(f"device copy bandwidth: {stream_gbs:.0f} GB/s (read + write)")
    tuned = None
    warm = args.warmup
    if args.unfused:
        eng.set_schedule(False)
    elif args.tile_rows or args.no_autotune:
        eng.set_schedule(args.mode, args.tile_rows, args.x_chunks)
        if args.xcd_group is not None:
            eng.set_xcd_group(args.xcd_group)
    else:
        # The autotuner's trial steps are part of the (untimed) warm-up.
        t0 = eng.timestep()
        best, trials = eng.autotune()
        tuned = {"mode": best[0], "tile_rows": best[1], "x_chunks": best[2],
                 "xcd_group": getattr(eng, "xcd_group", None),
                 "trials_ms": {f"m{m}_t{t}_c{c}": round(ms, 4) for m, t, c, ms in trials},
                 "xcd_trials_ms": {str(g): round(ms, 4) for g, ms in getattr(eng, "xcd_trials", [])}}
        log(f"autotune: mode {best[0]}, tile_rows {best[1]}, x_chunks {best[2]} ({best[3]:.3f} ms)")
        warm = max(0, args.warmup - (eng.timestep() - t0))
    eng.run(warm)
    eng.synchronize()
    log(f"setup + warm-up done ({eng.timestep()} steps)")

    energies = []
    t0 = time.perf_counter()
    if energy_every:
        for _ in range(steps // energy_every):
            eng.run(energy_every - 1)
            energies.append(eng.step_with_energy())
        eng.run(steps % energy_every)
    else:
        eng.run(steps)
    eng.synchronize()
    wall = time.perf_counter() - t0
    value = wl.cells * steps / wall / 1e9
    log(f"{steps} steps in {wall:.3f} s: {value:.2f} Gcell/s")
    vals = probes[0].values()
    import math
    assert all(math.isfinite(v) for v in vals), "non-finite probe"

    # Dominant kernel: mean launch duration from HIP events on the engine's
    # stream, algorithmic bytes from the lanes the engine holds.
    th, te, ts = eng.time_kernels(args.kernel_steps)
    hb, eb, cb = eng.traffic_model()
    if te == 0.0:
        name, dur, byt = "step_fused", th, hb
    elif th >= te:
        name, dur, byt = "h_update_pf" if eng.unfused_prefetch & 1 else "h_update", th, hb
    else:
        name, dur, byt = "e_update_pf" if eng.unfused_prefetch & 2 else "e_update", te, eb
    achieved = byt / (dur * 1e-3) / 1e9
    # HBM bytes per launch from the committed rocprof counter passes
    # (FETCH_SIZE, WRITE_SIZE, scaled by the copy-kernel calibration) of this
    # workload's dominant kernel, when there are some.
    traffic, traffic_cfg = None, None
    tpath = os.path.join(ROOT, "profiles", f"{name}_{wl.name}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        traffic, traffic_cfg = tr["traffic_bytes"], tr["config"]
    roof = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_measured_at": traffic_cfg,
            "bytes_per_launch": byt, "launch_ms": round(dur, 4),
            "step_ms_events": round(ts, 4), "device_copy_GBs": round(stream_gbs, 1),
            "frac_of_copy": round(achieved / stream_gbs, 4)}
    halves = f"H {th:.3f} ms, E {te:.3f} ms" if te else f"{th:.3f} ms"
    log(f"kernels: {halves}, step {ts:.3f} ms; dominant {name} {dur:.3f} ms, "
        f"{achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu_baseline:
        eng.close()
        cpu = cpu_baseline(wl, args.cpu_seconds, log)
    dispersive = sweep_line = None
    if args.workload == "cpml_box" and not args.no_dispersive:
        if args.no_cpu_baseline:
            eng.close()
        dispersive = child_line(log, ["--workload", "dispersive_box", "--no-cpu-baseline"],
                                ("metric", "value", "unit", "steps", "warmup", "ms_per_step",
                                 "dtype", "data", "config", "roofline"))
        # the batched leg of the 256-model patch sweep (each model of a batch
        # is bit-identical to its own engine: the tests pin that)
        sweep_line = child_line(log, ["--workload", "patch_sweep", "--skip-sequential"],
                                ("metric", "value", "unit", "steps", "dtype", "data", "config",
                                 "phases_s", "cell_updates", "batched_s", "roofline"))

    n = wl.grid.cells()
    line = {
        "metric": "Gcell-updates/s", "value": round(value, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": {"cpml_box": "synthetic (seeded graded mesh, FR-4 blocks, Gaussian source)",
                 "dispersive_box": "synthetic (seeded Lorentz/Drude spheres, conformal PEC "
                                   "cylinder, TF/SF plane wave 1-10 GHz)"}.get(
            args.workload, "synthetic (vacuum PEC box, Gaussian source)"),
        "config": {"workload": wl.name, "cells": list(n), "nodes": list(wl.grid.nodes()),
                   "cpml_cells": {"cpml_box": 8, "dispersive_box": 10}.get(args.workload, 0),
                   "energy_every": energy_every or None,
                   "probes": len(probes), "parallelism": "single GPU",
                   "schedule": "unfused" if args.unfused else "fused",
                   "mode": None if args.unfused else (tuned["mode"] if tuned else args.mode),
                   "autotune": tuned},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if dispersive is not None:
        line["dispersive_box"] = dispersive
    if sweep_line is not None:
        line["patch_sweep"] = sweep_line
    print(json.dumps(line), flush=True)
    return 0


def child_line(log, argv, keys):
    """This script run in a child process on another workload (after the
    CPML box's engine is closed), the listed figures of its line kept: the
    default run carries them beside the CPML box's."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__)] + list(argv)
    log("child: " + " ".join(cmd[1:]))
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        return {"error": "timed out after 600 s"}
    for l in r.stderr.splitlines():
        if l.startswith("[bench]"):
            log("  " + l)
    if r.returncode != 0:
        return {"error": f"exit status {r.returncode}", "stderr_tail": r.stderr[-400:]}
    try:
        d = json.loads(r.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": "no JSON line", "stdout_tail": r.stdout[-400:]}
    return {k: d.get(k) for k in keys}


if __name__ == "__main__":
    sys.exit(main())
