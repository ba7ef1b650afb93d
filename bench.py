#!/usr/bin/env python3
"""Throughput of the MI355X landscape step (FastScape_Execute_Step) on one GPU.

A "step" is one FastScape_Execute_Step over every landscape of a batch resident in HBM: uplift,
D8 receivers, closed-basin rerouting, drainage area, implicit stream power (Newton for n != 1), ADI
diffusion.  Synthetic, seeded inputs (fastscapelib_fortran/workloads.py); FP64 throughout.

Headline (``value``): the ragged batch -- 200 landscapes, nx and ny log-uniform in [65, 1025],
dx = 250 m, h = U[0,1) + a 10 m 2-D sinusoid of wavelength U[5, 50] km, n from {1, 1.5, 2},
m = 0.45 n, kf = 10**U(-6, -5), kd ~ U(1e-3, 1e-1), Newton tolerance 1e-8, BC 1010 / 1111
alternating, dt = 2e3 yr -- stepped as a batch, cell-steps/s over K timed steps of its 300 (after W
warm-up steps and the launch-shape autotuner), beside the same landscapes one at a time (``plain``, a
sample, each tuned the same way).

Further legs:
* uniform: 64 x 513 x 513 (dx = 200 m, seeds 0..63, n = 1), batched;
* mixed: one 4097 x 4097 multiple-flow landscape with deposition (g = 1) and 64 D8 landscapes of
  257 x 257 in one context;
* serving: Poisson arrivals (20 runs/s) over a window, grids log-uniform in [65, 2049], step counts
  log-uniform in [50, 2000], dt log-uniform in [1e3, 1e4] yr, D8 or multiple flow direction 50/50,
  parameters as the ragged batch, served by the continuous-batching loop on a landscape pool; p50/p99
  completion latency and cell-steps/s until a deadline;
* advect (uniform batch with advection first) and readouts (slope, curvature, chi, catchment, fluxes).

Prints ONE JSON line: metric, value, the plain rate and the batched/plain ratio, the roofline of the
dominant kernel (HIP events recorded around it on the stream it runs on, inside the timed region),
and the reference Fortran timed on this host's cores (one landscape per process) on a bounded sample
of the same batch.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fastscapelib-fortran_amd"))

from fastscapelib_fortran import workloads as W  # noqa: E402  (numpy recipes; does not touch the GPU)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-steps/s batched+served vs plain on 1 GPU; p50/p99 run latency"

# Algorithmic (compulsory) bytes of the single-kernel phases: every input read once and every
# output written once, per unit (node of the batch unless stated).
ALG_BYTES_PER_NODE = {
    "receivers": 52,     # uplift fused: R h,u,b  W ht,b, rec(i32), length
    "donors": 5,         # R rec  W dmask(u8)
    "lm_walk": 17,       # R (h, basin, preorder successor) record, candidate mask (closed-basin nodes ~ all)
    "area_prep": 65,     # R J, dmask, pos, donor pos, precip  W record (a0, 8 donor slots)
    "area_sweep": 48,    # R record  W Al
    "spl_prep": 48,      # R h, b, kf  W kfint, g, ht
    "spl_fact": 92,      # R J, rec, pos(x2), Al, length, kfint, ht(x2)  W area, lsed, record (node, rcv, f, ht)
    "spl_sweep": 44,     # dataflow sweep: R receiver position, factor, ht of the record  W Hl, Wl, HWl
    "spl_post": 100,     # R pos, Hl, HWl, ht, b, etot  W h, b, etot, erate, sedflux, hwater, lake_depth
}
ALG_BYTES_PER_EDGE = {}  # per basin-graph edge (the spanning-forest and tour phases are multi-kernel)

RAGGED_WORKLOAD = ("ragged batch 200 x log-uniform[65,1025]^2 (dx 250 m, noise + 10 m sinusoid), n in {1,1.5,2}, "
                   "Newton tol 1e-8, BC 1010/1111, dt 2e3")


def _sync():
    """Wait for the device (the library's own HIP runtime; a PyTorch build may bundle another)."""
    from fastscapelib_fortran import _lib
    _lib.device_synchronize()


GRAPHS = True  # replay captured HIP graphs (off: launch every kernel directly, for kernel-trace profilers)
TUNE = True    # run the launch-shape autotuner after the warm-up steps of a batch
SHAPES = None  # or launch shapes given on the command line (walk_gx, jump_blocks, tour_gx, nr_small_gx, area_wg, bor_gx[, flow_wg, flow_split, walk_big, mfd_split, tile_h, flow_nk])


def open_batch(fs, Ls, lanes=1):
    """One context for all landscapes, or (lanes > 1) fs.Lanes: several contexts stepping on their own
    streams, so one lane's per-landscape kernels overlap another's node-parallel ones."""
    sizes = [(L["nx"], L["ny"]) for L in Ls]
    return fs.Lanes(sizes, lanes) if lanes > 1 and len(Ls) > 1 else fs.Batch(sizes)


def _tune(b):
    if SHAPES:
        b.set_tuning(dict(zip(b.TUNING, SHAPES)))
        return b.tuning()
    return b.autotune() if TUNE else b.tuning()


def run_gpu(fs, Ls, steps, warmup, tune=True, lanes=1):
    with open_batch(fs, Ls, lanes) as b:
        b.graphs(GRAPHS)
        for k, L in enumerate(Ls):
            W.configure(b, L, k)
        if warmup:
            b.step(warmup)
        if tune:
            _tune(b)
            b.step(1)  # captures the graphs of the tuned shapes
        _sync()
        t0 = time.perf_counter()
        b.step(steps)
        _sync()
        dt = time.perf_counter() - t0
        stats = [b.stats(k) for k in range(len(Ls))]
        mem = b.memory()
    return dt, stats, mem


def run_headline(fs, Ls, steps, warmup, prof_steps, lanes=1):
    """The headline measurement.  A profiled window (every phase bracketed by HIP events, kernels
    launched one by one) gives the per-phase breakdown and names the dominant single-kernel phase;
    the timed region then replays each step as captured HIP graphs, with that one phase launched
    between its own HIP events on the step's stream, so its mean duration is measured live.  With
    lanes, the timers are lane 0's (its landscapes: ``sel``) while the other lanes run beside it."""
    with open_batch(fs, Ls, lanes) as b:
        multi = isinstance(b, fs.Lanes)
        sel = b.members[0] if multi else list(range(len(Ls)))
        lane_timers = (lambda: b.timers(lane=0)) if multi else b.timers
        b.graphs(GRAPHS)
        for k, L in enumerate(Ls):
            W.configure(b, L, k)
        if warmup:
            b.step(warmup)
        tuned = _tune(b)
        tuned = tuned if isinstance(tuned, list) else [tuned]
        b.profiling(True)
        b.reset_timers()
        b.step(prof_steps)
        prof = lane_timers()
        b.profiling(False)
        stats = [b.stats(k) for k in sel]
        dom = dominant_phase(prof, stats, sum(Ls[k]["nx"] * Ls[k]["ny"] for k in sel))
        b.time_phase(dom)
        b.step(1)  # captures the graphs
        b.reset_timers()
        _sync()
        t0 = time.perf_counter()
        b.step(steps)
        _sync()
        dt = time.perf_counter() - t0
        live = lane_timers()
        b.time_phase(None)
        stats_all = [b.stats(k) for k in range(len(Ls))]
        mem = b.memory()
    return dict(dt=dt, prof=prof, live=live, dom=dom, stats=stats_all, mem=mem, tuned=tuned, sel=sel,
                lanes=len(b.lanes) if multi else 1)


def plain_rate(fs, Ls, steps):
    """One landscape at a time (each alone in a context, after one untimed step and the same launch-shape
    tuning as the batch): cell-steps/s over the sample."""
    cells, secs = 0, 0.0
    for L in Ls:
        dt, _, _ = run_gpu(fs, [L], steps, 1, tune=True)
        cells += L["nx"] * L["ny"] * steps
        secs += dt
    return cells / secs


def spread(Ls, n):
    """n members spread over the batch sorted by size (smallest to largest), in batch order."""
    order = sorted(range(len(Ls)), key=lambda i: (Ls[i]["nx"] * Ls[i]["ny"], i))
    pick = {order[int(round(j * (len(order) - 1) / max(1, n - 1)))] for j in range(n)}
    return sorted(pick)


def uniform_leg(fs, steps, lanes):
    Ls = W.uniform_batch()
    dt, stats, mem = run_gpu(fs, Ls, steps, 2, tune=True, lanes=lanes)
    ncells = W.cells(Ls)
    return {"workload": "uniform batch 64 x 513x513 D8 n=1 g=0 kd=1e-2 fixed BC (dx 200 m, dt 1e3)",
            "landscapes": len(Ls), "cells": ncells, "lanes": lanes, "steps": steps,
            "ms_per_step": dt / steps * 1e3, "value": ncells * steps / dt, "hbm_bytes_resident": mem[0],
            "levels_per_landscape": int(np.mean([s["levels"] for s in stats])),
            "basins_per_landscape": int(np.mean([s["nbasins"] for s in stats]))}


def mixed_leg(fs, steps):
    """The large multiple-flow landscape and 64 small D8 landscapes, one context (one launch per phase)."""
    Ls = W.mixed_batch()
    dt, stats, mem = run_gpu(fs, Ls, steps, 1, tune=True)
    ncells = W.cells(Ls)
    return {"workload": "mixed batch: 1 x 4097^2 MFD p=1 g=1 (dx 100 m) + 64 x 257^2 D8 n=1, one context, dt 1e3",
            "cells": ncells, "steps": steps, "ms_per_step": dt / steps * 1e3, "value": ncells * steps / dt,
            "hbm_bytes_resident": mem[0], "gauss_seidel_iters_large": stats[0]["ngs"]}


def advect_leg(fs, steps, lanes):
    """The uniform batch with a velocity field (vx = 2 cm/yr, vy = -1 cm/yr, +-20 % noise), so every step
    starts with the advection of h, b and etot (fsx_set_v)."""
    Ls = W.uniform_batch()
    with open_batch(fs, Ls, lanes) as b:
        b.graphs(GRAPHS)
        for k, L in enumerate(Ls):
            W.configure(b, L, k)
            rng = np.random.default_rng(500 + k)
            nn = L["nx"] * L["ny"]
            b.set_v(k, 0.02 * (0.8 + 0.4 * rng.random(nn)), -0.01 * (0.8 + 0.4 * rng.random(nn)))
        b.step(2)
        _sync()
        t0 = time.perf_counter()
        b.step(steps)
        _sync()
        dt = time.perf_counter() - t0
    return {"workload": "uniform batch 64 x 513x513 with advection first (set_v)",
            "lanes": lanes, "steps": steps, "ms_per_step": dt / steps * 1e3, "value": W.cells(Ls) * steps / dt}


def readout_leg(fs):
    """The read-outs (slope, curvature, chi, catchment, fluxes) of every landscape of the uniform batch
    after two steps, one host call per landscape, each copying its nn values to host memory (PCIe
    included); then all landscapes in one launch."""
    from fastscapelib_fortran import _lib
    Ls = W.uniform_batch()
    nn = Ls[0]["nx"] * Ls[0]["ny"]
    res = {"workload": "64 x 513x513 after 2 steps, ms per landscape, including the copy of the field to host memory"}
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            W.configure(b, L, k)
        b.step(2)
        out = np.empty(nn, np.float64)
        calls = [("slope", lambda k: b.derived(k, _lib.FSX_SLOPE, out)),
                 ("curvature", lambda k: b.derived(k, _lib.FSX_CURVATURE, out)),
                 ("chi", lambda k: b.derived(k, _lib.FSX_CHI, out)),
                 ("catchment", lambda k: b.derived(k, _lib.FSX_CATCHMENT, out)),
                 ("fluxes", lambda k: b.fluxes(k))]
        for name, fn in calls:
            fn(0)
            t0 = time.perf_counter()
            for k in range(len(Ls)):
                fn(k)
            res[name] = round((time.perf_counter() - t0) / len(Ls) * 1e3, 4)
        batched = {}
        for name, fn in [("slope", lambda: b.derived_all(_lib.FSX_SLOPE)),
                         ("curvature", lambda: b.derived_all(_lib.FSX_CURVATURE)),
                         ("chi", lambda: b.derived_all(_lib.FSX_CHI)),
                         ("catchment", lambda: b.derived_all(_lib.FSX_CATCHMENT)),
                         ("fluxes", lambda: b.fluxes_all())]:
            fn()
            t0 = time.perf_counter()
            fn()
            batched[name] = round((time.perf_counter() - t0) / len(Ls) * 1e3, 4)
        res["batched"] = batched
        res["batched_note"] = "all landscapes in one launch, ms per landscape (the same copies to host)"
    return res, Ls[0]


def readout_cpu_baseline(L):
    """The same read-outs on the CPU, one landscape after one step, one core: the reference Fortran when
    its build is present (else the oracle's C restatement)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    use_ref = O.ref_available()
    if use_ref:
        o = O.FortranRef()
        o.setup(L["nx"], L["ny"])
    else:
        o = O.COracle(L["nx"], L["ny"])
    W.configure(o, L)
    o.execute_step()
    cpu = {"unit": "ms per landscape", "cores": 1, "kind": "reference" if use_ref else "port",
           "sample": f"1 x {L['nx']}x{L['ny']} after one step"}
    for name in ["slope", "curvature", "chi"] + (["catchment"] if use_ref else []):
        t0 = time.perf_counter()
        o.derived(name)
        cpu[name] = round((time.perf_counter() - t0) * 1e3, 4)
    t0 = time.perf_counter()
    o.fluxes()
    cpu["fluxes"] = round((time.perf_counter() - t0) * 1e3, 4)
    return cpu


HEAVY_CELLS = 1 << 20  # served runs above this many cells step in the first lane only


def serving_leg(fs, window, deadline, quantum, sample, lanes=3):
    """Arrivals of the serving stream during `window` seconds, served in real time by `lanes` pools (each
    stepping on its own stream, quanta launched as soon as a pool's last one finished); the loop stops at
    `deadline` seconds (runs still resident or waiting then are reported as unfinished)."""
    from fastscapelib_fortran import serving
    stream = [x for x in W.serving_stream(duration=window)]
    cap, per_node = serving.pool_capacity(0.8 / max(1, lanes))
    need = sum(L["nx"] * L["ny"] for _, _, L in stream)
    cap = int(min(cap, max(need, max(L["nx"] * L["ny"] for _, _, L in stream))))
    runs = [serving.Run(rid=i, nx=L["nx"], ny=L["ny"], steps=st, arrival=t, heavy=L["nx"] * L["ny"] > HEAVY_CELLS,
                        setup=(lambda p, k, L=L: W.configure(p, L, k))) for i, (t, st, L) in enumerate(stream)]
    offered = sum(L["nx"] * L["ny"] * st for _, st, L in stream)
    pools = [fs.Pool(cap, slots=256) for _ in range(max(1, lanes))]
    try:
        for p in pools:
            p.graphs(GRAPHS)
        srv = serving.Server(pools, quantum=quantum)
        rep = srv.serve(runs, realtime=True, deadline=deadline)
    finally:
        for p in pools:
            p.close()
    out = {"workload": f"Poisson 20 runs/s for {window:g} s; grids log-uniform[65,2049]^2, steps log-uniform[50,2000], "
                       "dt log-uniform[1e3,1e4], D8/MFD 50/50, n in {1,1.5,2}, Newton tol 1e-8",
           "runs": len(stream), "quantum": quantum, "lanes": len(pools), "pool_nodes": cap,
           "bytes_per_node": round(per_node, 1),
           "offered_cell_steps": offered, "deadline_s": deadline}
    out.update({k: rep[k] for k in ("runs_finished", "runs_unfinished", "cell_steps", "wall_s", "cell_steps_per_s",
                                    "p50_latency_s", "p99_latency_s", "busy_cell_steps_per_s")})
    out["latency_note"] = "completion latency (arrival to copy-out) of the runs finished by the deadline"
    pick = [stream[i][2] for i in spread([L for _, _, L in stream], sample)]
    out["plain"] = plain_rate(fs, pick, 5)
    out["plain_sample"] = f"{len(pick)} runs of the stream spread over its sizes, one at a time, 5 steps each"
    out["served_over_plain"] = out["busy_cell_steps_per_s"] / out["plain"]
    return out


def serving_ragged_leg(fs, rate, steps, quantum, lanes, plain):
    """The ragged batch's 200 landscapes as a served stream: Poisson arrivals at `rate` runs/s (seeded), each
    run one landscape for `steps` steps, served in real time by `lanes` pools (the landscapes above HEAVY_CELLS in
    the first); its rate over the busy time against the plain one-landscape-at-a-time rate of the same
    landscapes."""
    from fastscapelib_fortran import serving
    Ls = W.ragged_batch()
    rng = np.random.default_rng(2024)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(Ls)))
    cap, per_node = serving.pool_capacity(0.8 / max(1, lanes))
    cap = int(min(cap, W.cells(Ls)))
    runs = [serving.Run(rid=k, nx=L["nx"], ny=L["ny"], steps=steps, arrival=float(arrivals[k]),
                        heavy=L["nx"] * L["ny"] > HEAVY_CELLS, setup=(lambda p, i, L=L: W.configure(p, L, i)))
            for k, L in enumerate(Ls)]
    pools = [fs.Pool(cap, slots=256) for _ in range(max(1, lanes))]
    try:
        for p in pools:
            p.graphs(GRAPHS)
        rep = serving.Server(pools, quantum=quantum).serve(runs, realtime=True)
    finally:
        for p in pools:
            p.close()
    out = {"workload": f"{RAGGED_WORKLOAD}: each landscape a run of {steps} steps, Poisson arrivals {rate:g} runs/s",
           "runs": len(runs), "quantum": quantum, "lanes": len(pools), "pool_nodes": cap,
           "offered_cell_steps": W.cells(Ls) * steps}
    out.update({k: rep[k] for k in ("runs_finished", "runs_unfinished", "cell_steps", "wall_s", "cell_steps_per_s",
                                    "p50_latency_s", "p99_latency_s", "busy_cell_steps_per_s")})
    out["latency_note"] = "completion latency (arrival to copy-out) of every run"
    if plain:
        out["plain"] = plain
        out["served_over_plain"] = out["busy_cell_steps_per_s"] / plain
    return out


# ---- CPU baseline: the reference Fortran (oracle/_ref), one landscape per process -------------------
def _cpu_worker(args):
    L, steps, lib = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if lib:
        o = O.FortranRef(lib)
        o.setup(L["nx"], L["ny"])
    else:
        o = O.COracle(L["nx"], L["ny"])
    t0 = time.perf_counter()
    W.configure(o, L)
    for _ in range(steps):
        o.execute_step()
    o.field("h")
    return time.perf_counter() - t0


def cpu_baseline(Ls, procs, budget=6e6):
    """The ragged batch's landscapes on the host's cores, one landscape per process (the reference keeps
    its state in module variables), each for about `budget` cell-steps (its own step count)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    lib = O.REF8_LIB if O.ref_available(O.REF8_LIB) else None  # the build with Newton tolerance 1e-8
    if not lib:
        O.build_c()
    members = spread(Ls, procs)
    jobs = [(Ls[i], int(min(W.RAGGED_STEPS, max(2, round(budget / (Ls[i]["nx"] * Ls[i]["ny"]))))), lib)
            for i in members]
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        times = pool.map(_cpu_worker, jobs)
    makespan = time.perf_counter() - t0
    cs = sum(L["nx"] * L["ny"] * st for L, st, _ in jobs)
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"value": cs / makespan, "unit": "cell-steps/s", "cores": procs,
            "kind": "reference" if lib else "port", "host_cpu_count": os.cpu_count(), "host_cpus_usable": aff,
            "sample": f"{procs} landscapes of the ragged batch spread over its sizes, one process per core, each "
                      f"~{budget:.0e} cell-steps ({min(s for _, s, _ in jobs)}-{max(s for _, s, _ in jobs)} steps; "
                      + ("reference Fortran with Newton tolerance 1e-8, flang -O2" if lib else "C port") +
                      f"); makespan {makespan:.1f} s, per-process {min(times):.1f}-{max(times):.1f} s"}


def _alg_bytes(ph, stats, ncells):
    if ph in ALG_BYTES_PER_NODE:
        return ALG_BYTES_PER_NODE[ph] * ncells
    if ph in ALG_BYTES_PER_EDGE:
        return ALG_BYTES_PER_EDGE[ph] * sum(s["nconn"] for s in stats)
    return None


def dominant_phase(timers, stats, ncells):
    """The single-kernel phase with the largest time (phases with an algorithmic-bytes figure)."""
    best = None
    for ph, (ms, cnt) in timers.items():
        if cnt and _alg_bytes(ph, stats, ncells) is not None and (best is None or ms > best[1]):
            best = (ph, ms)
    return best[0]


def measured_traffic(ph, workload, lanes=1):
    """HBM bytes per launch of phase ph's kernel from the committed counter summary of this workload
    and lane count (FETCH_SIZE and WRITE_SIZE runs scaled by a copy of known size), or None."""
    for name in sorted(os.listdir(os.path.join(ROOT, "profiles"))):
        if not (name.startswith("hbm_traffic_") and name.endswith(".json")):
            continue
        try:
            with open(os.path.join(ROOT, "profiles", name)) as f:
                d = json.load(f)
        except (OSError, ValueError):
            continue
        k = d.get("kernels", {}).get(ph)
        if d.get("workload") == workload and d.get("lanes", 1) == lanes and k:
            return int(k["fetch_bytes"] + k["write_bytes"])
    return None


def roofline(ph, timers, stats, ncells):
    """achieved = algorithmic bytes per launch / mean launch time of phase ph."""
    ms, cnt = timers[ph]
    avg_ms = ms / cnt  # one timed launch per step (the donor phase's second launch when it runs twice)
    nbytes = _alg_bytes(ph, stats, ncells)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": ph,
            "avg_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": int(nbytes)}


def copy_leg(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (the HBM-counter calibration for the roofline's traffic)."""
    from fastscapelib_fortran import _lib
    x = _lib.DeviceBuffer(nbytes)
    y = _lib.DeviceBuffer(nbytes)
    _lib.device_copy(x, y, nbytes)  # first touch
    _lib.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        _lib.device_copy(y, x, nbytes)
    _lib.device_synchronize()
    dt = (time.perf_counter() - t0) / reps
    x.free()
    y.free()
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": dt * 1e3, "GBps": 2 * nbytes / dt / 1e9}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prof-steps", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=6)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--cpu-procs", type=int, default=0, help="0: the host's usable cores, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="ragged,plain,serve_ragged,uniform,mixed,serving,advect,readouts",
                    help="comma list of ragged (the headline), plain, serve_ragged (the ragged landscapes served as "
                         "an arriving stream), uniform, mixed, serving, advect, readouts; copy alone: HBM-counter "
                         "calibration")
    ap.add_argument("--serve-ragged-rate", type=float, default=50.0, help="arrivals per second of serve_ragged")
    ap.add_argument("--serve-ragged-steps", type=int, default=300, help="steps per run of serve_ragged")
    ap.add_argument("--uniform-steps", type=int, default=20)
    ap.add_argument("--mixed-steps", type=int, default=3)
    ap.add_argument("--serve-window", type=float, default=5.0)
    ap.add_argument("--serve-deadline", type=float, default=60.0)
    ap.add_argument("--quantum", type=int, default=8)
    ap.add_argument("--serve-lanes", type=int, default=3,
                    help="pools of the serving leg, each on its own stream (the first takes the large grids)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly instead of replaying HIP graphs (for kernel-trace profiling)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuner)")
    ap.add_argument("--lanes", type=int, default=2,
                    help="contexts stepping the batch side by side on their own streams (1: one context)")
    ap.add_argument("--shapes", default=None,
                    help="walk_gx,jump_blocks,tour_gx,nr_small_gx,area_wg,bor_gx[,flow_wg,flow_split,walk_big,mfd_split,"
                         "tile_h,flow_nk] instead of the autotuner (missing trailing values keep their defaults)")
    args = ap.parse_args()
    global GRAPHS, TUNE, SHAPES
    GRAPHS = not args.no_graphs
    TUNE = not args.no_tune and not args.shapes
    if args.shapes:
        SHAPES = [int(v) for v in args.shapes.split(",")]
    legs = set(args.legs.split(","))
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (landscapes shard "
              "across GPUs as independent batches, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    if legs == {"copy"}:
        print(json.dumps({"copy": copy_leg()}))
        return
    import fastscapelib_fortran as fs

    def note(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    Ls = W.ragged_batch()
    ncells = W.cells(Ls)
    out = {"metric": METRIC, "value": None, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded landscapes, fastscapelib_fortran/workloads.py)",
           "config": {"workload": RAGGED_WORKLOAD, "landscapes": len(Ls), "cells": ncells, "lanes": args.lanes,
                      "window": f"{args.steps} timed steps of the workload's 300, after {args.warmup} warm-up "
                                f"steps and the autotuner"}}
    if "ragged" in legs:
        note("ragged batch (headline)")
        r = run_headline(fs, Ls, args.steps, args.warmup, args.prof_steps, args.lanes)
        sel_cells = sum(Ls[k]["nx"] * Ls[k]["ny"] for k in r["sel"])
        out["value"] = ncells * args.steps / r["dt"]
        out["ms_per_step"] = r["dt"] / args.steps * 1e3
        out["config"]["lanes"] = r["lanes"]
        out["roofline"] = dict(roofline(r["dom"], r["live"], [r["stats"][k] for k in r["sel"]], sel_cells),
                               traffic=measured_traffic(r["dom"], RAGGED_WORKLOAD, r["lanes"]),
                               measured_on=(f"lane 0 of {r['lanes']}: {len(r['sel'])} landscapes, {sel_cells} nodes"
                                            if r["lanes"] > 1 else "all landscapes"))
        out["launch_shapes"] = {"per_lane": r["tuned"],
                                "chosen_by": "autotuner" if TUNE else ("given" if SHAPES else "defaults")}
        out["phase_ms_per_step"] = {p: round(ms / max(1, args.prof_steps), 4) for p, (ms, c) in r["prof"].items() if c}
        out["phase_ms_per_step_note"] = (f"profiled window of {args.prof_steps} steps, kernels launched one by one"
                                         + (f"; lane 0 of {r['lanes']} ({len(r['sel'])} landscapes), the others "
                                            "beside it" if r["lanes"] > 1 else ""))
        # every single-kernel phase with an algorithmic-bytes figure against the HBM peak, from the same window
        # (mean launch; the donor masks run twice per step)
        out["phase_hbm"] = {}
        for p, (ms, c) in r["prof"].items():
            nbytes = _alg_bytes(p, [r["stats"][k] for k in r["sel"]], sel_cells)
            if c and nbytes:
                gbs = nbytes / (ms / c * 1e-3) / 1e9
                out["phase_hbm"][p] = {"GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                                       "avg_launch_ms": round(ms / c, 4)}
        out["hbm_bytes_resident"] = r["mem"][0]
        out["newton_iters_per_step"] = float(np.sum([s["newton_iters"] for s in r["stats"]]))
        out["levels_max"] = int(max(s["levels"] for s in r["stats"]))
        out["basins_per_landscape"] = int(np.mean([s["nbasins"] for s in r["stats"]]))
    if "plain" in legs:
        note("plain (one landscape at a time)")
        pick = [Ls[i] for i in spread(Ls, args.plain_sample)]
        plain = plain_rate(fs, pick, args.plain_steps)
        out["plain"] = {"value": plain, "sample": f"{len(pick)} landscapes of the batch spread over its sizes, each "
                                                   f"alone, {args.plain_steps} steps after 1 warm-up step and the autotuner"}
        if out["value"]:
            out["batched_over_plain"] = out["value"] / plain
    if "serve_ragged" in legs:
        note("ragged landscapes served as a stream")
        if "plain" in out:
            plain_v = out["plain"]["value"]
        else:
            plain_v = plain_rate(fs, [Ls[i] for i in spread(Ls, args.plain_sample)], args.plain_steps)
        out["serve_ragged"] = serving_ragged_leg(fs, args.serve_ragged_rate, args.serve_ragged_steps, args.quantum,
                                                 args.lanes, plain_v)
    if "uniform" in legs:
        note("uniform batch")
        out["uniform"] = uniform_leg(fs, args.uniform_steps, args.lanes)
    if "mixed" in legs:
        note("mixed batch")
        out["mixed"] = mixed_leg(fs, args.mixed_steps)
    if "serving" in legs:
        note("serving")
        out["serving"] = serving_leg(fs, args.serve_window, args.serve_deadline, args.quantum, sample=6,
                                     lanes=args.serve_lanes)
    if "advect" in legs:
        note("advection")
        out["advect"] = advect_leg(fs, 10, args.lanes)
    if "readouts" in legs:
        note("read-outs")
        out["readouts"], L0 = readout_leg(fs)
        if not args.no_cpu:
            out["readouts"]["cpu_baseline"] = readout_cpu_baseline(L0)
    if not args.no_cpu:
        note("cpu baseline")
        aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        procs = args.cpu_procs or max(1, min(16, aff))
        out["cpu_baseline"] = cpu_baseline(Ls, procs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
