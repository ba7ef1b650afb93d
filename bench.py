#!/usr/bin/env python3
"""Throughput of the MI355X landscape step (FastScape_Execute_Step) on one GPU.

A "step" is one FastScape_Execute_Step over every landscape of a batch resident in HBM:
uplift, D8 receivers, closed-basin rerouting, drainage area, implicit stream power, ADI diffusion.
The default workload is the uniform batch: 64 independent 513x513 landscapes (dx = 200 m), seeds
0..63, h ~ U[0,1), kf log-uniform in [2e-6, 2e-5], uplift U ~ U[5e-4, 2e-3] m/yr inside and 0 on the
fixed boundary, m = 0.4, n = 1, kd = 1e-2, dt = 1e3 yr.  The same landscape run alone gives the plain
one-at-a-time figure.  Synthetic, seeded inputs; FP64 throughout.

Further legs (each reported beside the headline value):
* ragged: 200 landscapes, nx and ny log-uniform in [65, 1025], dx = 250 m, h = U[0,1) + a 10 m 2-D
  sinusoid of wavelength U[5, 50] km, n from {1, 1.5, 2}, m = 0.45 n, kf = 10**U(-6, -5),
  kd ~ U(1e-3, 1e-1), uplift 1e-3 m/yr off the fixed boundary, BC 1010 (even ids) / 1111 (odd),
  dt = 2e3 yr, Newton tolerance 1e-8; batched vs the same landscapes one at a time (a sample);
* serving: a Poisson stream of runs (20 runs/s) with grids log-uniform in [65, 513], step counts
  log-uniform in [20, 200], dt log-uniform in [1e3, 1e4], D8 or multiple flow direction (p = 1)
  50/50, other parameters as the ragged leg; served by the continuous-batching loop on a landscape
  pool; p50/p99 completion latency, and the same runs all submitted at once for the saturated rate;
* advect: the uniform batch with a velocity field, advection of h, b and etot first in every step;
* readouts: slope, curvature, chi, catchment and fluxes of every landscape of the uniform batch
  (ms per landscape), beside the reference Fortran on one of them.

Prints ONE JSON line: metric, value (batched cell-steps/s), plain rate and the batched/plain ratio,
the roofline of the dominant kernel (HIP events recorded around it on the stream it runs on during
the timed region), and the reference Fortran timed on this host's cores on a bounded sample.
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

HBM_PEAK_GBS = 8000.0
METRIC = "cell-steps/s batched+served vs plain on 1 GPU; p50/p99 run latency"

# Algorithmic (compulsory) bytes of the single-kernel phases: every input read once and every
# output written once, per unit (node of the batch unless stated).
ALG_BYTES_PER_NODE = {
    "uplift": 40,        # R h,u,b  W h,b
    "receivers": 20,     # R h  W rec(i32), length
    "donors": 5,         # R rec  W dmask(u8)
    "lm_walk": 17,       # R (h, basin, preorder successor) record, candidate mask (closed-basin nodes ~ all)
    "area_prep": 65,     # R J, dmask, pos, donor pos, precip  W record (a0, 8 donor slots)
    "area_sweep": 48,    # R record  W Al
    "spl_prep": 48,      # R h, b, kf  W kfint, g, ht
    "spl_fact": 92,      # R J, rec, pos(x2), Al, length, kfint, ht(x2)  W area, lsed, record (node, rcv, f, ht)
    "spl_sweep": 40,     # R record  W Hl, HWl (level order)
    "spl_post": 100,     # R pos, Hl, HWl, ht, b, etot  W h, b, etot, erate, sedflux, hwater, lake_depth
}
ALG_BYTES_PER_EDGE = {}  # per basin-graph edge (the spanning-forest and tour phases are multi-kernel)


def uniform_batch(ninst=64, n=513, dx=200.0):
    """Uniform batch recipe; per landscape: h noise, then kf, then U, from PCG64(seed)."""
    out = []
    for seed in range(ninst):
        rng = np.random.default_rng(seed)
        nn = n * n
        h = rng.random(nn)
        kf = 10.0 ** rng.uniform(np.log10(2e-6), np.log10(2e-5))
        U = rng.uniform(5e-4, 2e-3)
        u = np.full(nn, U)
        e = np.zeros((n, n), bool)
        e[0, :] = e[-1, :] = e[:, 0] = e[:, -1] = True
        u[e.ravel()] = 0.0
        out.append(dict(nx=n, ny=n, xl=dx * (n - 1), yl=dx * (n - 1), h=h, u=u, kf=kf, m=0.4, n=1.0, kd=1e-2,
                        dt=1e3, ibc=1111))
    return out


def _fixed_mask(nx, ny, ibc):
    """Nodes on a fixed (base-level) boundary for a BC code (digits bottom/right/top/left)."""
    c = f"{ibc:4d}"
    e = np.zeros((ny, nx), bool)
    if c[0] == "1":
        e[0, :] = True
    if c[1] == "1":
        e[:, -1] = True
    if c[2] == "1":
        e[-1, :] = True
    if c[3] == "1":
        e[:, 0] = True
    return e.ravel()


def _sinus_landscape(rng, nx, ny, dx, amp=10.0):
    lam = rng.uniform(5e3, 5e4)
    h = rng.random(nx * ny)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    return h + amp * (np.sin(2 * np.pi * x[None, :] / lam) * np.sin(2 * np.pi * y[:, None] / lam)).ravel()


def ragged_batch(ninst=200, lo=65, hi=1025, seed0=1000):
    """Ragged batch recipe; per landscape from PCG64(seed0 + id): nx, ny, h noise + sinusoid, then
    n, kf, kd."""
    out = []
    for i in range(ninst):
        rng = np.random.default_rng(seed0 + i)
        nx = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        ny = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        dx = 250.0
        h = _sinus_landscape(rng, nx, ny, dx)
        n = float(rng.choice([1.0, 1.5, 2.0]))
        kf = 10.0 ** rng.uniform(-6.0, -5.0)
        kd = rng.uniform(1e-3, 1e-1)
        ibc = 1010 if i % 2 == 0 else 1111
        u = np.full(nx * ny, 1e-3)
        u[_fixed_mask(nx, ny, ibc)] = 0.0
        out.append(dict(nx=nx, ny=ny, xl=dx * (nx - 1), yl=dx * (ny - 1), h=h, u=u, kf=kf, m=0.45 * n, n=n, kd=kd,
                        dt=2e3, ibc=ibc, newton_tol=1e-8))
    return out


def serving_stream(rate=20.0, duration=8.0, lo=65, hi=513, smin=20, smax=200, seed=7):
    """Arrivals of the serving leg: Poisson(rate) over `duration` seconds; per run grid, steps, dt,
    D8 or multiple flow direction, parameters as the ragged recipe (from PCG64(seed), then per run
    PCG64(seed*100000 + id))."""
    rng = np.random.default_rng(seed)
    t, runs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        i = len(runs)
        r = np.random.default_rng(seed * 100000 + i)
        nx = int(round(np.exp(r.uniform(np.log(lo), np.log(hi)))))
        ny = int(round(np.exp(r.uniform(np.log(lo), np.log(hi)))))
        steps = int(round(np.exp(r.uniform(np.log(smin), np.log(smax)))))
        dt = float(np.exp(r.uniform(np.log(1e3), np.log(1e4))))
        dx = 250.0
        h = _sinus_landscape(r, nx, ny, dx)
        n = float(r.choice([1.0, 1.5, 2.0]))
        kf = 10.0 ** r.uniform(-6.0, -5.0)
        kd = r.uniform(1e-3, 1e-1)
        mfd = bool(r.random() < 0.5)
        ibc = 1010 if i % 2 == 0 else 1111
        u = np.full(nx * ny, 1e-3)
        u[_fixed_mask(nx, ny, ibc)] = 0.0
        runs.append((t, steps, dict(nx=nx, ny=ny, xl=dx * (nx - 1), yl=dx * (ny - 1), h=h, u=u, kf=kf, m=0.45 * n,
                                    n=n, kd=kd, dt=dt, ibc=ibc, newton_tol=1e-8, p=1.0 if mfd else -2.0)))
    return runs


def configure(obj, L, *k):
    obj.set_xl_yl(*k, L["xl"], L["yl"])
    obj.set_dt(*k, L["dt"])
    obj.init_h(*k, L["h"])
    nn = L["nx"] * L["ny"]
    obj.set_erosional_parameters(*k, np.full(nn, L["kf"]), -1.0, L["m"], L["n"], np.full(nn, L["kd"]), -1.0, 0.0,
                                 0.0, L.get("p", -2.0))
    obj.set_u(*k, L["u"])
    obj.set_bc(*k, L["ibc"])
    if "newton_tol" in L and hasattr(obj, "set_newton_tol"):
        obj.set_newton_tol(*k, L["newton_tol"])


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


GRAPHS = True  # replay captured HIP graphs (off: launch every kernel directly, for kernel-trace profilers)
TUNE = True    # run the launch-shape autotuner after the warm-up steps of a batch
SHAPES = None  # or the headline batch's launch shapes, given (walk_gx, jump_blocks, tour_gx, nr_small_gx)


def open_batch(fs, Ls, lanes=1):
    """One context for all landscapes, or (lanes > 1) fs.Lanes: several contexts stepping on their own
    streams, so one lane's per-landscape kernels overlap another's node-parallel ones."""
    sizes = [(L["nx"], L["ny"]) for L in Ls]
    return fs.Lanes(sizes, lanes) if lanes > 1 else fs.Batch(sizes)


def run_gpu(fs, Ls, steps, warmup, profile, tune=False, lanes=1):
    with open_batch(fs, Ls, lanes) as b:
        b.graphs(GRAPHS)
        for k, L in enumerate(Ls):
            configure(b, L, k)
        if warmup:
            b.step(warmup)
        if tune and TUNE:
            b.autotune()
        if profile:
            b.profiling(True)
            b.reset_timers()
        _sync()
        t0 = time.perf_counter()
        b.step(steps)
        _sync()
        dt = time.perf_counter() - t0
        timers = b.timers() if profile else None
        stats = [b.stats(k) for k in range(len(Ls))]
        mem = b.memory()
    return dt, timers, stats, mem


def run_headline(fs, Ls, steps, warmup, prof_steps, lanes=1):
    """The headline measurement.  A profiled window (every phase bracketed by HIP events, kernels
    launched one by one) gives the per-phase breakdown and names the dominant single-kernel phase;
    the timed region then replays each step as captured HIP graphs, with that one phase launched
    between its own HIP events on the step's stream, so its mean duration is measured live.  With
    lanes, the timers are lane 0's (its landscapes: ``sel``) while the other lanes run beside it."""
    with open_batch(fs, Ls, lanes) as b:
        sel = b.members[0] if lanes > 1 else list(range(len(Ls)))
        lane_timers = (lambda: b.timers(lane=0)) if lanes > 1 else b.timers
        b.graphs(GRAPHS)
        for k, L in enumerate(Ls):
            configure(b, L, k)
        if warmup:
            b.step(warmup)
        if SHAPES:
            b.set_tuning(dict(zip(b.TUNING, SHAPES)))
        tuned = b.autotune() if TUNE else b.tuning()
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
        stats = [b.stats(k) for k in range(len(Ls))]
        mem = b.memory()
    return dt, prof, live, dom, stats, mem, tuned, sel


def plain_rate(fs, Ls, steps):
    """One landscape at a time (each after one untimed step): cell-steps/s over the sample."""
    cells, secs = 0, 0.0
    for L in Ls:
        dt, _, _, _ = run_gpu(fs, [L], steps, 1, profile=False)
        cells += L["nx"] * L["ny"] * steps
        secs += dt
    return cells / secs


def ragged_leg(fs, steps, sample, lanes=1):
    Ls = ragged_batch()
    ncells = sum(L["nx"] * L["ny"] for L in Ls)
    dt, _, stats, mem = run_gpu(fs, Ls, steps, 1, profile=False, tune=True, lanes=lanes)
    batched = ncells * steps / dt
    pick = Ls[::max(1, len(Ls) // sample)][:sample]
    plain = plain_rate(fs, pick, max(2, steps // 2))
    return {"workload": "ragged batch 200 x log-uniform[65,1025]^2, n in {1,1.5,2}, Newton tol 1e-8, BC 1010/1111",
            "landscapes": len(Ls), "cells": ncells, "lanes": lanes, "steps": steps, "ms_per_step": dt / steps * 1e3,
            "value": batched, "plain": plain, "plain_sample": f"{len(pick)} landscapes of the batch, one at a time",
            "batched_over_plain": batched / plain, "hbm_bytes_resident": mem[0],
            "newton_iters_per_step": float(np.sum([s["newton_iters"] for s in stats]))}


def advect_leg(fs, Ls, steps, lanes):
    """The uniform batch with a velocity field (vx = 2 cm/yr, vy = -1 cm/yr, +-20 % noise), so every step
    starts with the advection of h, b and etot (fsx_set_v)."""
    with open_batch(fs, Ls, lanes) as b:
        b.graphs(GRAPHS)
        for k, L in enumerate(Ls):
            configure(b, L, k)
            rng = np.random.default_rng(500 + k)
            nn = L["nx"] * L["ny"]
            b.set_v(k, 0.02 * (0.8 + 0.4 * rng.random(nn)), -0.01 * (0.8 + 0.4 * rng.random(nn)))
        b.step(2)
        _sync()
        t0 = time.perf_counter()
        b.step(steps)
        _sync()
        dt = time.perf_counter() - t0
    ncells = sum(L["nx"] * L["ny"] for L in Ls)
    return {"workload": f"uniform batch {len(Ls)} x {Ls[0]['nx']}x{Ls[0]['ny']} with advection first (set_v)",
            "lanes": lanes, "steps": steps, "ms_per_step": dt / steps * 1e3, "value": ncells * steps / dt}


def readout_leg(fs, Ls):
    """The read-outs (slope, curvature, chi, catchment, fluxes) of every landscape of the batch after two
    steps, one host call per landscape, each copying its nn values to host memory (PCIe included)."""
    from fastscapelib_fortran import _lib
    nn = Ls[0]["nx"] * Ls[0]["ny"]
    res = {"workload": f"{len(Ls)} x {Ls[0]['nx']}x{Ls[0]['ny']} after 2 steps, ms per landscape, "
                       "including the copy of the field to host memory"}
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            configure(b, L, k)
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
    return res


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
    configure(o, L)
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


def serving_leg(fs, quantum, duration, sample):
    from fastscapelib_fortran import serving
    stream = serving_stream(duration=duration)
    cap, per_node = serving.pool_capacity(0.5)
    need = sum(L["nx"] * L["ny"] for _, _, L in stream)
    cap = int(min(cap, max(need, 1)))

    def make(realtime):
        return [serving.Run(rid=i, nx=L["nx"], ny=L["ny"], steps=st, arrival=t,
                            setup=(lambda p, k, L=L: configure(p, L, k))) for i, (t, st, L) in enumerate(stream)]
    out = {"workload": f"Poisson 20 runs/s for {duration:g} s, grids log-uniform[65,513]^2, steps log-uniform[20,200],"
                       " D8/MFD 50/50", "runs": len(stream), "quantum": quantum, "pool_nodes": cap,
           "bytes_per_node": round(per_node, 1)}
    offered = sum(L["nx"] * L["ny"] * st for _, st, L in stream) / duration
    with fs.Pool(cap, slots=256) as p:
        p.graphs(GRAPHS)
        rep = serving.Server(p, quantum=quantum).serve(make(True), realtime=True)
    out.update({"offered_cell_steps_per_s": offered, "achieved_cell_steps_per_s": rep["cell_steps_per_s"],
                "p50_latency_s": rep["p50_latency_s"], "p99_latency_s": rep["p99_latency_s"],
                "wall_s": rep["wall_s"], "rejected": rep["rejected"]})
    with fs.Pool(cap, slots=256) as p:
        p.graphs(GRAPHS)
        sat = serving.Server(p, quantum=quantum).serve(make(False), realtime=False)
    out["saturated_cell_steps_per_s"] = sat["cell_steps_per_s"]
    pick = [L for _, _, L in stream][::max(1, len(stream) // sample)][:sample]
    out["plain"] = plain_rate(fs, pick, 10)
    out["served_over_plain"] = sat["cell_steps_per_s"] / out["plain"]
    return out


# ---- CPU baseline: the reference Fortran (oracle/_ref), one landscape per process -------------------
def _cpu_worker(args):
    L, steps, use_ref = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if use_ref:
        o = O.FortranRef()
        o.setup(L["nx"], L["ny"])
    else:
        o = O.COracle(L["nx"], L["ny"])
    t0 = time.perf_counter()
    configure(o, L)
    for _ in range(steps):
        o.execute_step()
    o.field("h")
    return time.perf_counter() - t0


def cpu_baseline(Ls, steps, procs):
    ref_lib = os.path.join(ROOT, "oracle", "_ref", "libfsref.so")
    use_ref = os.path.exists(ref_lib)
    if not use_ref:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        O.build_c()
    jobs = [(Ls[i % len(Ls)], steps, use_ref) for i in range(procs)]
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        times = pool.map(_cpu_worker, jobs)
    makespan = time.perf_counter() - t0
    cells = sum(L["nx"] * L["ny"] for L, _, _ in jobs)
    return {"value": cells * steps / makespan, "unit": "cell-steps/s", "cores": procs,
            "kind": "reference" if use_ref else "port",
            "sample": f"{procs} x {Ls[0]['nx']}x{Ls[0]['ny']} landscapes of the batch, {steps} steps each, one "
                      f"process per core ({'unmodified reference Fortran, flang -O2' if use_ref else 'C port'}); "
                      f"makespan {makespan:.1f} s, per-process {min(times):.1f}-{max(times):.1f} s"}


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
    path = os.path.join(ROOT, "profiles", "hbm_traffic_uniform_batch_64x513.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    k = d.get("kernels", {}).get(ph)
    if d.get("workload") != workload or d.get("lanes", 1) != lanes or not k:
        return None
    return int(k["fetch_bytes"] + k["write_bytes"])


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
    import torch
    n = nbytes // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": dt * 1e3, "GBps": 2 * nbytes / dt / 1e9}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prof-steps", type=int, default=10)
    ap.add_argument("--ninst", type=int, default=64)
    ap.add_argument("--n", type=int, default=513)
    ap.add_argument("--plain-steps", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--cpu-procs", type=int, default=0, help="0: min(16, host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="uniform,plain,ragged,serving,advect,readouts",
                    help="comma list of uniform, plain, ragged, serving, copy (copy alone: HBM-counter calibration)")
    ap.add_argument("--ragged-steps", type=int, default=10)
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--quantum", type=int, default=8)
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly instead of replaying HIP graphs (for kernel-trace profiling)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuner)")
    ap.add_argument("--lanes", type=int, default=2,
                    help="contexts stepping the batch side by side on their own streams (1: one context)")
    ap.add_argument("--shapes", default=None,
                    help="walk_gx,jump_blocks,tour_gx,nr_small_gx for the headline batch instead of the autotuner")
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

    Ls = uniform_batch(args.ninst, args.n)
    workload = f"uniform batch {args.ninst} x {args.n}x{args.n} D8 n=1 g=0 kd=1e-2 fixed BC"
    ncells = sum(L["nx"] * L["ny"] for L in Ls)
    dt, timers, live, dom, stats, mem, tuned, sel = run_headline(fs, Ls, args.steps, args.warmup, args.prof_steps,
                                                                 args.lanes)
    value = ncells * args.steps / dt
    sel_cells = sum(Ls[k]["nx"] * Ls[k]["ny"] for k in sel)

    # plain path: one landscape at a time (the first of the batch)
    plain, pdt = None, None
    if "plain" in legs:
        pdt, _, _, _ = run_gpu(fs, Ls[:1], args.plain_steps, 1, profile=False)
        plain = Ls[0]["nx"] * Ls[0]["ny"] * args.plain_steps / pdt

    phases = {p: round(ms / max(1, args.prof_steps), 4) for p, (ms, c) in timers.items() if c}
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded uniform-noise landscapes)",
        "config": {"workload": workload,
                   "landscapes": args.ninst, "cells": ncells, "lanes": args.lanes},
        "plain": ({"value": plain, "ms_per_step": pdt / args.plain_steps * 1e3,
                   "workload": f"1 x {args.n}x{args.n}, same parameters"} if plain else None),
        "batched_over_plain": value / plain if plain else None,
        "roofline": dict(roofline(dom, live, [stats[k] for k in sel], sel_cells),
                         traffic=measured_traffic(dom, workload, args.lanes),
                         measured_on=f"lane 0 of {args.lanes}: {len(sel)} landscapes" if args.lanes > 1 else "all"),
        "launch_shapes": {"per_lane": tuned,
                          "chosen_by": "autotuner" if TUNE else ("given" if SHAPES else "defaults")},
        "phase_ms_per_step": phases,
        "phase_ms_per_step_note": f"profiled window of {args.prof_steps} steps, kernels launched one by one"
                                  + (f"; lane 0 of {args.lanes} ({len(sel)} landscapes), the others beside it"
                                     if args.lanes > 1 else ""),
        "step_model_frac": round(336.0 * ncells / (dt / args.steps) / 1e9 / HBM_PEAK_GBS, 5),
        "hbm_bytes_resident": mem[0],
        "basins_per_landscape": int(np.mean([s["nbasins"] for s in stats])),
        "levels_per_landscape": int(np.mean([s["levels"] for s in stats])),
    }
    if "ragged" in legs:
        out["ragged"] = ragged_leg(fs, args.ragged_steps, sample=8, lanes=args.lanes)
    if "serving" in legs:
        out["serving"] = serving_leg(fs, args.quantum, args.serve_seconds, sample=8)
    if "advect" in legs:
        out["advect"] = advect_leg(fs, Ls, 20, args.lanes)
    if "readouts" in legs:
        out["readouts"] = readout_leg(fs, Ls)
        if not args.no_cpu:
            out["readouts"]["cpu_baseline"] = readout_cpu_baseline(Ls[0])
    if not args.no_cpu:
        procs = args.cpu_procs or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(Ls, args.cpu_steps, procs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
