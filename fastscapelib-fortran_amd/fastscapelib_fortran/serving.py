"""Continuous batching of landscape runs on one GPU.

The reference runs one landscape per process, start to finish (src/FastScape_api.f90:201-263 called
in the user's time loop).  A server here keeps every admitted run's state resident in one landscape
pool (``Pool``) and steps all of them with one set of kernel launches per step.  Runs of any grid size
and step count arrive at any time; at every quantum boundary the server

* retires the runs that reached their step count (copies their outputs out, frees their nodes),
* admits waiting runs in arrival order while a landscape slot and a contiguous node range of
  nx*ny nodes are free (the pool's arena is sized from free HBM and the library's bytes per node),
  letting a later, smaller run go ahead of one that does not fit yet,
* advances every resident run by q = min(quantum, fewest steps left) steps.

With several pools (lanes) each lane runs its own quanta on its own stream and is attended to as soon as its
quantum finishes, so the device keeps stepping the other lanes while the host retires and admits.

Each run's result does not depend on which other runs shared its steps (the kernels index every
landscape by its own node range and reduce per landscape).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib


@dataclass
class Run:
    """One landscape run: grid, step count and a ``setup(pool, k)`` that configures landscape k of
    the pool (xl/yl, dt, init_h, erosional parameters, uplift, BC ...).  ``outputs`` names the fields
    copied out when it completes: h, b, etot, erate, area, lake_depth, rec, ndon."""
    rid: int
    nx: int
    ny: int
    steps: int
    setup: Callable
    arrival: float = 0.0
    outputs: tuple = ("h",)
    heavy: bool = False  # a run whose steps take long (large grid): served in the first lane only
    admitted: Optional[float] = None
    done: Optional[float] = None
    left: int = 0
    slot: int = -1
    result: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.nx * self.ny

    @property
    def latency(self) -> float:
        return self.done - self.arrival


_FIELDS = {"h": _lib.FSX_H, "b": _lib.FSX_B, "etot": _lib.FSX_ETOT, "erate": _lib.FSX_ERATE,
           "area": _lib.FSX_AREA, "lake_depth": _lib.FSX_LAKE_DEPTH}
_IFIELDS = {"rec": _lib.FSX_REC, "ndon": _lib.FSX_NDON}  # receivers and donor counts of the last step


def pool_capacity(hbm_fraction=0.8, slots=256):
    """Nodes a pool may reserve: the given fraction of free HBM over the library's bytes per node
    (measured on a small pool, multiple-flow buffers included)."""
    from . import Pool
    free, _ = _lib.device_memory()
    with Pool(4096, 1) as p:
        k = p.add(3, 3)
        p.set_erosional_parameters(k, np.zeros(9), -1.0, 0.4, 1.0, np.zeros(9), -1.0, 0.0, 0.0, 1.0)  # MFD buffers
        total, _ = p.memory()
    per_node = total / 4096.0
    return int(free * hbm_fraction / per_node), per_node


class _Lane:
    """One pool of the server and the quantum it has in flight."""

    def __init__(self, pool):
        self.pool = pool
        self.active: Dict[int, Run] = {}
        self.q = 0  # steps in flight (0: idle)
        self.cells = 0

    def free(self) -> int:
        cap = getattr(self.pool, "capacity", 0)
        return cap - sum(r.cells for r in self.active.values())


class Server:
    """Continuous-batching loop over one or several landscape pools (see module docstring).  With
    several pools ("lanes", each stepping on its own HIP stream) a lane's next quantum is launched as soon
    as its last one finished and its runs were retired and admitted, while the other lanes' quanta run:
    the host's copy-outs, admissions and graph captures overlap the device's steps, and one lane's
    per-landscape kernels overlap another's node-parallel ones."""

    def __init__(self, pool, quantum: int = 8, clock=time.perf_counter):
        self.lanes = [_Lane(p) for p in (pool if isinstance(pool, (list, tuple)) else [pool])]
        self.pool = self.lanes[0].pool
        self.quantum = int(quantum)
        self.clock = clock
        self.queue: deque = deque()
        self.finished: List[Run] = []
        self.rejected: List[Run] = []
        self.cell_steps = 0
        self.quanta = 0
        self.busy = 0.0  # wall time with runs resident

    @property
    def active(self) -> Dict[int, Run]:
        out = {}
        for j, ln in enumerate(self.lanes):
            out.update({(j, k): r for k, r in ln.active.items()})
        return out

    def submit(self, run: Run) -> None:
        run.left = run.steps
        self.queue.append(run)

    def _admit(self, ln: _Lane) -> None:
        """Waiting runs, in arrival order, that fit the idle lane ln (a later, smaller run may go ahead of
        one that does not fit yet); a run that fits no empty pool is rejected.  With several lanes, heavy
        runs go to the first lane and the others to the rest, so that the short steps of small grids are
        not held to the pace of the largest ones."""
        waiting = deque()
        split = len(self.lanes) > 1
        first = ln is self.lanes[0]
        while self.queue:
            run = self.queue.popleft()
            if split and run.heavy != first:
                waiting.append(run)
                continue
            k = ln.pool.add(run.nx, run.ny)
            if k is None:
                if all(not x.active and not x.pool.nn for x in self.lanes):  # fits no empty pool
                    run.done = self.clock()
                    self.rejected.append(run)
                else:
                    waiting.append(run)
                continue
            run.slot = k
            run.setup(ln.pool, k)
            run.admitted = self.clock()
            if run.left == 0:
                self._retire(ln, run)
            else:
                ln.active[k] = run
        self.queue = waiting

    def _retire(self, ln: _Lane, run: Run) -> None:
        for name in run.outputs:
            if name in _IFIELDS:
                run.result[name] = ln.pool.ifield(run.slot, _IFIELDS[name])
            else:
                run.result[name] = ln.pool.field(run.slot, _FIELDS[name])
        ln.pool.remove(run.slot)
        run.done = self.clock()
        ln.active.pop(run.slot, None)
        self.finished.append(run)

    def _complete(self, ln: _Lane) -> None:
        """The lane's quantum finished: count its steps, retire the runs that reached their step count."""
        q = ln.q
        ln.q = 0
        self.quanta += 1
        for run in list(ln.active.values()):
            run.left -= q
            self.cell_steps += run.cells * q
            if run.left == 0:
                self._retire(ln, run)

    def tick(self) -> int:
        """One pass over the lanes: a lane whose quantum finished is retired, refilled from the queue and
        given its next quantum (the lane with the most free nodes is served first).  Returns the steps
        launched (0 when nothing was launched)."""
        launched = 0
        for ln in sorted(self.lanes, key=lambda x: -x.free()):
            if ln.q:
                ready = getattr(ln.pool, "ready", None)
                if ready is not None and not ready():
                    continue  # still stepping
                self._complete(ln)
            self._admit(ln)
            if not ln.active:
                continue
            q = min(self.quantum, min(r.left for r in ln.active.values()))
            if hasattr(ln.pool, "ready"):
                ln.pool.step(q, sync=False)
            else:
                ln.pool.step(q)
            ln.q = q
            launched += q
        return launched

    def serve(self, schedule: List[Run], realtime: bool = True, deadline: Optional[float] = None) -> dict:
        """Drive a whole arrival schedule (runs with ``arrival`` in seconds from the start).  With
        ``realtime`` the loop waits for arrivals on the wall clock; otherwise every run is available
        at time 0.  With ``deadline`` (seconds) the loop stops then, leaving unfinished runs where
        they are (counted in the report)."""
        pending = sorted(schedule, key=lambda r: r.arrival)
        t0 = self.clock()
        i = 0
        base = t0 if realtime else None
        for r in pending:
            r.arrival = (base + r.arrival) if realtime else t0
        last = t0
        while i < len(pending) or self.queue or any(ln.active or ln.q for ln in self.lanes):
            now = self.clock()
            if any(ln.q for ln in self.lanes):
                self.busy += now - last
            last = now
            if deadline is not None and now - t0 >= deadline:
                break
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i])
                i += 1
            if self.tick() == 0:
                if any(ln.q for ln in self.lanes):
                    time.sleep(0.0002)  # every lane is stepping
                elif i < len(pending) and not self.queue:
                    time.sleep(max(0.0, min(0.05, pending[i].arrival - self.clock())))
        for ln in self.lanes:  # let the quanta in flight finish (their steps are not counted past the deadline)
            if ln.q and hasattr(ln.pool, "synchronize"):
                ln.pool.synchronize()
        rep = self.report(self.clock() - t0)
        rep["runs_unfinished"] = len(pending) - len(self.finished) - len(self.rejected)
        rep["lanes"] = len(self.lanes)
        return rep

    def report(self, wall: float) -> dict:
        lat = np.array([r.latency for r in self.finished]) if self.finished else np.full(1, np.nan)
        return {"runs": len(self.finished), "runs_finished": len(self.finished), "rejected": len(self.rejected),
                "cell_steps": int(self.cell_steps), "wall_s": wall,
                "cell_steps_per_s": self.cell_steps / wall if wall > 0 else 0.0,
                "busy_cell_steps_per_s": self.cell_steps / self.busy if self.busy > 0 else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "quanta": self.quanta}
