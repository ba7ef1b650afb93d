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


class Server:
    """Continuous-batching loop over a landscape pool (see module docstring)."""

    def __init__(self, pool, quantum: int = 8, clock=time.perf_counter):
        self.pool = pool
        self.quantum = int(quantum)
        self.clock = clock
        self.queue: deque = deque()
        self.active: Dict[int, Run] = {}
        self.finished: List[Run] = []
        self.rejected: List[Run] = []
        self.cell_steps = 0
        self.quanta = 0
        self.busy = 0.0  # wall time of the quanta (admission, steps and retirement while runs are resident)

    def submit(self, run: Run) -> None:
        run.left = run.steps
        self.queue.append(run)

    def _admit(self) -> None:
        waiting = deque()
        while self.queue:
            run = self.queue.popleft()
            k = self.pool.add(run.nx, run.ny)
            if k is None:
                if not self.active and not self.pool.nn:  # does not fit even an empty pool
                    run.done = self.clock()
                    self.rejected.append(run)
                else:
                    waiting.append(run)
                continue
            run.slot = k
            run.setup(self.pool, k)
            run.admitted = self.clock()
            if run.left == 0:
                self._retire(run)
            else:
                self.active[k] = run
        self.queue = waiting

    def _retire(self, run: Run) -> None:
        for name in run.outputs:
            if name in _IFIELDS:
                run.result[name] = self.pool.ifield(run.slot, _IFIELDS[name])
            else:
                run.result[name] = self.pool.field(run.slot, _FIELDS[name])
        self.pool.remove(run.slot)
        run.done = self.clock()
        self.active.pop(run.slot, None)
        self.finished.append(run)

    def tick(self) -> int:
        """Admit, advance every resident run by one quantum, retire finished runs.  Returns the
        number of steps taken (0 when nothing is resident)."""
        t0 = self.clock()
        self._admit()
        if not self.active:
            return 0
        q = min(self.quantum, min(r.left for r in self.active.values()))
        self.pool.step(q)
        self.quanta += 1
        for run in list(self.active.values()):
            run.left -= q
            self.cell_steps += run.cells * q
            if run.left == 0:
                self._retire(run)
        self.busy += self.clock() - t0
        return q

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
        while i < len(pending) or self.queue or self.active:
            now = self.clock()
            if deadline is not None and now - t0 >= deadline:
                break
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i])
                i += 1
            if self.tick() == 0 and i < len(pending) and not self.queue:
                time.sleep(max(0.0, min(0.05, pending[i].arrival - self.clock())))
        rep = self.report(self.clock() - t0)
        rep["runs_unfinished"] = len(pending) - len(self.finished) - len(self.rejected)
        return rep

    def report(self, wall: float) -> dict:
        lat = np.array([r.latency for r in self.finished]) if self.finished else np.full(1, np.nan)
        return {"runs": len(self.finished), "runs_finished": len(self.finished), "rejected": len(self.rejected),
                "cell_steps": int(self.cell_steps), "wall_s": wall,
                "cell_steps_per_s": self.cell_steps / wall if wall > 0 else 0.0,
                "busy_cell_steps_per_s": self.cell_steps / self.busy if self.busy > 0 else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "quanta": self.quanta}
