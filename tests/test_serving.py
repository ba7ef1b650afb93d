"""Host logic of the continuous-batching server, on a stand-in pool (no GPU): admission by free
slots and contiguous node ranges, backfill, quanta never overshooting a run's step count,
retirement, rejection of runs that can never fit, latency bookkeeping."""
from __future__ import annotations

import numpy as np
import pytest


class FakePool:
    """Arena of `capacity` nodes with first-fit ranges, like the library's pool."""

    def __init__(self, capacity, slots):
        self.capacity, self.slots = capacity, slots
        self.nn, self.ranges, self.steps = {}, {}, {}
        self.log = []

    def add(self, nx, ny):
        n = nx * ny
        if len(self.nn) >= self.slots:
            return None
        used = sorted(self.ranges.values())
        off = 0
        for a, b in used:
            if a - off >= n:
                break
            off = b
        if off + n > self.capacity:
            return None
        k = next(s for s in range(self.slots) if s not in self.nn)
        self.nn[k], self.ranges[k], self.steps[k] = n, (off, off + n), 0
        return k

    def remove(self, k):
        del self.nn[k], self.ranges[k], self.steps[k]

    def step(self, q):
        assert q > 0
        for k in self.nn:
            self.steps[k] += q
        self.log.append((q, sorted(self.nn)))

    def field(self, k, f):
        return np.full(self.nn[k], float(self.steps[k]))


class FakeAsyncPool(FakePool):
    """A pool whose quanta run asynchronously: ready() turns true on the second poll after a step."""

    def __init__(self, capacity, slots):
        super().__init__(capacity, slots)
        self.polls = 0

    def step(self, q, sync=True):
        super().step(q)
        self.polls = 2

    def ready(self):
        self.polls -= 1
        return self.polls <= 0


@pytest.fixture
def serving():
    from fastscapelib_fortran import serving
    return serving


def _runs(serving, spec):
    return [serving.Run(rid=i, nx=nx, ny=ny, steps=st, setup=lambda p, k: None, arrival=t)
            for i, (nx, ny, st, t) in enumerate(spec)]


def test_every_run_gets_exactly_its_steps(serving):
    pool = FakePool(capacity=10_000, slots=4)
    srv = serving.Server(pool, quantum=5)
    runs = _runs(serving, [(50, 50, 13, 0), (40, 40, 7, 0), (30, 30, 22, 0), (60, 60, 3, 0), (20, 20, 9, 0),
                           (70, 70, 11, 0)])
    rep = srv.serve(runs, realtime=False)
    assert rep["runs"] == 6 and rep["rejected"] == 0
    for r in runs:
        assert r.result["h"][0] == r.steps, r.rid  # stepped exactly `steps` times while resident
    assert rep["cell_steps"] == sum(r.nx * r.ny * r.steps for r in runs)
    assert all(q <= 5 for q, _ in pool.log)


def test_admission_respects_memory_and_backfills(serving):
    pool = FakePool(capacity=5000, slots=8)
    srv = serving.Server(pool, quantum=4)
    # 60x60 = 3600 fits only alone; 30x30 = 900 runs fit beside it
    runs = _runs(serving, [(60, 60, 8, 0), (60, 60, 8, 0), (30, 30, 4, 0)])
    srv.serve(runs, realtime=False)
    first = pool.log[0][1]
    assert first == sorted([runs[0].slot, runs[2].slot])  # the small run went ahead of the second big one
    assert runs[2].admitted < runs[1].admitted
    assert all(sum(pool.nn.get(k, 0) for k in ks) <= 5000 for _, ks in pool.log)
    assert runs[1].admitted >= runs[0].done


def test_never_fitting_run_is_rejected(serving):
    pool = FakePool(capacity=1000, slots=2)
    srv = serving.Server(pool, quantum=4)
    runs = _runs(serving, [(100, 100, 5, 0), (20, 20, 5, 0)])
    rep = srv.serve(runs, realtime=False)
    assert rep["rejected"] == 1 and rep["runs"] == 1


def test_latency_is_completion_minus_arrival(serving):
    t = [0.0]

    def clock():
        t[0] += 0.001
        return t[0]
    pool = FakePool(capacity=10_000, slots=4)
    srv = serving.Server(pool, quantum=3, clock=clock)
    runs = _runs(serving, [(10, 10, 6, 0), (10, 10, 9, 0)])
    rep = srv.serve(runs, realtime=False)
    for r in runs:
        assert r.done > r.admitted >= r.arrival and r.latency == r.done - r.arrival
    assert rep["p99_latency_s"] >= rep["p50_latency_s"] > 0


def test_lanes_deal_balances_cells_and_covers_every_landscape():
    """Lanes.deal (host logic only): largest first to the least-loaded lane, every id exactly once."""
    from fastscapelib_fortran import Lanes
    sizes = [(200, 200), (100, 100), (100, 100), (100, 100), (100, 100), (50, 50)]
    m = Lanes.deal(sizes, 2)
    assert sorted(k for lane in m for k in lane) == list(range(len(sizes)))
    loads = [sum(sizes[k][0] * sizes[k][1] for k in lane) for lane in m]
    assert max(loads) - min(loads) <= 50 * 50
    assert all(lane == sorted(lane) for lane in m)
    assert Lanes.deal(sizes, 10) == [[0], [1], [2], [3], [4], [5]]  # no empty lanes
    assert Lanes.deal([(513, 513)] * 64, 2) == [list(range(0, 64, 2)), list(range(1, 64, 2))]


def test_lanes_step_every_run_exactly_and_overlap(serving):
    """Two lanes: each run lives in one lane and gets exactly its steps, heavy runs in the first lane and
    the others in the second; both lanes have quanta in flight at once."""
    pools = [FakeAsyncPool(capacity=6000, slots=4), FakeAsyncPool(capacity=6000, slots=4)]
    srv = serving.Server(pools, quantum=4)
    runs = _runs(serving, [(50, 50, 13, 0), (40, 40, 7, 0), (30, 30, 22, 0), (60, 60, 3, 0), (20, 20, 9, 0),
                           (70, 70, 11, 0), (45, 45, 5, 0), (25, 25, 17, 0)])
    for r in runs:
        r.heavy = r.nx >= 50
    seen_both = []
    orig = srv.tick

    def tick():
        n = orig()
        seen_both.append(all(ln.q for ln in srv.lanes))
        return n
    srv.tick = tick
    rep = srv.serve(runs, realtime=False)
    assert rep["runs"] == len(runs) and rep["rejected"] == 0 and rep["lanes"] == 2
    for r in runs:
        assert r.result["h"][0] == r.steps, r.rid
    assert rep["cell_steps"] == sum(r.nx * r.ny * r.steps for r in runs)
    assert any(seen_both)  # quanta of both lanes in flight together
    for p in pools:
        assert p.log and all(q <= 4 for q, _ in p.log)
    heavy = {r.slot for r in runs if r.heavy}
    assert all(set(ks) <= heavy for _, ks in pools[0].log)  # only heavy runs in the first lane
