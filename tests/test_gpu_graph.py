"""Steps replayed from a captured HIP graph give the eager launches' results, replay after replay,
with the host's own work and synchronisations between replays (bitwise without deposition; 1e-12
relative with it, whose lake volumes are summed with device atomics)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape, rel_err

pytestmark = pytest.mark.gpu


def _batch(fs, Ls, graphs):
    b = fs.Batch([(L["nx"], L["ny"]) for L in Ls])
    for k, L in enumerate(Ls):
        apply(b, L, k)
    b.graphs(graphs)
    return b


def test_graph_replays_equal_eager_launches():
    import fastscapelib_fortran as fs
    Ls = [landscape(57, 44, seed=21), landscape(40, 61, seed=22, ibc=1010, n=1.5, m=0.675, kf=2e-6),
          landscape(45, 45, seed=23, p=1.0, g=1.0, kfsed=1e-5, kdsed=2e-2), landscape(33, 70, seed=24, p=-1.0)]
    print("graph vs eager, 4 landscapes, 6 replays", flush=True)
    with _batch(fs, Ls, True) as bg, _batch(fs, Ls, False) as be:
        for rep in range(6):
            bg.step(1)
            be.step(1)
            for k, L in enumerate(Ls):
                hg, he = bg.h(k), be.h(k)
                if L["g1"] > 0:
                    assert rel_err(hg, he) <= 1e-12, (rep, k)
                else:
                    assert np.array_equal(hg, he), (rep, k)
                assert np.array_equal(bg.field(k, 6), be.field(k, 6)) or L["g1"] > 0
            # host work between replays: a parameter change is picked up by the next replay
            if rep == 2:
                for b in (bg, be):
                    b.set_dt(0, 500.0)
        assert bg.get_step(0) == 6


def test_live_phase_timing_inside_graph_replays():
    """Splitting the step's graph around one directly launched, event-timed phase changes nothing."""
    import fastscapelib_fortran as fs
    Ls = [landscape(64, 48, seed=31), landscape(50, 50, seed=32, p=1.0)]
    print("graph split around lm_walk", flush=True)
    with _batch(fs, Ls, True) as bt, _batch(fs, Ls, False) as be:
        bt.time_phase("lm_walk")
        bt.reset_timers()
        for rep in range(4):
            bt.step(2)
            be.step(2)
            for k in range(len(Ls)):
                assert np.array_equal(bt.h(k), be.h(k)), (rep, k)
        ms, cnt = bt.timers()["lm_walk"]
        assert cnt == 8 and ms > 0.0
        bt.time_phase(None)
        bt.step(1)
        be.step(1)
        assert np.array_equal(bt.h(0), be.h(0))
