"""Continuous batching on the GPU: landscapes added to and removed from a pool between steps, and
runs served as they arrive, each give the same result as the run alone (bitwise without deposition,
1e-12 relative with it: lake volumes are summed with device atomics)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape, rel_err

pytestmark = pytest.mark.gpu


def _alone(fs, L, steps):
    with fs.Batch([(L["nx"], L["ny"])]) as b:
        apply(b, L, 0)
        b.step(steps)
        return b.h(0)


def _same(got, want, L):
    if L["g1"] > 0:
        assert rel_err(got, want) <= 1e-12
    else:
        assert np.array_equal(got, want)


def test_pool_add_remove_reuse():
    import fastscapelib_fortran as fs
    A = landscape(40, 30, seed=1)
    B = landscape(64, 50, seed=2, ibc=1010, sinus=(10.0, 5000.0))
    C = landscape(33, 45, seed=3, n=1.5, m=0.675, kf=2e-6)
    D = landscape(50, 39, seed=4, p=1.0)  # multiple flow direction, placed into B's freed range
    print("pool add/remove", flush=True)
    with fs.Pool(40 * 30 + 64 * 50 + 33 * 45, slots=4) as p:
        ka, kb, kc = p.add(40, 30), p.add(64, 50), p.add(33, 45)
        assert p.add(10, 10) is None  # arena full
        for k, L in [(ka, A), (kb, B), (kc, C)]:
            apply(p, L, k)
        p.step(5)
        hb = p.h(kb)
        p.remove(kb)
        kd = p.add(50, 39)
        assert kd is not None
        apply(p, D, kd)
        p.step(7)
        assert p.get_step(ka) == 12 and p.get_step(kd) == 7
        _same(p.h(ka), _alone(fs, A, 12), A)
        _same(p.h(kc), _alone(fs, C, 12), C)
        _same(p.h(kd), _alone(fs, D, 7), D)
    _same(hb, _alone(fs, B, 5), B)


def test_served_runs_equal_alone():
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import serving
    specs = [dict(nx=45, ny=38, seed=10, steps=12), dict(nx=70, ny=52, seed=11, steps=7, ibc=1010, n=2.0, m=0.9),
             dict(nx=33, ny=33, seed=12, steps=20, p=1.0), dict(nx=60, ny=41, seed=13, steps=9, g=1.0),
             dict(nx=52, ny=47, seed=14, steps=15, p=-1.0), dict(nx=38, ny=64, seed=15, steps=5, ibc=101)]
    Ls, runs = [], []
    for i, sp in enumerate(specs):
        sp = dict(sp)
        st = sp.pop("steps")
        L = landscape(sp.pop("nx"), sp.pop("ny"), seed=sp.pop("seed"), dx=250.0, kf=2e-6, sinus=(10.0, 4000.0),
                      **sp)
        Ls.append(L)
        runs.append(serving.Run(rid=i, nx=L["nx"], ny=L["ny"], steps=st,
                                setup=(lambda p, k, L=L: apply(p, L, k)), outputs=("h", "area")))
    cap = 2 * max(L["nx"] * L["ny"] for L in Ls)  # forces runs to wait for memory
    print(f"serving {len(runs)} runs, capacity {cap} nodes", flush=True)
    with fs.Pool(cap, slots=3) as p:
        srv = serving.Server(p, quantum=4)
        srv.submit(runs[0])
        srv.submit(runs[1])
        srv.tick()
        for r in runs[2:]:  # arrive while others are resident
            srv.submit(r)
            srv.tick()
        while srv.queue or srv.active:
            srv.tick()
        rep = srv.report(1.0)
    assert rep["runs"] == len(runs) and rep["rejected"] == 0
    for r, L in zip(runs, Ls):
        _same(r.result["h"], _alone(fs, L, r.steps), L)
