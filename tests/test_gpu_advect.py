"""Advection on the GPU (fsx_set_v; Advect, src/Advect.f90:1-103) against the oracle.

* Advection alone (no erosion, no uplift) in a batch where some landscapes advect and others do not:
  h, b and etot bit for bit (the row and column solves keep tridag's operations and order), the
  others untouched; with captured graphs and with kernels launched one by one.
* The whole step with advection first, single and multiple flow: receivers bit for bit, h and b
  within 1e-10, drainage area within 1e-12 (as the other parity tests).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape, rel_err

pytestmark = pytest.mark.gpu


def _vel(nx, ny, seed):
    rng = np.random.default_rng(seed)
    vx = (rng.random(nx * ny) - 0.3) * 0.02
    vy = (rng.random(nx * ny) - 0.6) * 0.015
    vx[::7] = 0.0
    return vx, vy


@pytest.mark.parametrize("graphs", [True, False])
def test_advection_alone_bitwise(oracle, graphs):
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    sizes = [(37, 31), (64, 20), (15, 70), (33, 33)]
    adv = [True, False, True, True]
    rng = np.random.default_rng(11)
    hs = [rng.random(nx * ny) * 10.0 for nx, ny in sizes]
    print(f"advection alone, graphs={graphs}", flush=True)
    with fs.Batch(sizes) as b:
        b.graphs(graphs)
        for k, (nx, ny) in enumerate(sizes):
            b.set_xl_yl(k, 100.0 * (nx - 1), 90.0 * (ny - 1))
            b.set_dt(k, 1e3)
            b.init_h(k, hs[k])
            b.set_basement(k, hs[k] - 1.0)
            if adv[k]:
                b.set_v(k, *_vel(nx, ny, k))
        b.step(3)
        for k, (nx, ny) in enumerate(sizes):
            o = oracle.COracle(nx, ny)
            o.set_xl_yl(100.0 * (nx - 1), 90.0 * (ny - 1))
            o.set_dt(1e3)
            o.init_h(hs[k])
            o.set_basement(hs[k] - 1.0)
            if adv[k]:
                o.set_v(*_vel(nx, ny, k))
            for _ in range(3):
                o.execute_step()
            assert np.array_equal(b.h(k), o.field("h")), k
            assert np.array_equal(b.field(k, _lib.FSX_B), o.field("b")), k
            assert np.array_equal(b.field(k, _lib.FSX_ETOT), o.field("etot")), k
            if not adv[k]:
                assert np.array_equal(b.h(k), hs[k]), k


CASES = [(45, 38, 1111, -2.0), (52, 41, 1010, -2.0), (40, 47, 101, 1.0)]


def test_step_with_advection_matches_oracle(oracle):
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    Ls = [landscape(nx, ny, seed=nx + ny, dx=250.0, ibc=ibc, p=p, sinus=(5.0, 4000.0)) for nx, ny, ibc, p in CASES]
    steps = 3
    print(f"step with advection: {len(Ls)} landscapes, {steps} steps", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
            b.set_v(k, *_vel(L["nx"], L["ny"], 100 + k))
        b.step(steps)
        for k, L in enumerate(Ls):
            o = oracle.COracle(L["nx"], L["ny"])
            apply(o, L)
            o.set_v(*_vel(L["nx"], L["ny"], 100 + k))
            for _ in range(steps):
                o.execute_step()
            assert np.array_equal(b.rec(k), o.ifield("rec")), k
            assert rel_err(b.h(k), o.field("h")) <= 1e-10, k
            assert rel_err(b.field(k, _lib.FSX_B), o.field("b")) <= 1e-10, k
            assert rel_err(b.field(k, _lib.FSX_AREA), o.field("area")) <= 1e-12, k
