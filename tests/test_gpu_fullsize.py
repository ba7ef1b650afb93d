"""The benchmark configurations at their full sizes.

* The uniform batch (64 landscapes of 513x513, D8, n=1, seeded as bench.py seeds it): two members are
  stepped beside the other landscapes of the batch and compared with the oracle run alone (receivers
  and donors bit for bit, h within 1e-10, area within 1e-12); the whole batch, stepped as three lanes,
  is checked for properties that hold at any size -- every receiver chain ends at a fixed boundary
  node, and the drainage area held by the base-level nodes is twice the grid's area (the reference
  finally adds a base-level node's accumulated area to itself, src/FlowRouting.f90:136-139).
* The large landscape (4097x4097, multiple flow direction p=1, deposition g=1, kf=kfsed=5e-6,
  kd=kdsed=5e-2, uplift on the central half): one step, then the multiple-flow drainage area that
  reaches the sinks equals the grid's area (the flow weights of a node sum to one) and every field is
  finite.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

H_TOL = 1e-10
A_TOL = 1e-12


def _uniform(seed, n=513, dx=200.0):
    """One landscape of the uniform batch: h noise, then kf, then U, from PCG64(seed)."""
    rng = np.random.default_rng(seed)
    nn = n * n
    h = rng.random(nn)
    kf = 10.0 ** rng.uniform(np.log10(2e-6), np.log10(2e-5))
    U = rng.uniform(5e-4, 2e-3)
    u = np.full(nn, U)
    e = np.zeros((n, n), bool)
    e[0, :] = e[-1, :] = e[:, 0] = e[:, -1] = True
    u[e.ravel()] = 0.0
    return dict(nx=n, ny=n, xl=dx * (n - 1), yl=dx * (n - 1), h=h, u=u, kf=np.full(nn, kf), kfsed=-1.0, m=0.4,
                n=1.0, kd=np.full(nn, 1e-2), kdsed=-1.0, g1=0.0, g2=0.0, p=-2.0, ibc=1111, dt=1e3)


def _apply(obj, L, *k):
    obj.set_xl_yl(*k, L["xl"], L["yl"])
    obj.set_dt(*k, L["dt"])
    obj.init_h(*k, L["h"])
    obj.set_erosional_parameters(*k, L["kf"], L["kfsed"], L["m"], L["n"], L["kd"], L["kdsed"], L["g1"], L["g2"],
                                 L["p"])
    obj.set_u(*k, L["u"])
    obj.set_bc(*k, L["ibc"])


def test_uniform_batch_members_match_oracle(oracle):
    import fastscapelib_fortran as fs
    steps, members = 3, [0, 5]
    Ls = [_uniform(s) for s in range(8)]
    print(f"uniform batch: 8 x 513x513 in one context, {steps} steps, oracle members {members}", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            _apply(b, L, k)
        b.step(steps)
        got = {k: (b.rec(k), b.donors(k), b.h(k), b.field(k, 1), b.field(k, 6)) for k in members}
    for k in members:
        o = oracle.COracle(Ls[k]["nx"], Ls[k]["ny"])
        _apply(o, Ls[k])
        for _ in range(steps):
            o.execute_step()
        rec, (ndon, don), h, bb, area = got[k]
        assert np.array_equal(rec, o.ifield("rec")), k
        rnd, rdon = o.donors()
        assert np.array_equal(ndon, rnd) and np.array_equal(don, rdon), k
        assert rel_err(h, o.field("h")) <= H_TOL, k
        assert rel_err(bb, o.field("b")) <= H_TOL, k
        assert rel_err(area, o.field("area")) <= A_TOL, k


def test_uniform_batch_invariants_all_64():
    import fastscapelib_fortran as fs
    n, dx = 513, 200.0
    Ls = [_uniform(s) for s in range(64)]
    print("uniform batch: 64 x 513x513 as 3 lanes, 2 steps", flush=True)
    edge = np.zeros((n, n), bool)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    edge = edge.ravel()
    with fs.Lanes([(L["nx"], L["ny"]) for L in Ls], 3) as b:
        for k, L in enumerate(Ls):
            _apply(b, L, k)
        b.step(2)
        for k in range(64):
            rec = b.rec(k).astype(np.int64)
            area = b.field(k, 6)
            roots = np.flatnonzero(rec == np.arange(n * n))
            assert edge[roots].all(), k  # after LocalMinima only fixed boundary nodes drain nowhere
            r = rec.copy()
            for _ in range(40):  # pointer doubling: 2^40 hops
                r = r[r]
            assert np.array_equal(r[roots], roots) and edge[r].all(), k  # every chain ends at a boundary node
            total = 2.0 * dx * dx * n * n
            assert abs(area[roots].sum() - total) <= 1e-12 * total, k
            assert np.isfinite(b.h(k)).all(), k


def test_large_mfd_landscape_one_step():
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    n, dx = 4097, 100.0
    nn = n * n
    rng = np.random.default_rng(4097)
    h = rng.random(nn)
    u = np.zeros((n, n))
    u[n // 4:3 * n // 4, n // 4:3 * n // 4] = 2e-3
    print("large landscape: 4097x4097, multiple flow p=1, g=1, 1 step", flush=True)
    with fs.Batch([(n, n)]) as b:
        b.set_xl_yl(0, dx * (n - 1), dx * (n - 1))
        b.set_dt(0, 1e3)
        b.init_h(0, h)
        b.set_erosional_parameters(0, np.full(nn, 5e-6), 5e-6, 0.4, 1.0, np.full(nn, 5e-2), 5e-2, 1.0, 1.0, 1.0)
        b.set_u(0, u.ravel())
        b.set_bc(0, 1111)
        b.step(1)
        assert b.get_step(0) == 1
        area = b.field(0, 6)
        mnrec = b.ifield(0, _lib.FSX_MNREC)
        hh = b.h(0)
        assert np.isfinite(hh).all() and np.isfinite(area).all()
        sinks = mnrec == 0
        total = dx * dx * nn
        assert abs(area[sinks].sum() - total) <= 1e-9 * total
        assert area.min() >= dx * dx * (1.0 - 1e-12)
