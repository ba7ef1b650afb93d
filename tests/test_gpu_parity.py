"""Parity of the HIP path (through the C ABI) with the oracle on seeded landscapes.

Bar: receivers, donors and the edge-sort permutation bit-exact; elevation and the other FP64 fields
within 1e-10 relative (they are bit-identical unless libm pow and the device pow round differently);
drainage area within 1e-12 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape, rel_err

pytestmark = pytest.mark.gpu

H_TOL = 1e-10
A_TOL = 1e-12


@pytest.fixture(scope="module")
def fs():
    import fastscapelib_fortran as fs
    return fs


from fastscapelib_fortran import _lib  # noqa: E402


def _tie_heavy(rng, n, kind):
    if kind == 0:
        return rng.integers(0, 5, n).astype(float)
    if kind == 1:
        return rng.random(n)
    if kind == 2:
        return rng.integers(0, max(1, n // 3), n).astype(float)
    a = rng.integers(0, 50, n).astype(float)
    a[rng.random(n) < 0.2] = -1e10
    return a


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 63, 64, 65, 700, 2047, 2048, 2049, 5000, 40000, 150000])
def test_indexx_matches_reference_sort(fs, oracle, n):
    rng = np.random.default_rng(n)
    for kind in range(4):
        a = _tie_heavy(rng, n, kind)
        got = fs.indexx(a)
        ref = oracle.indexx(a)
        assert np.array_equal(got, ref), f"n={n} kind={kind}: first mismatch at {np.argmax(got != ref)}"


def _route_gpu(fs, L):
    with fs.Batch([(L["nx"], L["ny"])]) as b:
        apply(b, L, 0)
        b.route()
        rec = b.rec(0)
        ndon, don = b.donors(0)
        length = b.field(0, _lib.FSX_LENGTH)
        area = b.field(0, _lib.FSX_AREA)
        hw = b.field(0, _lib.FSX_HWATER)
        order = b.order(0)
    return rec, length, ndon, don, area, hw, order


@pytest.mark.parametrize("ibc", [1111, 1010, 101, 1000, 11, 1100])
@pytest.mark.parametrize("shape", [(101, 101), (57, 33)])
def test_routing_matches_oracle(fs, oracle, ibc, shape):
    nx, ny = shape
    L = landscape(nx, ny, seed=ibc + nx, sinus=(5.0, 8000.0), ibc=ibc)
    rec, length, ndon, don, area, hw, order = _route_gpu(fs, L)
    dx = L["xl"] / (nx - 1)
    dy = L["yl"] / (ny - 1)
    r_rec, r_len, r_ndon, r_don, r_stack, _ = oracle.local_minima(L["h"], nx, ny, dx, dy, ibc)
    assert np.array_equal(rec, r_rec), f"receivers differ at {np.flatnonzero(rec != r_rec)[:10]}"
    assert np.array_equal(length, r_len)
    assert np.array_equal(ndon, r_ndon)
    assert np.array_equal(don, r_don)
    # order: a permutation in which every receiver precedes its donors
    pos = np.empty(nx * ny, np.int64)
    pos[order] = np.arange(nx * ny)
    assert np.array_equal(np.sort(order), np.arange(nx * ny))
    assert np.all(pos[rec] <= pos)
    # lake surface and area, as the reference computes them along its stack
    h = L["h"]
    r_hw = h.copy()
    for v in r_stack:
        r = r_rec[v]
        if r_hw[r] > r_hw[v]:
            r_hw[v] = r_hw[r] + 1e-8
    assert np.array_equal(hw, r_hw)
    a = np.full(nx * ny, dx * dy)
    for v in r_stack[::-1]:
        a[r_rec[v]] += a[v]
    assert rel_err(area, a) <= A_TOL


def _bowl(n, seed, depth, rad, noise):
    """A closed bowl (Gaussian depression of the given depth and radius, in grid fractions) on a
    plateau of uniform noise: the bowl is one large basin ringed by many small pits."""
    rng = np.random.default_rng(seed)
    x = np.arange(n) - (n - 1) / 2
    r2 = (x[None, :] ** 2 + x[:, None] ** 2) / (rad * n) ** 2
    return (100.0 - depth * np.exp(-r2) + noise * rng.random((n, n))).ravel()


@pytest.mark.parametrize("n,depth,rad,noise,min_edges", [(65, 20.0, 0.3, 0.5, 100), (97, 30.0, 0.25, 1.0, 20),
                                                         (161, 30.0, 0.3, 0.5, 30), (257, 60.0, 0.3, 0.3, 100)])
def test_routing_basin_with_many_neighbours(fs, oracle, n, depth, rad, noise, min_edges):
    """One closed basin borders dozens of others, so its pass table outgrows the walk's per-thread
    table and is found through the hash index; at 161 and 257 nodes a side the bowl holds more than
    2048 nodes and is searched node-parallel (preorder ranks, hash table by atomics): receivers,
    lengths and donors stay bit-exact."""
    L = landscape(n, n, seed=5, ibc=1111)
    L["h"] = _bowl(n, 5, depth, rad, noise)
    dx = L["xl"] / (n - 1)
    _, cb, cn, _ = oracle.connect_basins(L["h"], n, n, dx, dx, 1111)
    real = cn[:, 0] >= 0
    assert np.bincount(cb[real, 0]).max() >= min_edges  # the case the test is for
    if n >= 161:  # the bowl is a basin of more than 2048 nodes (roots of the steepest-descent forest)
        root, _ = oracle.find_receiver(L["h"], n, n, dx, dx, 1111)
        for _ in range(64):
            root = root[root]
        assert np.bincount(root).max() > 2048
    rec, length, ndon, don, area, hw, order = _route_gpu(fs, L)
    r_rec, r_len, r_ndon, r_don, _, _ = oracle.local_minima(L["h"], n, n, dx, dx, 1111)
    assert np.array_equal(rec, r_rec), f"receivers differ at {np.flatnonzero(rec != r_rec)[:10]}"
    assert np.array_equal(length, r_len)
    assert np.array_equal(ndon, r_ndon) and np.array_equal(don, r_don)
    got, ref = _run_both(fs, oracle, L, 5)
    _check(got, ref, 5)


def _run_both(fs, oracle, L, steps):
    o = oracle.COracle(L["nx"], L["ny"])
    apply(o, L)
    with fs.Batch([(L["nx"], L["ny"])]) as b:
        apply(b, L, 0)
        for _ in range(steps):
            o.execute_step()
        b.step(steps)
        got = {k: b.field(0, f) for k, f in [("h", 0), ("b", 1), ("etot", 4), ("erate", 5), ("area", 6),
                                             ("lake_depth", 7)]}
        got["rec"] = b.rec(0)
        got["ndon"], got["don"] = b.donors(0)
        got["step"] = b.get_step(0)
    ref = {k: o.field(k) for k in ["h", "b", "etot", "erate", "area", "lake_depth"]}
    ref["rec"] = o.ifield("rec")
    ref["ndon"], ref["don"] = o.donors()
    return got, ref


def _check(got, ref, steps):
    assert got["step"] == steps
    assert np.array_equal(got["rec"], ref["rec"]), "receivers differ"
    assert np.array_equal(got["ndon"], ref["ndon"]) and np.array_equal(got["don"], ref["don"]), "donors differ"
    for k in ["h", "b"]:
        assert rel_err(got[k], ref[k]) <= H_TOL, k
    assert rel_err(got["area"], ref["area"]) <= A_TOL
    scale = max(1.0, float(np.abs(ref["h"]).max()))
    for k in ["etot", "erate", "lake_depth"]:
        assert np.max(np.abs(got[k] - ref[k])) <= H_TOL * scale * max(1.0, float(np.abs(ref[k]).max())), k


def test_plain_101_d8_200_steps(fs, oracle):
    """The reference's own CPU case: 101x101, dx=1 km, D8, kf=1e-5, m=0.4, n=1, kd=1e-2, 200 steps."""
    L = landscape(101, 101, seed=0)
    got, ref = _run_both(fs, oracle, L, 200)
    _check(got, ref, 200)


@pytest.mark.parametrize("n,ibc,g", [(1.0, 1010, 0.0), (1.5, 1111, 0.0), (2.0, 1010, 0.0), (1.0, 1111, 1.0),
                                     (1.5, 101, 0.5)])
def test_nonlinear_and_bc_variants(fs, oracle, n, ibc, g):
    L = landscape(73, 49, seed=int(n * 10) + ibc, dx=250.0, kf=2e-6, m=0.45 * n, n=n, ibc=ibc, dt=2e3, g=g,
                  sinus=(10.0, 6000.0), kd=3e-2)
    got, ref = _run_both(fs, oracle, L, 25)
    _check(got, ref, 25)


def test_deeper_than_lds_level_histogram(fs, oracle):
    """A 40001x3 strip fixed on its left edge only, rising to the right: receiver chains of ~40000
    nodes, so the level order holds more levels than the LDS histogram (32768) and the deep levels
    go through the global-memory counts."""
    nx, ny = 40001, 3
    L = landscape(nx, ny, seed=77, dx=100.0, ibc=1, kf=1e-6, kd=1e-3)
    L["h"] = (np.tile(np.arange(nx, dtype=np.float64), ny) * 0.5 + L["h"] * 0.1)
    L["u"] = np.full(nx * ny, 1e-4)
    L["u"][np.arange(nx * ny) % nx == 0] = 0.0
    print("deep strip 40001x3, ibc=0001, 2 steps", flush=True)
    got, ref = _run_both(fs, oracle, L, 2)
    _check(got, ref, 2)
    with fs.Batch([(nx, ny)]) as b:
        apply(b, L, 0)
        b.step(1)
        assert b.stats(0)["levels"] > 32768


@pytest.mark.parametrize("shape,ibc,deep", [((521, 261), 1010, False), ((50001, 3), 1, True)])
def test_large_landscape_level_order(fs, oracle, shape, ibc, deep):
    """Landscapes of 131072 nodes and more are level-sorted by one workgroup per 16384-node segment.
    521x261: a 2-D landscape of 9 segments; 50001x3: a strip whose ~50000 levels are deeper than the
    segments' LDS bins (8192) and the one-workgroup sort's LDS histogram (32768).  Elevations match the
    oracle, the order is a permutation in which every receiver precedes its donors, and a batch with a
    small landscape beside the large one gives the large one's result bit for bit."""
    nx, ny = shape
    if deep:
        L = landscape(nx, ny, seed=78, dx=100.0, ibc=ibc, kf=1e-6, kd=1e-3)
        L["h"] = (np.tile(np.arange(nx, dtype=np.float64), ny) * 0.5 + L["h"] * 0.1)
        L["u"] = np.full(nx * ny, 1e-4)
        L["u"][np.arange(nx * ny) % nx == 0] = 0.0
        steps = 2
    else:
        L = landscape(nx, ny, seed=5, dx=250.0, ibc=ibc, sinus=(10.0, 20000.0), kf=2e-6)
        steps = 4
    print(f"large landscape {nx}x{ny}, ibc={ibc}, {steps} steps", flush=True)
    got, ref = _run_both(fs, oracle, L, steps)
    _check(got, ref, steps)
    small = landscape(65, 90, seed=1, dx=250.0, ibc=1010, sinus=(10.0, 5000.0))
    with fs.Batch([(nx, ny), (65, 90)]) as b:
        apply(b, L, 0)
        apply(b, small, 1)
        b.step(steps)
        h_mix = b.h(0)
        rec, order = b.rec(0), b.order(0)
        levels = b.stats(0)["levels"]
    assert np.array_equal(h_mix, got["h"])
    pos = np.empty(nx * ny, np.int64)
    pos[order] = np.arange(nx * ny)
    assert np.array_equal(np.sort(order), np.arange(nx * ny))
    assert np.all(pos[rec] <= pos)
    if deep:
        assert levels > 32768


def test_batch_equals_alone(fs, oracle):
    """Ragged batch: each landscape's result equals its run alone, bit for bit."""
    Ls = [landscape(65, 90, seed=1, dx=250.0, ibc=1010, sinus=(10.0, 5000.0)),
          landscape(120, 70, seed=2, dx=250.0, n=1.5, m=0.675, kf=2e-6),
          landscape(40, 40, seed=3)]
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        b.step(15)
        mix = [b.h(k) for k in range(len(Ls))]
        mixrec = [b.rec(k) for k in range(len(Ls))]
    for k, L in enumerate(Ls):
        with fs.Batch([(L["nx"], L["ny"])]) as b1:
            apply(b1, L, 0)
            b1.step(15)
            assert np.array_equal(b1.h(0), mix[k])
            assert np.array_equal(b1.rec(0), mixrec[k])


def test_f2py_surface_roundtrip(fs, oracle):
    """The fastscape_* free functions drive the same step as the reference's module."""
    L = landscape(61, 45, seed=7)
    fs.fastscape_init()
    fs.fastscape_set_nx_ny(61, 45)
    fs.fastscape_setup()
    fs.fastscape_set_xl_yl(L["xl"], L["yl"])
    fs.fastscape_set_dt(L["dt"])
    fs.fastscape_init_h(L["h"])
    fs.fastscape_set_erosional_parameters(L["kf"], -1.0, 0.4, 1.0, L["kd"], -1.0, 0.0, 0.0, -2.0)
    fs.fastscape_set_u(L["u"])
    fs.fastscape_set_bc(1111)
    for _ in range(10):
        fs.fastscape_execute_step()
    h = np.empty(61 * 45)
    fs.fastscape_copy_h(h)
    assert fs.fastscape_get_step() == 10
    assert fs.fastscape_get_sizes() == (61, 45)
    fs.fastscape_destroy()
    o = oracle.COracle(61, 45)
    apply(o, L)
    for _ in range(10):
        o.execute_step()
    assert rel_err(h, o.field("h")) <= H_TOL


# ---- multiple flow direction -------------------------------------------------------------------
def _mfd_case(nx, ny, seed, ibc, p, g=0.0, n=1.0):
    return landscape(nx, ny, seed=seed, dx=250.0, kf=2e-6, m=0.45 * n, n=n, ibc=ibc, dt=2e3, g=g,
                     sinus=(10.0, 4000.0), kd=3e-2, p=p, kfsed=(1e-5 if g else -1.0), kdsed=(2e-2 if g else -1.0))


@pytest.mark.parametrize("ibc", [1111, 1010, 101, 1000, 11])
@pytest.mark.parametrize("p", [1.0, -1.0])
def test_mfd_routing_matches_oracle(fs, oracle, ibc, p):
    """Multiple-flow receivers bit-exact (slot order included); the order puts every node before
    its receivers and its single-flow receiver; drainage area within 1e-12."""
    L = _mfd_case(41, 33, seed=ibc, ibc=ibc, p=p)
    print(f"mfd routing ibc={ibc} p={p}", flush=True)
    o = oracle.COracle(41, 33)
    apply(o, L)
    o.execute_step()  # routing inside the oracle's step runs on the post-uplift h
    with fs.Batch([(41, 33)]) as b:
        apply(b, L, 0)
        b.step(1)
        mn, mr = b.mrec(0)
        order = b.morder(0)
        area = b.field(0, _lib.FSX_AREA)
        rec = b.rec(0)
        hw = b.field(0, _lib.FSX_HWATER)
    omr, _, _ = o.mrec()
    omn = o.ifield("mnrec")
    assert np.array_equal(mn, omn)
    for v in range(mn.size):
        assert np.array_equal(mr[v, :mn[v]], omr[v, :omn[v]]), v
    assert np.all(mr[np.arange(8)[None, :] >= mn[:, None]] == -1)
    assert np.array_equal(hw, o.field("hwater"))
    assert rel_err(area, o.field("area")) <= A_TOL
    pos = np.empty(mn.size, np.int64)
    pos[order] = np.arange(mn.size)
    assert np.array_equal(np.sort(order), np.arange(mn.size))
    for v in range(mn.size):
        for t in mr[v, :mn[v]]:
            assert pos[t] > pos[v]
        if rec[v] != v:
            assert pos[rec[v]] > pos[v]


@pytest.mark.parametrize("ibc,p,g,n", [(1111, 1.0, 0.0, 1.0), (1010, 1.0, 1.0, 1.0), (1111, -1.0, 0.0, 1.5),
                                       (101, 1.0, 0.5, 2.0), (1000, 0.5, 0.0, 1.0)])
def test_mfd_steps_match_oracle(fs, oracle, ibc, p, g, n):
    """15 steps.  Without deposition h is identical up to libm pow rounding and area is held to
    1e-12.  With deposition (g > 0) the erosion push and the lake volumes are sums over donors in a
    different order, so h agrees to ~1e-13; multiple-flow weights are slopes between filled
    heights only 1e-8 apart inside lakes, which turns that into ~1e-12 differences of the next
    step's weights and areas: there the area bar is 1e-10 (the one-step area, on identical h, is
    held to 1e-12 by test_mfd_routing_matches_oracle)."""
    L = _mfd_case(53, 37, seed=7 + ibc, ibc=ibc, p=p, g=g, n=n)
    print(f"mfd steps ibc={ibc} p={p} g={g} n={n}", flush=True)
    got, ref = _run_both(fs, oracle, L, 15)
    assert got["step"] == 15
    assert np.array_equal(got["rec"], ref["rec"]), "receivers differ"
    for k in ["h", "b"]:
        assert rel_err(got[k], ref[k]) <= H_TOL, k
    assert rel_err(got["area"], ref["area"]) <= (A_TOL if g == 0 else 1e-10)
    scale = max(1.0, float(np.abs(ref["h"]).max()))
    for k in ["etot", "erate", "lake_depth"]:
        assert np.max(np.abs(got[k] - ref[k])) <= H_TOL * scale * max(1.0, float(np.abs(ref[k]).max())), k


def test_mfd_and_sfd_mixed_batch_equal_alone(fs):
    """Single- and multiple-flow landscapes in one batch: each equals its run alone, bit for bit
    without deposition; with deposition (g > 0) the lake volumes are sums whose order follows the
    device's atomic adds, so there the bar is 1e-12 relative."""
    Ls = [_mfd_case(45, 30, seed=11, ibc=1111, p=1.0, g=1.0), landscape(50, 41, seed=12),
          _mfd_case(33, 52, seed=13, ibc=1010, p=-1.0, n=1.5)]
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        b.step(10)
        mix = [b.h(k) for k in range(len(Ls))]
    for k, L in enumerate(Ls):
        with fs.Batch([(L["nx"], L["ny"])]) as b1:
            apply(b1, L, 0)
            b1.step(10)
            if L["g1"] > 0:
                assert rel_err(b1.h(0), mix[k]) <= 1e-12, k
            else:
                assert np.array_equal(b1.h(0), mix[k]), k
