"""GPU read-outs (fsx_copy_derived: slope, curvature, chi; fsx_get_fluxes) against the oracle.

Each case steps a batch of landscapes on the GPU and the same landscapes in the oracle, then checks

* the kernels, on the GPU's own state: the oracle's slope / curvature of the GPU's h, its chi over
  the GPU's receivers, order, area and lengths, its single-flow fluxes over the GPU's receivers,
  order, u and erosion rate.  Slope, curvature and chi within 1e-12 relative to the field's largest
  value (the GPU's atan / pow may differ from glibc's in the last place, chi sums them along a
  path); the catchment exactly, against the oracle's labels of the receivers that the last step
  found before closed-basin rerouting (find_receiver on the h it started from plus u dt); fluxes within 1e-12 + 2 nn eps of the sum of the magnitudes of their terms (the GPU adds in
  another order -- a tree reduction, and donors' shares by atomics -- and a sum of nn terms in index
  order is itself only good to (nn-1) eps of that magnitude);
* end to end, against the oracle's own state (which the GPU's matches to 1e-10 in h): slope,
  curvature and chi within 1e-7 of the field's largest value, fluxes within 1e-8 of the sum of the
  magnitudes of their terms (multiple flow is checked this way only);
* that chi, the catchment and the fluxes refuse a landscape that no step has routed;
* that the batched entry points (all landscapes in one launch) give the per-landscape results: bit for
  bit for the fields, to rounding for the fluxes (their atomics add in any order).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape

pytestmark = pytest.mark.gpu

CASES = [  # nx, ny, ibc, p, g
    (41, 37, 1111, -2.0, 0.0),
    (64, 45, 101, -2.0, 0.0),
    (57, 33, 1010, -2.0, 0.0),
    (50, 61, 1111, 1.0, 1.0),
    (47, 40, 1010, -1.0, 0.0),
]


def _sum_tol(nn):
    return 1e-12 + 2.0 * nn * np.finfo(np.float64).eps


def _close(got, want, tol):
    scale = max(float(np.abs(want).max()), 1e-300)
    return float(np.abs(np.asarray(got) - np.asarray(want)).max()) <= tol * scale


def test_readouts_match_oracle(oracle):
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    Ls = [landscape(nx, ny, seed=nx * ny, dx=250.0, ibc=ibc, p=p, g=g, kfsed=1e-5 if g else -1.0,
                    sinus=(5.0, 4000.0)) for nx, ny, ibc, p, g in CASES]
    steps = 3
    print(f"read-outs: {len(Ls)} landscapes, {steps} steps", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        b.step(steps - 1)
        h_prev = [b.h(k) for k in range(len(Ls))]
        b.step(1)
        for k, (L, (nx, ny, ibc, p, g)) in enumerate(zip(Ls, CASES)):
            dx, dy = L["xl"] / (nx - 1), L["yl"] / (ny - 1)
            o = oracle.COracle(nx, ny)
            apply(o, L)
            for _ in range(steps):
                o.execute_step()
            h = b.h(k)
            sl, cu, ch = b.slope(k), b.curvature(k), b.chi(k)
            # kernels on the GPU's own state
            assert _close(sl, oracle.slope(h, nx, ny, dx, dy), 1e-12), k
            assert _close(cu, oracle.curvature(h, nx, ny, dx, dy), 1e-12), k
            rec, order, area, length = b.rec(k), b.order(k), b.field(k, _lib.FSX_AREA), b.field(k, _lib.FSX_LENGTH)
            assert _close(ch, oracle.chi(order, rec, area, length, dx, dy, L["m"], L["n"]), 1e-12), k
            # catchment: the trees of the receivers the last step found before rerouting (on h + u dt)
            rec_pre, _ = oracle.find_receiver(h_prev[k] + L["u"] * L["dt"], nx, ny, dx, dy, ibc)
            assert np.array_equal(b.catchment(k), oracle.catchment(rec_pre)), k
            fl = b.fluxes(k)
            u, er = b.field(k, _lib.FSX_U), b.field(k, _lib.FSX_ERATE)
            surf = L["xl"] * L["yl"] / (nx - 1) / (ny - 1)
            mag = (np.abs(u).sum() * surf, np.abs(er).sum() * surf, 2.0 * np.abs(er).sum() * surf)
            if p < -1.5:
                _, bc = oracle.bc_decode(ibc, nx, ny)
                want = oracle.fluxes_sfd(order, rec, u, er, bc, surf)
                for q in range(3):
                    assert abs(fl[q] - want[q]) <= _sum_tol(nx * ny) * mag[q], (k, q, fl, want)
            # end to end against the oracle's own state
            assert _close(sl, o.derived("slope"), 1e-7), k
            assert _close(cu, o.derived("curvature"), 1e-7), k
            assert _close(ch, o.derived("chi"), 1e-7), k
            want = o.fluxes()
            for q in range(3):
                assert abs(fl[q] - want[q]) <= 1e-8 * mag[q], (k, q, fl, want)
        # the batched entry points
        for which in [_lib.FSX_SLOPE, _lib.FSX_CURVATURE, _lib.FSX_CHI, _lib.FSX_CATCHMENT]:
            outs = b.derived_all(which)
            for k in range(len(Ls)):
                assert np.array_equal(outs[k], b.derived(k, which)), (which, k)
        part = b.derived_all(_lib.FSX_CHI, [3, 1])
        assert np.array_equal(part[0], b.chi(3)) and np.array_equal(part[1], b.chi(1))
        fa = b.fluxes_all()
        for k, (nx, ny, *_r) in enumerate(CASES):
            one = b.fluxes(k)
            assert all(abs(fa[k, q] - one[q]) <= _sum_tol(nx * ny) * max(abs(one[q]), 1.0) for q in range(3)), k


def test_readouts_need_a_routed_step():
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    L = landscape(20, 20, seed=1)
    with fs.Batch([(20, 20)]) as b:
        apply(b, L, 0)
        assert np.isfinite(b.slope(0)).all()  # slope and curvature need only h
        with pytest.raises(fs.FastscapeError):
            b.chi(0)
        with pytest.raises(fs.FastscapeError):
            b.fluxes(0)
        with pytest.raises(fs.FastscapeError):
            b.catchment(0)
        with pytest.raises(fs.FastscapeError):
            b.derived_all(_lib.FSX_CHI)
        assert np.isnan(b.fluxes_all()).all()
        b.step(1)
        assert np.isfinite(b.chi(0)).all() and all(np.isfinite(b.fluxes(0)))


def test_readouts_uniform_batch_member_full_size(oracle):
    """A 513x513 member of the uniform batch after two steps, on the GPU's state (as above)."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    from test_gpu_fullsize import _apply, _uniform
    Ls = [_uniform(s) for s in range(4)]
    print("read-outs: 4 x 513x513, 2 steps", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            _apply(b, L, k)
        b.step(2)
        k, L, n = 3, Ls[3], 513
        dx = L["xl"] / (n - 1)
        h, rec, order = b.h(k), b.rec(k), b.order(k)
        area, length = b.field(k, _lib.FSX_AREA), b.field(k, _lib.FSX_LENGTH)
        assert _close(b.slope(k), oracle.slope(h, n, n, dx, dx), 1e-12)
        assert _close(b.curvature(k), oracle.curvature(h, n, n, dx, dx), 1e-12)
        assert _close(b.chi(k), oracle.chi(order, rec, area, length, dx, dx, L["m"], L["n"]), 1e-12)
        u, er = b.field(k, _lib.FSX_U), b.field(k, _lib.FSX_ERATE)
        _, bc = oracle.bc_decode(L["ibc"], n, n)
        surf = dx * dx
        want = oracle.fluxes_sfd(order, rec, u, er, bc, surf)
        got = b.fluxes(k)
        mag = (np.abs(u).sum() * surf, np.abs(er).sum() * surf, 2.0 * np.abs(er).sum() * surf)
        for q in range(3):
            assert abs(got[q] - want[q]) <= _sum_tol(n * n) * mag[q], (q, got, want)


def test_vtk_snapshot_of_a_batch_member(tmp_path):
    """Batch.vtk writes the member's own fields, grid spacing and step into the reference's file layout."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    from fastscapelib_fortran.vtk import write_vtk
    Ls = [landscape(30, 20, seed=1), landscape(41, 27, seed=2, dx=150.0)]
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        b.step(3)
        f = b.field(1, _lib.FSX_ERATE)
        paths = b.vtk(1, f, -2.0, directory=str(tmp_path / "gpu"))
        want = write_vtk(3, 41, 27, Ls[1]["xl"], Ls[1]["yl"], b.h(1), b.field(1, _lib.FSX_B), f, -2.0,
                         directory=str(tmp_path / "want"))
    assert [p.rsplit("/", 1)[1] for p in paths] == ["Topography0000003.vtk", "Basement0000003.vtk",
                                                    "SeaLevel0000003.vtk"]
    for p, q in zip(paths, want):
        assert open(p, "rb").read() == open(q, "rb").read()


def test_batched_readouts_over_lanes():
    """Lanes.derived_all / fluxes_all give each landscape's own read-outs, in landscape order."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    Ls = [landscape(30 + 3 * k, 25 + k, seed=k) for k in range(5)]
    with fs.Lanes([(L["nx"], L["ny"]) for L in Ls], 2) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        b.step(2)
        chi = b.derived_all(_lib.FSX_CHI, [4, 0, 3])
        assert all(np.array_equal(c, b.chi(k)) for c, k in zip(chi, [4, 0, 3]))
        fa = b.fluxes_all()
        for k, L in enumerate(Ls):
            one = b.fluxes(k)
            assert all(abs(fa[k, q] - one[q]) <= _sum_tol(L["nx"] * L["ny"]) * max(abs(one[q]), 1.0) for q in range(3))


def test_pool_readouts_with_free_slots_and_relayout():
    """A pool with free slots and a removed landscape: the batched read-outs address landscapes by slot id
    (one output entry per slot on the library side), and a landscape whose level bounds moved with the
    re-layout refuses chi and fluxes until it is stepped again, then gives the values of a fresh step."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    Ls = [landscape(30 + 4 * k, 22 + 3 * k, seed=40 + k, dx=300.0) for k in range(3)]
    with fs.Pool(40000, slots=16) as p:
        ids = [p.add(L["nx"], L["ny"]) for L in Ls]
        for k, L in zip(ids, Ls):
            apply(p, L, k)
        p.step(2)
        chi_before = {k: p.chi(k) for k in ids}
        fl_before = {k: p.fluxes(k) for k in ids}
        fa = p.fluxes_all()
        assert fa.shape == (3, 3)
        for row, k in zip(fa, sorted(ids)):
            assert np.allclose(row, fl_before[k], rtol=1e-12, atol=0.0)
        p.remove(ids[0])  # lowest node offset: the others' per-landscape offsets move
        rest = ids[1:]
        for k in rest:
            with pytest.raises(fs.FastscapeError):
                p.chi(k)
            with pytest.raises(fs.FastscapeError):
                p.fluxes(k)
        assert np.isnan(p.fluxes_all()).all()
        # the same landscapes set up again and stepped twice in a fresh pool give the reference values
        p.step(0)
        with fs.Pool(40000, slots=16) as q:
            qid = [q.add(L["nx"], L["ny"]) for L in Ls[1:]]
            for k, L in zip(qid, Ls[1:]):
                apply(q, L, k)
            q.step(2)
            want_chi = [q.chi(k) for k in qid]
            want_fl = [q.fluxes(k) for k in qid]
        # p's landscapes are now at step 2; step p's and q's copies once more and compare
        for k in rest:
            assert np.array_equal(chi_before[k], want_chi[rest.index(k)])
            assert np.allclose(fl_before[k], want_fl[rest.index(k)], rtol=1e-12, atol=0.0)
        p.step(1)
        outs = p.derived_all(_lib.FSX_CHI)
        assert len(outs) == len(rest)
        for o, k in zip(outs, sorted(rest)):
            assert np.array_equal(o, p.chi(k))
        fa = p.fluxes_all()
        assert fa.shape == (len(rest), 3) and np.isfinite(fa).all()
