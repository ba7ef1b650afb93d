"""Every benchmark configuration at its own sizes and step counts, against the reference Fortran.

The fixtures (tests/golden/configs/*.npz, written by tests/golden/make_configs.py from the reference
built from its own sources; the members with Newton tolerance 1e-8 from the build with that
tolerance) hold sampled members of each configuration after its full step count: h, b, etot and
drainage area at a seeded sample of nodes, and SHA-256 digests of the receivers and donor counts of
the last step.  Here the whole configuration is stepped on the GPU, the sampled members are compared
with their fixtures (receivers and donor counts bit for bit, h and b within 1e-10 relative, area
within 1e-12 relative -- 1e-8 for multiple-flow runs of hundreds of steps, see the serving test --,
etot within 1e-10 of the elevation scale), and every member is held to the
invariants of a routed landscape (only fixed boundary nodes drain nowhere, every receiver chain ends on
one, the base-level nodes collect twice the grid's area).
"""
from __future__ import annotations

import glob
import hashlib
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CFG = os.path.join(ROOT, "tests", "golden", "configs")
H_TOL = 1e-10
A_TOL = 1e-12


def _fixtures(config):
    out = {}
    for p in sorted(glob.glob(os.path.join(CFG, f"{config}_*.npz"))):
        g = dict(np.load(p, allow_pickle=False))
        out[int(g["member"])] = g
    return out


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def _check_member(name, g, h, b, etot, area, rec, ndon, area_tol=A_TOL):
    idx = g["idx"]
    bad = np.flatnonzero(rec[idx] != g["rec_s"])
    assert bad.size == 0, (name, "receivers differ at sampled nodes", idx[bad[:5]])
    assert _digest(rec) == str(g["rec_sha"]), (name, "receivers")
    assert _digest(ndon) == str(g["ndon_sha"]), (name, "donor counts")
    for k, got, tol in (("h", h, H_TOL), ("b", b, H_TOL), ("area", area, area_tol)):
        want = g[k]
        err = float(np.max(np.abs(got[idx] - want) / np.maximum(np.abs(want), 1e-300)))
        assert err <= tol, (name, k, err)
    scale = max(1.0, float(np.abs(g["h"]).max()))
    assert float(np.max(np.abs(etot[idx] - g["etot"]))) <= H_TOL * scale * max(1.0, float(np.abs(g["etot"]).max())), \
        (name, "etot")


def _fetch(b, k):
    from fastscapelib_fortran import _lib
    ndon, _ = b.donors(k)
    return (b.h(k), b.field(k, _lib.FSX_B), b.field(k, _lib.FSX_ETOT), b.field(k, _lib.FSX_AREA), b.rec(k), ndon)


def _invariants(name, L, rec, area):
    from fastscapelib_fortran.workloads import fixed_mask
    nx, ny = L["nx"], L["ny"]
    nn = nx * ny
    fixed = fixed_mask(nx, ny, L["ibc"])
    rec = rec.astype(np.int64)
    roots = np.flatnonzero(rec == np.arange(nn))
    assert fixed[roots].all(), (name, "a closed basin is left")
    r = rec.copy()
    for _ in range(int(np.ceil(np.log2(nn))) + 1):  # pointer doubling
        r = r[r]
    assert fixed[r].all() and np.array_equal(r[roots], roots), (name, "a receiver chain does not end on the boundary")
    dx, dy = L["xl"] / (nx - 1), L["yl"] / (ny - 1)
    total = 2.0 * dx * dy * nn
    assert abs(area[roots].sum() - total) <= 1e-11 * total, (name, "area")


def test_uniform_batch_500_steps():
    """64 x 513 x 513, 500 steps, as two lanes (the benchmark's layout)."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import workloads as W
    fx = _fixtures("uniform")
    assert fx, "missing fixtures"
    Ls = W.uniform_batch()
    print(f"uniform batch: {len(Ls)} landscapes, {W.UNIFORM_STEPS} steps, reference members {sorted(fx)}", flush=True)
    with fs.Lanes([(L["nx"], L["ny"]) for L in Ls], 2) as b:
        for k, L in enumerate(Ls):
            W.configure(b, L, k)
        b.step(W.UNIFORM_STEPS)
        for k, L in enumerate(Ls):
            h, bb, et, area, rec, ndon = _fetch(b, k)
            assert np.isfinite(h).all(), k
            _invariants(f"uniform {k}", L, rec, area)
            if k in fx:
                _check_member(f"uniform {k}", fx[k], h, bb, et, area, rec, ndon)


def test_ragged_batch_300_steps():
    """200 ragged landscapes (n from {1, 1.5, 2}, Newton tolerance 1e-8, BC 1010 / 1111), 300 steps,
    in one context with the launch shapes chosen by the autotuner."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import workloads as W
    fx = _fixtures("ragged")
    assert fx, "missing fixtures"
    Ls = W.ragged_batch()
    print(f"ragged batch: {len(Ls)} landscapes, {W.RAGGED_STEPS} steps, reference members {sorted(fx)}", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            W.configure(b, L, k)
        b.step(2)
        b.autotune()
        b.step(W.RAGGED_STEPS - 2)
        for k, L in enumerate(Ls):
            h, bb, et, area, rec, ndon = _fetch(b, k)
            assert np.isfinite(h).all(), k
            _invariants(f"ragged {k}", L, rec, area)
            if k in fx:
                _check_member(f"ragged {k} ({L['nx']}x{L['ny']} n={L['n']} ibc={L['ibc']})", fx[k], h, bb, et, area,
                              rec, ndon)


def test_mixed_batch():
    """The 4097 x 4097 multiple-flow landscape with deposition and 64 D8 landscapes of 257 x 257 in one
    context: all step together for 3 steps (the large landscape is compared with the reference there),
    then the large one is set inactive and the small ones complete their 200 steps in the same launches."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import workloads as W
    fl = _fixtures("mixed_large")
    fs_ = _fixtures("mixed_small")
    assert fl and fs_, "missing fixtures"
    Ls = W.mixed_batch()
    pre = int(fl[0]["steps"])
    print(f"mixed batch: 4097^2 MFD g=1 + 64 x 257^2 D8, {pre} steps together, then {W.MIXED_STEPS} for the small",
          flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            W.configure(b, L, k)
        b.step(pre)
        h, bb, et, area, rec, ndon = _fetch(b, 0)
        _check_member("mixed large", fl[0], h, bb, et, area, rec, ndon)
        # multiple flow: base-level nodes receive the whole grid's area once (no self-addition)
        from fastscapelib_fortran.workloads import fixed_mask
        L0 = Ls[0]
        dx = L0["xl"] / (L0["nx"] - 1)
        total = dx * dx * L0["nx"] * L0["ny"]
        assert abs(area[fixed_mask(L0["nx"], L0["ny"], 1111)].sum() - total) <= 1e-9 * total
        print("mixed batch: large landscape checked, small ones continue", flush=True)
        b.set_active(0, False)
        b.step(W.MIXED_STEPS - pre)
        for k in range(1, len(Ls)):
            h, bb, et, area, rec, ndon = _fetch(b, k)
            _invariants(f"mixed small {k - 1}", Ls[k], rec, area)
            if k - 1 in fs_:
                _check_member(f"mixed small {k - 1}", fs_[k - 1], h, bb, et, area, rec, ndon)
        assert b.get_step(0) == pre and b.get_step(1) == W.MIXED_STEPS


def test_served_runs_match_reference():
    """Runs drawn from the serving stream's distributions (grids to 2049, D8 and multiple flow, n from
    {1, 1.5, 2}, Newton tolerance 1e-8) served by the continuous-batching loop together with other runs
    of the stream, each result compared with the reference run alone."""
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import serving
    from fastscapelib_fortran import workloads as W
    fx = _fixtures("serving")
    assert fx, "missing fixtures"
    others = [i for i in range(40) if i not in fx and W.serving_run(i)[1]["nx"] * W.serving_run(i)[1]["ny"]
              * W.serving_run(i)[0] <= 2e7][:12]
    ids = sorted(fx) + others
    runs = []
    for i in ids:
        steps, L = W.serving_run(i)
        runs.append(serving.Run(rid=i, nx=L["nx"], ny=L["ny"], steps=steps, arrival=0.0,
                                setup=(lambda p, k, L=L: W.configure(p, L, k)),
                                outputs=("h", "b", "etot", "area", "rec", "ndon")))
    print(f"serving: {len(runs)} runs, reference runs {sorted(fx)}", flush=True)
    need = max(sum(r.cells for r in runs) // 2, max(r.cells for r in runs))
    with fs.Pool(need, slots=64) as p:
        rep = serving.Server(p, quantum=16).serve(runs, realtime=False)
    assert rep["runs"] == len(runs) and rep["rejected"] == 0
    by = {r.rid: r for r in runs}
    for i, g in fx.items():
        r = by[i].result
        mfd = W.serving_run(i)[1]["p"] > 0
        # multiple flow: the area sums donors in a fixed direction order where the reference follows its
        # mstack, so h agrees to ~1e-13 rather than bitwise; the flow weights are slope ratios between
        # neighbours that may be nearly level, so over hundreds of steps the area drifts to ~1e-9 while h
        # stays within 1e-10 (on identical h the area agrees to 1e-12: test_gpu_parity's multiple-flow tests)
        _check_member(f"served run {i} ({'MFD' if mfd else 'D8'})", g, r["h"], r["b"], r["etot"], r["area"], r["rec"],
                      r["ndon"], area_tol=1e-8 if mfd else A_TOL)
