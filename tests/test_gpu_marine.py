"""Marine transport (src/Marine.f90:1-167, SiltSandCouplingDiffusion :171-499) on the GPU, through the C ABI.

* Every marine fixture's landscape (tests/golden/marine_*.npz: single and multiple flow, n = 1, 1.5, 2,
  with and without deposition, five boundary conditions, three sea levels, non-default porosities and
  an initial silt fraction) is stepped 20 times and held to the oracle, which tests/test_oracle.py
  pins bit for bit to the reference on the same runs: h, b within 1e-10 relative, etot, erate and
  Sedflux within 1e-10 of the elevation scale, the silt fraction within 1e-10.  The shelf flux sums
  a node's donors in direction order where the reference follows its stack, so results agree to
  rounding rather than bitwise.
* A marine landscape stepped in a batch beside landscapes without marine transport (single and
  multiple flow) equals the same landscape stepped alone, bit for bit, and leaves the others as they
  are without it.
* The f2py-style calls: fastscape_set_marine_parameters, fastscape_init_f / fastscape_copy_f (default
  silt fraction 0.5).
* Errors: marine transport without the erosional parameters is refused before the step; a silt / sand
  diffusion that does not converge within 1000 iterations (where the reference STOPs) fails the step.
"""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

from conftest import ROOT, apply, landscape, rel_err, shelf

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")
TOL = 1e-10
STEPS = 20


def _load(name):
    g = dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))
    L = {k[3:]: g[k] for k in g if k.startswith("in_")}
    for k in ["nx", "ny", "ibc"]:
        L[k] = int(L[k])
    for k in ["xl", "yl", "kfsed", "m", "n", "kdsed", "g1", "g2", "p", "dt"]:
        L[k] = float(L[k])
    return L


def _state(obj, k=None):
    from fastscapelib_fortran import _lib
    if k is None:
        return {f: obj.field(f) for f in ("h", "b", "etot", "erate", "sedflux", "fmix")}
    ids = {"h": _lib.FSX_H, "b": _lib.FSX_B, "etot": _lib.FSX_ETOT, "erate": _lib.FSX_ERATE,
           "sedflux": _lib.FSX_SEDFLUX, "fmix": _lib.FSX_FMIX}
    return {f: obj.field(k, i) for f, i in ids.items()}


def _close(name, s, got, want):
    assert rel_err(got["h"], want["h"]) <= TOL, (name, s, "h", rel_err(got["h"], want["h"]))
    assert rel_err(got["b"], want["b"]) <= TOL, (name, s, "b")
    scale = max(1.0, float(np.abs(want["h"]).max()))
    for k in ("etot", "erate", "sedflux"):
        err = float(np.max(np.abs(got[k] - want[k])))
        assert err <= TOL * scale * max(1.0, float(np.abs(want[k]).max())), (name, s, k, err)
    assert float(np.max(np.abs(got["fmix"] - want["fmix"]))) <= TOL, (name, s, "fmix")


MARINE = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "marine_*.npz")))


@pytest.mark.parametrize("name", MARINE)
def test_marine_runs_match_oracle(oracle, name):
    import fastscapelib_fortran as fs
    L = _load(name)
    print(f"marine {name}: {L['nx']}x{L['ny']} ibc={L['ibc']} p={L['p']} n={L['n']} g={L['g1']} "
          f"sea level={L['marine'][0]}, {STEPS} steps", flush=True)
    o = oracle.COracle(L["nx"], L["ny"])
    apply(o, L)
    with fs.Batch([(L["nx"], L["ny"])]) as b:
        apply(b, L, 0)
        for s in range(1, STEPS + 1):
            o.execute_step()
            b.step(1)
            if s in (1, 2, 5, 10, STEPS):
                _close(name, s, _state(b, 0), _state(o))


def test_marine_landscape_in_a_batch_equals_alone():
    import fastscapelib_fortran as fs
    M = shelf(37, 45, seed=71, dx=1000.0, kf=1e-5, dt=1e3, kd=1e-2, ibc=1000, n=1.5, m=0.6)
    S = landscape(40, 33, seed=72, ibc=1010)
    F = landscape(29, 31, seed=73, p=1.0, g=1.0)
    Ls = [S, M, F]
    print("marine batch: 3 landscapes, 6 steps", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b, fs.Batch([(M["nx"], M["ny"])]) as a, \
            fs.Batch([(S["nx"], S["ny"]), (F["nx"], F["ny"])]) as c:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        apply(a, M, 0)
        apply(c, S, 0)
        apply(c, F, 1)
        for x in (a, b, c):
            x.step(6)
        got, alone = _state(b, 1), _state(a, 0)
        for f in got:
            assert np.array_equal(got[f], alone[f]), f
        for kb, kc in ((0, 0), (2, 1)):
            x, y = _state(b, kb), _state(c, kc)
            for f in ("h", "b", "etot", "erate", "sedflux"):
                assert np.array_equal(x[f], y[f]), (kb, f)


def test_f2py_marine_calls(oracle):
    import fastscapelib_fortran as fs
    L = shelf(31, 41, seed=81, dx=1000.0, kf=1e-5, dt=1e3, kd=1e-2, ibc=1100, fmix=None)
    nn = L["nx"] * L["ny"]
    fs.fastscape_init()
    fs.fastscape_set_nx_ny(L["nx"], L["ny"])
    fs.fastscape_setup()
    try:
        f0 = np.empty(nn)
        fs.fastscape_copy_f(f0)
        assert np.all(f0 == 0.5)  # SetUp default
        fs.fastscape_set_xl_yl(L["xl"], L["yl"])
        fs.fastscape_set_dt(L["dt"])
        fs.fastscape_init_h(L["h"])
        fs.fastscape_set_erosional_parameters(L["kf"], L["kfsed"], L["m"], L["n"], L["kd"], L["kdsed"], L["g1"],
                                              L["g2"], L["p"])
        fs.fastscape_set_marine_parameters(*L["marine"])
        fin = np.random.default_rng(3).random(nn)
        fs.fastscape_init_f(fin)
        fs.fastscape_set_u(L["u"])
        fs.fastscape_set_bc(L["ibc"])
        for _ in range(3):
            fs.fastscape_execute_step()
        h = np.empty(nn)
        f = np.empty(nn)
        fs.fastscape_copy_h(h)
        fs.fastscape_copy_f(f)
    finally:
        fs.fastscape_destroy()
    o = oracle.COracle(L["nx"], L["ny"])
    apply(o, dict(L, fmix=fin))
    for _ in range(3):
        o.execute_step()
    assert rel_err(h, o.field("h")) <= TOL
    assert float(np.max(np.abs(f - o.field("fmix")))) <= TOL


def test_marine_errors():
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    L = shelf(31, 41, seed=91, dx=1000.0, kf=1e-5, dt=1e3, kd=1e-2, ibc=1000)
    print("marine errors: no routing; no convergence", flush=True)
    with fs.Batch([(L["nx"], L["ny"])]) as b:
        b.set_xl_yl(0, L["xl"], L["yl"])
        b.set_dt(0, L["dt"])
        b.init_h(0, L["h"])
        b.set_marine_parameters(0, *L["marine"])
        with pytest.raises(_lib.FastscapeError, match="erosional parameters"):
            b.step(1)
        assert b.get_step(0) == 0
    # a diffusivity at which the reference's silt / sand iterations do not settle (it STOPs after 1000)
    bad = shelf(41, 61, seed=321, dx=1000.0, kf=1e-5, dt=1e3, kd=1e-2, ibc=101, n=2.0, m=0.8,
                marine=(0.5, 0.1, 0.2, 1e3, 1e3, 0.6, 1e2, 4e2, 1e2))
    with fs.Batch([(bad["nx"], bad["ny"])]) as b:
        apply(b, bad, 0)
        with pytest.raises(_lib.FastscapeError, match="did not converge"):
            for _ in range(20):  # one synchronous step at a time: stop at the step that fails
                b.step(1)
