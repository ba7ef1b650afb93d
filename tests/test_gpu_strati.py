"""Stratigraphic horizons on the GPU (run_Strati's reflector update inside every step; slopes and distances to
shore of an output step computed on the GPU; fields and files completed by fastscapelib_fortran.strati), through
the C ABI, against the oracle on the reference's stratigraphy fixtures (tests/golden/strati_*.npz, which
tests/test_strati.py holds the oracle and the writer to byte for byte): at every output step the reflectors,
the erosion below them, the ten fields and the basement's distance to shore within 1e-10 (relative, of the
elevation scale for differences), and the same files written; then the f2py-style fastscape_strati call.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import ROOT, apply
from test_strati import _load, mask_first_eroded

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _close(a, b, scale, what):
    err = float(np.max(np.abs(np.asarray(a) - np.asarray(b))))
    assert err <= TOL * scale, (what, err)


@pytest.mark.parametrize("name", ["strati_marine_1000.npz", "strati_land_1111.npz"])
def test_strati_matches_oracle(oracle, tmp_path, name):
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    L, sched, steps, ref = _load(name)
    nn = L["nx"] * L["ny"]
    print(f"strati {name}: {L['nx']}x{L['ny']} schedule {sched}, {steps} steps", flush=True)
    o = oracle.COracle(L["nx"], L["ny"])
    apply(o, L)
    o.strati(*sched)
    with fs.Batch([(L["nx"], L["ny"])]) as b:
        apply(b, L, 0)
        b.strati(0, *sched, directory=str(tmp_path / "VTK"))
        for s in range(1, steps + 1):
            o.execute_step()
            b.step(1)
            refl, fields, iref, out = o.strati_state()
            st = b._strati[0]
            scale = max(1.0, float(np.abs(o.field("h")).max()))
            _close(b.strati_read(0, _lib.FSX_STRATI_REFLECTORS, sched[1]), refl, scale, (s, "reflectors"))
            _close(b.strati_read(0, _lib.FSX_STRATI_ERODED, sched[1]), fields[:, 9], scale, (s, "eroded"))
            if out:
                assert st["last"] == s and st["iref"] == iref, (s, st["iref"], iref)
                for i in range(sched[1]):
                    for f in range(10):
                        x = fields[i, f]
                        _close(st["fields"][i, f], x, max(1.0, float(np.abs(x).max())), (s, i, f))
                distb = oracle.distance_to_shore(o.field("b"), o.ifield("rec"), L["nx"], L["ny"], L["xl"], L["yl"])
                got = b.strati_read(0, _lib.FSX_STRATI_SHORE, sched[1] + 1)[-1]
                _close(got, distb, max(1.0, float(distb.max())), (s, "basement distance"))
    names = sorted(os.listdir(tmp_path / "VTK"))
    assert names == sorted(ref)
    # same layout and text as the reference's files; numbers equal where the states agree bit for bit
    for n in names:
        data = (tmp_path / "VTK" / n).read_bytes()
        assert len(data) == len(ref[n]), n
        if n.startswith("Strati-"):
            ours = mask_first_eroded(n, data, nn).split(b"\n")
            want = mask_first_eroded(n, ref[n], nn).split(b"\n")
            assert [w for w in ours if not w[:1].isdigit() and w[:1] not in b" -."] == \
                   [w for w in want if not w[:1].isdigit() and w[:1] not in b" -."], n


def test_f2py_strati(tmp_path, monkeypatch):
    import fastscapelib_fortran as fs
    L, sched, steps, ref = _load("strati_land_1111.npz")
    monkeypatch.chdir(tmp_path)
    fs.fastscape_init()
    fs.fastscape_set_nx_ny(L["nx"], L["ny"])
    fs.fastscape_setup()
    try:
        fs.fastscape_set_xl_yl(L["xl"], L["yl"])
        fs.fastscape_set_dt(L["dt"])
        fs.fastscape_init_h(L["h"])
        fs.fastscape_set_erosional_parameters(L["kf"], L["kfsed"], L["m"], L["n"], L["kd"], L["kdsed"], L["g1"],
                                              L["g2"], L["p"])
        fs.fastscape_set_u(L["u"])
        fs.fastscape_set_bc(L["ibc"])
        fs.fastscape_strati(*sched)
        for _ in range(steps):
            fs.fastscape_execute_step()
        assert fs.fastscape_get_step() == steps
    finally:
        fs.fastscape_destroy()
    assert sorted(os.listdir(tmp_path / "VTK")) == sorted(ref)
