"""Stratigraphic horizons: the oracle's run_Strati / Strati fields (oracle/fastscape_oracle.c) written by
fastscapelib_fortran.strati, against the files the reference itself writes (tests/golden/strati_*.npz, from
tests/golden/make_golden.py strati): every Horizon*.vtk, Strati-*.vtk and CUBE.vtk byte for byte, except
the erosion below the first reflector (field 10 of reflector 1), which the reference accumulates from its
reflector array before it first sets it (src/FastScape_ctx.f90:817,822 read what the allocation of :796 left),
so that its value there is whatever that memory held.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import ROOT, apply

GOLDEN = os.path.join(ROOT, "tests", "golden")
ERODED = b"SCALARS A.ThicknessErodedBelow(m) float 1\nLOOKUP_TABLE default\n"


def _load(name):
    g = dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))
    L = {k[3:]: g[k] for k in g if k.startswith("in_")}
    for k in ["nx", "ny", "ibc"]:
        L[k] = int(L[k])
    for k in ["xl", "yl", "kfsed", "m", "n", "kdsed", "g1", "g2", "p", "dt"]:
        L[k] = float(L[k])
    files = {str(n): g["file_" + str(n)].tobytes() for n in g["names"]}
    sched = g["sched"]
    return L, (int(sched[0]), int(sched[1]), int(sched[2]), float(sched[3])), int(g["steps"]), files


def mask_first_eroded(name, data, nn1):
    """The file with field 10 of reflector 1 set to zero bytes / '0' lines."""
    if name.startswith("Strati-"):
        lines = data.split(b"\n")
        at = lines.index(ERODED.split(b"\n")[0].replace(b" float", b"      float")) + 2
        for q in range(at + nn1, at + 2 * nn1):
            lines[q] = b"0"
        return b"\n".join(lines)
    if name == "CUBE.vtk" or name.startswith("Horizon001-"):
        at = data.index(ERODED) + len(ERODED)
        return data[:at] + b"\0" * (4 * nn1) + data[at + 4 * nn1:]
    return data


def oracle_files(oracle, L, sched, steps, directory):
    """Run the oracle with stratigraphy and write every output step with fastscapelib_fortran.strati."""
    from fastscapelib_fortran import strati as S
    o = oracle.COracle(L["nx"], L["ny"])
    apply(o, L)
    o.strati(*sched)
    for s in range(1, steps + 1):
        o.execute_step()
        refl, fields, iref, out = o.strati_state()
        if out:
            distb = oracle.distance_to_shore(o.field("b"), o.ifield("rec"), L["nx"], L["ny"], L["xl"], L["yl"])
            S.write_output(s, L["nx"], L["ny"], L["xl"], L["yl"], refl, fields, o.field("b"), distb, sched[3], iref,
                           directory)
    return {n: open(os.path.join(directory, n), "rb").read() for n in sorted(os.listdir(directory))}


@pytest.mark.parametrize("name", ["strati_marine_1000.npz", "strati_land_1111.npz"])
def test_strati_files_match_reference(oracle, tmp_path, name):
    L, sched, steps, ref = _load(name)
    ours = oracle_files(oracle, L, sched, steps, str(tmp_path / "VTK"))
    assert sorted(ours) == sorted(ref)
    nn1 = L["nx"] * L["ny"]
    for n in ref:
        assert mask_first_eroded(n, ours[n], nn1) == mask_first_eroded(n, ref[n], nn1), n


def test_fortran_e_format():
    """E10.4 as the reference's runtime writes it (values seen in its Strati files)."""
    from fastscapelib_fortran.strati import _e10_4
    assert _e10_4(0.0) == "0.0000E+00"
    assert _e10_4(1.0) == "0.1000E+01"
    assert _e10_4(-1234.56) == "-.1235E+04"
    assert _e10_4(0.00099996) == "0.1000E-02"
    assert _e10_4(2.5e120) == "0.2500+121"


def test_distance_to_shore_brute_force(oracle):
    """distance_to_shore against a direct evaluation of its definition on a small surface."""
    nx, ny, xl, yl = 9, 7, 800.0, 600.0
    rng = np.random.default_rng(4)
    h = rng.random(nx * ny) - 0.5
    rec = rng.integers(0, nx * ny, nx * ny).astype(np.int32)
    x = np.tile(xl * np.arange(nx) / (nx - 1), ny)
    y = np.repeat(yl * np.arange(ny) / (ny - 1), nx)
    d = np.full(nx * ny, 2 * (xl ** 2 + yl ** 2))
    for ij in range(nx * ny):
        r = rec[ij]
        if h[ij] * h[r] < 0:
            rat = h[ij] / (h[r] - h[ij]) if h[r] != h[ij] else 0.5
            xs, ys = x[ij] + rat * (x[r] - x[ij]), y[ij] + rat * (y[r] - y[ij])
            d = np.minimum(d, (x - xs) ** 2 + (y - ys) ** 2)
    want = np.where(d > 0, np.sqrt(d), d)
    assert np.allclose(oracle.distance_to_shore(h, rec, nx, ny, xl, yl), want, rtol=1e-13, atol=0)
