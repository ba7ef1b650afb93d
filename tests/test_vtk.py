"""The legacy-VTK writer (fastscapelib_fortran/vtk.py; Make_VTK, src/FastScape_ctx.f90:692-777) against the
files the reference build itself writes from the same state: byte for byte, with a positive vertical
exaggeration (topography only) and a negative one (topography, basement and sea level).  Skipped when the
reference build is absent."""
from __future__ import annotations

import os

import numpy as np
import pytest


@pytest.mark.parametrize("vex", [2.0, -3.0])
def test_vtk_files_match_reference(oracle, tmp_path, monkeypatch, vex):
    if not oracle.ref_available():
        pytest.skip("reference build absent (make -C oracle ref needs the reference sources)")
    from fastscapelib_fortran.vtk import write_vtk
    nx, ny, xl, yl = 23, 17, 2200.0, 1600.0
    f = oracle.FortranRef()
    f.setup(nx, ny)
    rng = np.random.default_rng(8)
    f.set_xl_yl(xl, yl)
    f.set_dt(1e3)
    f.init_h(rng.random(nx * ny) * 10.0)
    f.set_erosional_parameters(np.full(nx * ny, 2e-5), -1.0, 0.4, 1.0, np.full(nx * ny, 1e-2), -1.0, 0.0, 0.0, -2.0)
    f.set_u(np.full(nx * ny, 1e-3))
    f.set_bc(1111)
    for _ in range(12):
        f.execute_step()
    field = f.field("erate")
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    monkeypatch.chdir(ref_dir)
    f.vtk(field, vex)
    names = sorted(os.listdir(ref_dir / "VTK"))
    assert names == (["Topography0000012.vtk"] if vex > 0 else
                     ["Basement0000012.vtk", "SeaLevel0000012.vtk", "Topography0000012.vtk"])
    paths = write_vtk(f.get_step(), nx, ny, xl, yl, f.field("h"), f.field("b"), field, vex,
                      directory=str(tmp_path / "ours"))
    assert sorted(os.path.basename(p) for p in paths) == names
    for n in names:
        assert (tmp_path / "ours" / n).read_bytes() == (ref_dir / "VTK" / n).read_bytes(), n
