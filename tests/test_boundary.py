"""The drop-in boundary without a GPU: the C-ABI library loads and exports every entry point its
header declares, host-side argument checks answer with the documented codes, the f2py-style facade
keeps the reference's call-order errors, and bench.py refuses more than one GPU."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fastscape_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fsx_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from fastscapelib_fortran import _lib
    return _lib


def test_header_declares_the_step_surface():
    names = _declared()
    for need in ["fsx_create", "fsx_destroy", "fsx_set_xl_yl", "fsx_set_dt", "fsx_set_bc",
                 "fsx_set_erosional_parameters", "fsx_init_h", "fsx_set_field", "fsx_execute_steps",
                 "fsx_copy_field", "fsx_get_step", "fsx_indexx", "fsx_route"]:
        assert need in names


def test_library_exports_every_declared_symbol(lib):
    so = C.CDLL(lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(so, n)]
    assert not missing, missing
    # and the Python binding gives every one of them a signature
    for n in _declared():
        assert getattr(lib.lib, n).restype is not None, n


def test_host_side_argument_checks(lib):
    L = lib.lib
    out = C.c_void_p()
    two, five = (C.c_int * 1)(2), (C.c_int * 1)(5)
    assert L.fsx_create(1, two, five, C.byref(out)) == -1
    assert b"nx and ny" in L.fsx_last_error()
    assert L.fsx_create(0, five, five, C.byref(out)) == -1
    assert L.fsx_indexx(-1, None, None) == -1
    assert L.fsx_indexx(0, None, None) == 0
    assert L.fsx_destroy(None) == 0
    assert L.fsx_autotune(None, 1, None) == -1
    assert L.fsx_get_tuning(None, None) == -1
    assert L.fsx_set_tuning(None, None) == -1
    assert L.fsx_copy_derived(None, 0, 0, None, 0) == -1
    assert L.fsx_get_fluxes(None, 0, None) == -1
    assert L.fsx_version().startswith(b"fastscape")


def test_no_cpu_fallback_without_a_device(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(lib.FastscapeError):
        import fastscapelib_fortran as fs
        fs.Batch([(5, 5)])


def test_facade_call_order_errors():
    import fastscapelib_fortran as fs
    fs.fastscape_destroy()
    with pytest.raises(RuntimeError, match="SetUp"):
        fs.fastscape_execute_step()
    fs.fastscape_init()
    with pytest.raises(RuntimeError, match="nx"):
        fs.fastscape_setup()
    with pytest.raises(RuntimeError, match="SetUp"):
        fs.fastscape_copy_h(np.zeros(4))
    with pytest.raises(RuntimeError, match="SetUp"):
        fs.fastscape_get_fluxes()
    with pytest.raises(RuntimeError, match="SetUp"):
        fs.fastscape_set_marine_parameters(0.0, 0.0, 0.0, 1e3, 1e3, 0.5, 1e2, 3e2, 1.5e2)
    with pytest.raises(RuntimeError, match="SetUp"):
        fs.fastscape_copy_f(np.zeros(4))
    with pytest.raises(RuntimeError, match="SetUp"):
        fs.fastscape_strati(10, 2, 5, 1.0)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_oracle_library_exports(oracle):
    so = C.CDLL(oracle.build_c())
    for n in ["fso_create", "fso_execute_step", "fso_local_minima", "fso_indexx", "fso_find_receiver",
              "fso_find_donor", "fso_find_stack", "fso_copy_field", "fso_slope", "fso_curvature", "fso_chi",
              "fso_fluxes_sfd", "fso_fluxes_mfd", "fso_get_fluxes", "fso_catchment"]:
        assert hasattr(so, n), n
