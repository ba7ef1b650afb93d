"""Shared test set-up: the `gpu` marker, import paths, and the seeded landscape recipes."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fastscapelib-fortran_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build_c()
    return O


def edge_mask(nx, ny):
    m = np.zeros((ny, nx), bool)
    m[0, :] = m[-1, :] = True
    m[:, 0] = m[:, -1] = True
    return m.ravel()


def landscape(nx, ny, seed, *, dx=1000.0, u=1e-3, kf=1e-5, m=0.4, n=1.0, kd=1e-2, ibc=1111, dt=1e3, g=0.0,
              sinus=None, kfsed=-1.0, kdsed=-1.0, p=-2.0):
    """A seeded landscape description: h i.i.d. uniform [0,1) (+ optional 2-D sinusoid), uplift u inside
    and 0 on the grid edge (fixed base level), uniform kf/kd."""
    rng = np.random.default_rng(seed)
    h = rng.random(nx * ny)
    if sinus is not None:
        amp, lam = sinus
        x = np.arange(nx) * dx
        y = np.arange(ny) * dx
        h = h + amp * (np.sin(2 * np.pi * x[None, :] / lam) * np.sin(2 * np.pi * y[:, None] / lam)).ravel()
    uu = np.full(nx * ny, u)
    uu[edge_mask(nx, ny)] = 0.0
    return dict(nx=nx, ny=ny, xl=dx * (nx - 1), yl=dx * (ny - 1), h=h, u=uu, kf=np.full(nx * ny, kf), kfsed=kfsed,
                m=m, n=n, kd=np.full(nx * ny, kd), kdsed=kdsed, g1=g, g2=g, p=p, ibc=ibc, dt=dt)


def shelf(nx, ny, seed, *, marine=(0.0, 0.0, 0.0, 1e3, 1e3, 0.5, 1e2, 3e2, 1.5e2), fmix=None, **kw):
    """A landscape with a sea: the southern half lowered by 1000 m with no uplift there (as the
    reference's Margin example, examples/Margin.f90), and the marine parameters (sea level, porosities,
    e-folding depths, silt ratio, mixed-layer thickness, diffusivities); fmix: initial silt fraction."""
    L = landscape(nx, ny, seed, **kw)
    south = np.repeat(np.arange(ny) < ny // 2, nx)
    L["h"] = np.where(south, L["h"] - 1000.0, L["h"])
    L["u"] = np.where(south, 0.0, L["u"])
    L["marine"] = np.asarray(marine, np.float64)
    if fmix is not None:
        L["fmix"] = np.asarray(fmix, np.float64)
    return L


def apply(obj, L, k=None):
    """Configure an oracle / reference / Batch instance from a landscape description."""
    a = (k,) if k is not None else ()
    obj.set_xl_yl(*a, L["xl"], L["yl"])
    obj.set_dt(*a, L["dt"])
    obj.init_h(*a, L["h"])
    obj.set_erosional_parameters(*a, L["kf"], L["kfsed"], L["m"], L["n"], L["kd"], L["kdsed"], L["g1"], L["g2"],
                                 L["p"])
    obj.set_u(*a, L["u"])
    obj.set_bc(*a, L["ibc"])
    if "newton_tol" in L:
        obj.set_newton_tol(*a, float(L["newton_tol"]))
    if "marine" in L:
        obj.set_marine_parameters(*a, *[float(x) for x in L["marine"]])
    if "fmix" in L:
        obj.init_f(*a, L["fmix"])


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0
