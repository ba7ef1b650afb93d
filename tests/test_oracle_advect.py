"""Pins the oracle's Advect (oracle/fastscape_oracle.c, advect; src/Advect.f90:1-103) against

1. the unmodified reference compiled from its own sources (oracle/_ref): h, b, etot bit for bit after
   steps with advection alone and with the whole step (single and multiple flow), skipped when the
   reference build is absent; and
2. the implicit upwind discretisation of dh/dt + v dh/dx = 0 on each grid row,
   (h'_i - h_i)/dt + v (h'_i - h'_{i-1})/dx = 0 for v > 0 and (h'_{i+1} - h'_i) for v < 0, with the
   first and last node of a row held, solved by LAPACK's banded solver (scipy.linalg.solve_banded):
   agreement to 1e-13, so a wrong upwind side, sign or end row fails; a zero velocity leaves the
   surface unchanged, and b <= h after the step.
"""
from __future__ import annotations

import numpy as np
import pytest


def _setup(x, nx, ny, h, vx, vy, full, p=-2.0, ibc=1111):
    x.set_xl_yl(100.0 * (nx - 1), 90.0 * (ny - 1))
    x.set_dt(1e3)
    x.init_h(h)
    if full:
        x.set_erosional_parameters(np.full(nx * ny, 2e-5), -1.0, 0.4, 1.0, np.full(nx * ny, 1e-2), -1.0, 0.0, 0.0, p)
        x.set_u(np.full(nx * ny, 1e-3))
    x.set_bc(ibc)
    x.set_v(vx, vy)


@pytest.mark.parametrize("full,p,ibc", [(False, -2.0, 1111), (True, -2.0, 1111), (True, -2.0, 1010), (True, 1.0, 101)])
def test_advect_matches_reference_live(oracle, full, p, ibc):
    if not oracle.ref_available():
        pytest.skip("reference build absent (make -C oracle ref needs the reference sources)")
    f = oracle.FortranRef()
    nx, ny = 37, 31
    f.setup(nx, ny)
    o = oracle.COracle(nx, ny)
    rng = np.random.default_rng(ibc + int(full))
    h = rng.random(nx * ny) * 10.0
    vx = (rng.random(nx * ny) - 0.3) * 0.02
    vy = (rng.random(nx * ny) - 0.6) * 0.015
    vx[::7] = 0.0
    for x in (f, o):
        _setup(x, nx, ny, h, vx, vy, full, p, ibc)
        for _ in range(3):
            x.execute_step()
    for name in ["h", "b", "etot"]:
        assert np.array_equal(o.field(name), f.field(name)), name


def test_advect_rows_solve_the_upwind_scheme(oracle):
    from scipy.linalg import solve_banded
    nx, ny, dt = 29, 5, 1e3
    dx = 100.0
    rng = np.random.default_rng(3)
    h = rng.random(nx * ny) * 10.0
    vx = (rng.random(nx * ny) - 0.5) * 0.2
    vx[::5] = 0.0
    o = oracle.COracle(nx, ny)
    o.set_xl_yl(dx * (nx - 1), 90.0 * (ny - 1))
    o.set_dt(dt)
    o.init_h(h)
    o.set_v(vx, np.zeros(nx * ny))  # y velocity 0: the column pass is the identity
    o.execute_step()
    got = o.field("h").reshape(ny, nx)
    H, V = h.reshape(ny, nx), vx.reshape(ny, nx)
    for j in range(ny):
        c = V[j] * dt / dx
        ab = np.zeros((3, nx))  # rows: super, diagonal, sub (LAPACK band storage)
        ab[1] = 1.0 + np.abs(c)
        ab[2, :-1] = np.where(c[1:] > 0, -c[1:], 0.0)   # coefficient of h'_{i-1} in row i
        ab[0, 1:] = np.where(c[:-1] < 0, c[:-1], 0.0)   # coefficient of h'_{i+1} in row i
        ab[1, 0] = ab[1, -1] = 1.0
        ab[0, 1] = 0.0
        ab[2, -2] = 0.0
        want = solve_banded((1, 1), ab, H[j])
        assert np.allclose(got[j], want, rtol=1e-13, atol=1e-12), j
    assert np.array_equal(got[:, 0], H[:, 0]) and np.array_equal(got[:, -1], H[:, -1])


def test_advect_zero_velocity_is_identity_and_clamps_b(oracle):
    nx, ny = 17, 13
    rng = np.random.default_rng(4)
    h = rng.random(nx * ny)
    o = oracle.COracle(nx, ny)
    o.set_xl_yl(1600.0, 1200.0)
    o.set_dt(1e3)
    o.init_h(h)
    o.set_basement(h + 0.5)
    o.set_v(np.zeros(nx * ny), np.zeros(nx * ny))
    o.execute_step()
    assert np.array_equal(o.field("h"), h)
    assert np.array_equal(o.field("b"), h)
