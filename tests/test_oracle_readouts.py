"""Pins the oracle's read-outs (slope, curvature, chi, fluxes: oracle/fastscape_oracle.c after
fso_copy_field) against

1. the unmodified reference compiled from its own sources (oracle/_ref), after steps of the SFD and
   MFD paths under several boundary conditions: bit for bit (skipped when the reference build is
   absent), and the catchment's grouping (the reference labels each tree of the receivers found
   before closed-basin rerouting with a random number, the oracle with (outlet + 1) / nn); and
2. what the mathematics fixes: the slope of a plane (atan of the gradient's norm, in degrees), the
   curvature of a plane (0) and of a parabola in x (c/2 / (1 + (2cx)^2)^(3/2)), chi on a tree with a
   constant area (half the depth when (a0/a)^(m/n) = 1/2 and lengths are 1, restarting from 0 below
   the a0 threshold), and the fluxes when every rate has one sign (the boundary receives twice the
   total along single flow -- a base-level node adds its own flux to itself -- and nothing when
   every rate is negative).
"""
from __future__ import annotations

import numpy as np
import pytest


def _ref():
    import oracle as O
    if not O.ref_available():
        pytest.skip("reference build absent (make -C oracle ref needs the reference sources)")
    return O.FortranRef()


CASES = [  # nx, ny, p, ibc, g
    (41, 37, -2.0, 1111, 0.0),
    (41, 37, -2.0, 101, 0.0),
    (30, 45, -2.0, 1010, 0.0),
    (30, 45, 1.0, 1010, 1.0),
    (33, 29, -1.0, 1111, 0.0),
]


@pytest.mark.parametrize("nx,ny,p,ibc,g", CASES)
def test_readouts_match_reference_live(oracle, nx, ny, p, ibc, g):
    f = _ref()
    f.setup(nx, ny)
    o = oracle.COracle(nx, ny)
    rng = np.random.default_rng(nx * ny + ibc)
    h = rng.random(nx * ny) * 10.0
    for x in (f, o):
        x.set_xl_yl(100.0 * (nx - 1), 90.0 * (ny - 1))
        x.set_dt(1e3)
        x.init_h(h)
        x.set_erosional_parameters(np.full(nx * ny, 2e-5), 1e-5 if g else -1.0, 0.4, 1.0, np.full(nx * ny, 1e-2),
                                   -1.0, g, g, p)
        x.set_u(np.full(nx * ny, 1e-3))
        x.set_bc(ibc)
        for _ in range(3):
            x.execute_step()
    for name in ["slope", "curvature", "chi"]:
        assert np.array_equal(o.derived(name), f.derived(name)), name
    assert o.fluxes() == f.fluxes()


@pytest.mark.parametrize("ibc,p", [(1111, -2.0), (1010, -2.0), (101, 1.0)])
def test_catchment_groups_match_reference_live(oracle, ibc, p):
    f = _ref()
    nx, ny, dt = 37, 29, 1e3
    f.setup(nx, ny)
    rng = np.random.default_rng(ibc + 7)
    h = rng.random(nx * ny) * 10.0
    u = np.full(nx * ny, 1e-3)
    f.set_xl_yl(100.0 * (nx - 1), 90.0 * (ny - 1))
    f.set_dt(dt)
    f.init_h(h)
    f.set_erosional_parameters(np.full(nx * ny, 2e-5), -1.0, 0.4, 1.0, np.full(nx * ny, 1e-2), -1.0, 0.0, 0.0, p)
    f.set_u(u)
    f.set_bc(ibc)
    f.execute_step()
    h1 = f.field("h")
    f.execute_step()
    cat = f.derived("catchment")
    rec_pre, _ = oracle.find_receiver(h1 + u * dt, nx, ny, 100.0, 90.0, ibc)
    lab = oracle.catchment(rec_pre)
    _, ia = np.unique(lab, return_inverse=True)
    _, ib = np.unique(cat, return_inverse=True)
    assert len(set(zip(ia.tolist(), ib.tolist()))) == ia.max() + 1 == ib.max() + 1  # the same partition
    roots = rec_pre == np.arange(nx * ny)
    assert np.array_equal(lab[roots], (np.flatnonzero(roots) + 1) / (nx * ny))


def test_slope_and_curvature_of_a_plane(oracle):
    nx, ny, dx, dy = 23, 17, 50.0, 40.0
    x = np.arange(nx) * dx
    y = np.arange(ny) * dy
    alpha, beta = 0.03, -0.02
    h = (alpha * x[None, :] + beta * y[:, None]).ravel()
    s = oracle.slope(h, nx, ny, dx, dy).reshape(ny, nx)
    want = np.degrees(np.arctan(np.hypot(alpha, beta)))
    assert np.allclose(s[1:-1, 1:-1], want, rtol=1e-12, atol=0)
    assert (s[0] == 0).all() and (s[-1] == 0).all() and (s[:, 0] == 0).all() and (s[:, -1] == 0).all()
    c = oracle.curvature(h, nx, ny, dx, dy).reshape(ny, nx)
    assert np.abs(c).max() < 1e-17


def test_curvature_of_a_parabola(oracle):
    nx, ny, dx, dy = 21, 9, 10.0, 12.0
    cc = 1e-3
    x = (np.arange(nx) - 7) * dx
    h = np.broadcast_to(cc * x[None, :] ** 2, (ny, nx)).ravel().copy()
    got = oracle.curvature(h, nx, ny, dx, dy).reshape(ny, nx)
    want = (cc / 2.0) / (1.0 + (2.0 * cc * x) ** 2) ** 1.5
    assert np.allclose(got[1:-1, 1:-1], np.broadcast_to(want[None, 1:-1], (ny - 2, nx - 2)), rtol=1e-9, atol=1e-15)


def _random_tree(n, rng):
    """A random forest given as rec (roots point to themselves), a receivers-first order and depths."""
    rec = np.arange(n)
    roots = rng.choice(n, size=3, replace=False)
    others = np.setdiff1d(np.arange(n), roots)
    rng.shuffle(others)
    placed = list(roots)
    for v in others:
        rec[v] = placed[rng.integers(len(placed))]
        placed.append(v)
    depth = np.zeros(n, int)
    for v in placed:
        depth[v] = 0 if rec[v] == v else depth[rec[v]] + 1
    return rec, np.array(placed, np.int32), depth


def test_chi_on_a_constant_area_tree(oracle):
    rng = np.random.default_rng(5)
    n, dx, dy = 400, 3.0, 2.0
    rec, order, depth = _random_tree(n, rng)
    a0 = dx * dy * 10.0
    a = np.full(n, 4.0 * a0)  # (a0/a)^(m/n) = (1/4)^(1/2) = 1/2
    length = np.where(rec == np.arange(n), 0.0, 1.0)
    chi = oracle.chi(order, rec, a, length, dx, dy, 0.4, 0.8)
    assert np.array_equal(chi, 0.5 * depth)
    # nodes at or below the threshold hold 0 and their donors start again from 0
    low = rng.random(n) < 0.2
    a2 = np.where(low, a0, 4.0 * a0)
    chi2 = oracle.chi(order, rec, a2, length, dx, dy, 0.4, 0.8)
    run = np.zeros(n)
    for v in order:  # distance in hops to the nearest node at or below the threshold, or to the root
        run[v] = 0.0 if (low[v] or rec[v] == v) else run[rec[v]] + 1.0
    assert np.array_equal(chi2, 0.5 * run)


def test_fluxes_single_flow_closed_forms(oracle):
    rng = np.random.default_rng(9)
    n = 300
    rec, order, _ = _random_tree(n, rng)
    roots = rec == np.arange(n)
    bc = roots.astype(np.uint8)
    u = rng.random(n)
    e = rng.random(n) * 1e-3
    t, er, b = oracle.fluxes_sfd(order, rec, u, e, bc, 2.5)
    assert np.isclose(t, u.sum() * 2.5, rtol=1e-13) and np.isclose(er, e.sum() * 2.5, rtol=1e-13)
    assert np.isclose(b, 2.0 * er, rtol=1e-12)
    _, er2, b2 = oracle.fluxes_sfd(order, rec, u, -e, bc, 2.5)
    assert b2 == 0.0 and er2 < 0
