"""Seeded synthetic landscape workloads of the benchmark configurations (shared by bench.py, the tests
and the fixture generators, so that all three step the same landscapes).

A landscape is a dict: nx, ny, xl, yl, dt, ibc, h (initial elevation, nx*ny values, x fastest), u
(uplift rate per node), kf and kd (a scalar or nx*ny values), kfsed, m, n, kdsed, g1, g2, p (the nine
arguments of fastscape_set_erosional_parameters; p < -1.5 selects single flow direction), and
optionally newton_tol (the stream-power Newton tolerance; the reference's is 1e-3).

* ``plain_cpu()``       -- the reference's own CPU case: 101 x 101, dx = 1 km, D8, 200 steps;
* ``uniform_batch()``   -- 64 x 513 x 513, dx = 200 m, seeds 0..63, kf log-uniform in [2e-6, 2e-5],
  U ~ U[5e-4, 2e-3] m/yr, n = 1, 500 steps;
* ``ragged_batch()``    -- 200 landscapes, nx, ny log-uniform in [65, 1025], dx = 250 m, noise + a 10 m
  2-D sinusoid of wavelength U[5, 50] km, n from {1, 1.5, 2}, m = 0.45 n, Newton tolerance 1e-8,
  BC 1010 / 1111 alternating, dt = 2e3 yr, 300 steps;
* ``mixed_batch()``     -- one 4097 x 4097 multiple-flow landscape with deposition (g = 1) and 64 D8
  landscapes of 257 x 257 with the CPU case's parameters, dt = 1e3 yr, 200 steps;
* ``serving_stream()``  -- Poisson arrivals of runs, grids log-uniform in [65, 2049], step counts
  log-uniform in [50, 2000], dt log-uniform in [1e3, 1e4] yr, parameters as the ragged batch, D8 or
  multiple flow direction 50/50.
"""
from __future__ import annotations

import numpy as np

PLAIN_STEPS, UNIFORM_STEPS, RAGGED_STEPS, MIXED_STEPS = 200, 500, 300, 200


def fixed_mask(nx, ny, ibc):
    """Nodes on a fixed (base-level) boundary for a BC code (digits bottom/right/top/left)."""
    c = f"{int(ibc):4d}"
    e = np.zeros((ny, nx), bool)
    if c[0] == "1":
        e[0, :] = True
    if c[1] == "1":
        e[:, -1] = True
    if c[2] == "1":
        e[-1, :] = True
    if c[3] == "1":
        e[:, 0] = True
    return e.ravel()


def _sinus(rng, nx, ny, dx, amp=10.0):
    """h = U[0,1) noise + amp * sin(2 pi x / lam) * sin(2 pi y / lam), lam ~ U[5, 50] km (lam drawn first)."""
    lam = rng.uniform(5e3, 5e4)
    h = rng.random(nx * ny)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    return h + amp * (np.sin(2 * np.pi * x[None, :] / lam) * np.sin(2 * np.pi * y[:, None] / lam)).ravel()


def _d8(nx, ny, dx, h, u, kf, m, n, kd, dt, ibc, **kw):
    L = dict(nx=nx, ny=ny, xl=dx * (nx - 1), yl=dx * (ny - 1), h=h, u=u, kf=kf, kfsed=-1.0, m=m, n=n, kd=kd,
             kdsed=-1.0, g1=0.0, g2=0.0, p=-2.0, dt=dt, ibc=ibc)
    L.update(kw)
    return L


def plain_cpu(n=101, seed=0, dx=1000.0):
    """The reference's CPU case: h = U[0,1) from PCG64(seed), U = 1e-3 m/yr inside, 0 on the fixed edge."""
    rng = np.random.default_rng(seed)
    h = rng.random(n * n)
    u = np.full(n * n, 1e-3)
    u[fixed_mask(n, n, 1111)] = 0.0
    return _d8(n, n, dx, h, u, 1e-5, 0.4, 1.0, 1e-2, 1e3, 1111)


def uniform_member(seed, n=513, dx=200.0):
    """One landscape of the uniform batch: h noise, then kf, then U, from PCG64(seed)."""
    rng = np.random.default_rng(seed)
    h = rng.random(n * n)
    kf = 10.0 ** rng.uniform(np.log10(2e-6), np.log10(2e-5))
    U = rng.uniform(5e-4, 2e-3)
    u = np.full(n * n, U)
    u[fixed_mask(n, n, 1111)] = 0.0
    return _d8(n, n, dx, h, u, kf, 0.4, 1.0, 1e-2, 1e3, 1111)


def uniform_batch(ninst=64, n=513, dx=200.0):
    return [uniform_member(s, n, dx) for s in range(ninst)]


def ragged_member(i, lo=65, hi=1025, seed0=1000):
    """Landscape i of the ragged batch, from PCG64(seed0 + i): nx, ny, h (sinusoid then noise), n, kf,
    kd; BC 1010 (periodic east-west, fixed north-south) for even i, 1111 for odd i; uplift 1e-3 m/yr off
    the fixed boundary."""
    rng = np.random.default_rng(seed0 + i)
    nx = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    ny = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    dx = 250.0
    h = _sinus(rng, nx, ny, dx)
    n = float(rng.choice([1.0, 1.5, 2.0]))
    kf = 10.0 ** rng.uniform(-6.0, -5.0)
    kd = rng.uniform(1e-3, 1e-1)
    ibc = 1010 if i % 2 == 0 else 1111
    u = np.full(nx * ny, 1e-3)
    u[fixed_mask(nx, ny, ibc)] = 0.0
    return _d8(nx, ny, dx, h, u, kf, 0.45 * n, n, kd, 2e3, ibc, newton_tol=1e-8)


def ragged_batch(ninst=200, lo=65, hi=1025, seed0=1000):
    return [ragged_member(i, lo, hi, seed0) for i in range(ninst)]


def mixed_large(n=4097, dx=100.0, seed=4097):
    """The multiple-flow landscape of the mixed batch: h = U[0,1), U = 2e-3 m/yr on the central half of
    each direction, p = 1, g = 1, kf = kfsed = 5e-6, kd = kdsed = 5e-2, n = 1, m = 0.4, fixed boundary."""
    rng = np.random.default_rng(seed)
    h = rng.random(n * n)
    u = np.zeros((n, n))
    q = n // 4
    u[q:n - q, q:n - q] = 2e-3
    return dict(nx=n, ny=n, xl=dx * (n - 1), yl=dx * (n - 1), h=h, u=u.ravel(), kf=5e-6, kfsed=5e-6, m=0.4, n=1.0,
                kd=5e-2, kdsed=5e-2, g1=1.0, g2=1.0, p=1.0, dt=1e3, ibc=1111)


def mixed_small(i, n=257):
    """D8 landscape i of the mixed batch: the CPU case's parameters on 257 x 257, h from PCG64(5000 + i)."""
    return plain_cpu(n=n, seed=5000 + i)


def mixed_batch(nsmall=64, large=4097, small=257):
    """[large multiple-flow landscape] + nsmall D8 landscapes, in that order."""
    return [mixed_large(large)] + [mixed_small(i, small) for i in range(nsmall)]


def serving_run(i, seed=7, lo=65, hi=2049, smin=50, smax=2000):
    """Run i of the serving stream, from PCG64(seed * 100000 + i): (steps, landscape)."""
    r = np.random.default_rng(seed * 100000 + i)
    nx = int(round(np.exp(r.uniform(np.log(lo), np.log(hi)))))
    ny = int(round(np.exp(r.uniform(np.log(lo), np.log(hi)))))
    steps = int(round(np.exp(r.uniform(np.log(smin), np.log(smax)))))
    dt = float(np.exp(r.uniform(np.log(1e3), np.log(1e4))))
    dx = 250.0
    h = _sinus(r, nx, ny, dx)
    n = float(r.choice([1.0, 1.5, 2.0]))
    kf = 10.0 ** r.uniform(-6.0, -5.0)
    kd = r.uniform(1e-3, 1e-1)
    mfd = bool(r.random() < 0.5)
    ibc = 1010 if i % 2 == 0 else 1111
    u = np.full(nx * ny, 1e-3)
    u[fixed_mask(nx, ny, ibc)] = 0.0
    L = _d8(nx, ny, dx, h, u, kf, 0.45 * n, n, kd, dt, ibc, newton_tol=1e-8)
    L["p"] = 1.0 if mfd else -2.0
    return steps, L


def serving_arrivals(rate=20.0, duration=600.0, seed=7):
    """Poisson(rate) arrival times over `duration` seconds, from PCG64(seed)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        out.append(t)


def serving_stream(rate=20.0, duration=600.0, seed=7, **kw):
    """[(arrival s, steps, landscape)] of the serving stream over `duration` seconds."""
    return [(t,) + serving_run(i, seed, **kw) for i, t in enumerate(serving_arrivals(rate, duration, seed))]


def configure(obj, L, *k):
    """Set a landscape up on a Batch / Pool / Lanes (k = its id) or an oracle (no k)."""
    nn = L["nx"] * L["ny"]
    obj.set_xl_yl(*k, L["xl"], L["yl"])
    obj.set_dt(*k, L["dt"])
    obj.init_h(*k, L["h"])
    kf = np.broadcast_to(np.asarray(L["kf"], np.float64), (nn,))
    kd = np.broadcast_to(np.asarray(L["kd"], np.float64), (nn,))
    obj.set_erosional_parameters(*k, np.ascontiguousarray(kf), L["kfsed"], L["m"], L["n"], np.ascontiguousarray(kd),
                                 L["kdsed"], L["g1"], L["g2"], L["p"])
    obj.set_u(*k, L["u"])
    obj.set_bc(*k, L["ibc"])
    if "newton_tol" in L and hasattr(obj, "set_newton_tol"):
        obj.set_newton_tol(*k, float(L["newton_tol"]))


def cells(Ls):
    return sum(L["nx"] * L["ny"] for L in Ls)
