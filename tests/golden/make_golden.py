"""Writes tests/golden/*.npz from the reference Fortran compiled from its own sources
(oracle/_ref/libfsref.so, built by `make -C oracle ref`).  Each fixture holds the seeded inputs and
the reference's outputs, so tests can check the oracle and the GPU path without the reference.

    python tests/golden/make_golden.py          # fixtures of the unmodified reference build
    python tests/golden/make_golden.py tol8     # fixtures of the build with Newton tolerance 1e-8
                                                # (make -C oracle ref8), names tol8_*
    python tests/golden/make_golden.py marine   # runs with the marine transport, names marine_*, and
                                                # vectors of its silt / sand diffusion (silt_sand.npz)
    python tests/golden/make_golden.py strati   # the files the reference's stratigraphy writes
                                                # (strati_*.npz)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import oracle as O  # noqa: E402
from conftest import landscape, shelf  # noqa: E402


def run_ref(L, steps_to_save, lib=O.REF_LIB):
    f = O.FortranRef(lib)
    f.setup(L["nx"], L["ny"])
    f.set_xl_yl(L["xl"], L["yl"])
    f.set_dt(L["dt"])
    f.init_h(L["h"])
    f.set_erosional_parameters(L["kf"], L["kfsed"], L["m"], L["n"], L["kd"], L["kdsed"], L["g1"], L["g2"], L["p"])
    f.set_u(L["u"])
    f.set_bc(L["ibc"])
    if "marine" in L:
        f.set_marine_parameters(*L["marine"])
    if "fmix" in L:
        f.init_f(L["fmix"])
    out = {}
    last = max(steps_to_save)
    fields = ["h", "b", "etot", "erate", "area", "lake_depth"] + (["fmix"] if "marine" in L else [])
    for s in range(1, last + 1):
        f.execute_step()
        if s in steps_to_save:
            for k in fields:
                out[f"{k}_{s}"] = f.field(k)
    return out


def inputs(L):
    return {"in_" + k: np.asarray(v) for k, v in L.items()}


def main():
    # the reference's own CPU case: 101x101, dx = 1 km, D8, 200 steps
    L = landscape(101, 101, seed=0)
    np.savez_compressed(os.path.join(HERE, "plain_101_d8.npz"), **inputs(L), **run_ref(L, {1, 10, 50, 200}))

    # short runs across boundary conditions, exponents and deposition (single flow direction)
    cases = {
        "sfd_1010_n1": dict(ibc=1010),
        "sfd_0101_n15": dict(ibc=101, n=1.5, m=0.675),
        "sfd_1000_n2": dict(ibc=1000, n=2.0, m=0.9),
        "sfd_1111_g1": dict(ibc=1111, g=1.0),
        "mfd_1111_p1_g1": dict(ibc=1111, g=1.0, p=1.0, kfsed=1e-5, kdsed=2e-2),
        "mfd_1010_p1": dict(ibc=1010, p=1.0),
        "mfd_1111_pneg": dict(ibc=1111, p=-1.0, n=1.5, m=0.675),
    }
    for name, kw in cases.items():
        L = landscape(47, 31, seed=len(name), dx=250.0, kf=2e-6, dt=2e3, sinus=(10.0, 4000.0), kd=3e-2, **kw)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **inputs(L), **run_ref(L, {1, 5, 20}))

    # routing snapshots: receivers, donors and stack after LocalMinima on fixed inputs
    f = O.FortranRef()
    snaps = {}
    for ibc in (1111, 1010, 101, 1000, 11):
        L = landscape(33, 27, seed=ibc, dx=500.0, sinus=(6.0, 5000.0), ibc=ibc)
        dx = L["xl"] / 32
        dy = L["yl"] / 26
        rec, length, ndon, don, stack = f.routing(L["h"], 33, 27, dx, dy, ibc)
        snaps.update({f"h_{ibc}": L["h"], f"rec_{ibc}": rec, f"length_{ibc}": length, f"ndon_{ibc}": ndon,
                      f"don_{ibc}": don, f"stack_{ibc}": stack})
    np.savez_compressed(os.path.join(HERE, "routing_33x27.npz"), **snaps)

    # the Numerical Recipes edge sort on tie-heavy keys
    rng = np.random.default_rng(5)
    sorts = {}
    for i, n in enumerate([1, 2, 7, 8, 9, 50, 300, 2500]):
        a = rng.integers(0, max(2, n // 4), n).astype(float)
        a[rng.random(n) < 0.1] = -1e10
        sorts[f"keys_{i}"] = a
        sorts[f"perm_{i}"] = f.indexx(a)
    np.savez_compressed(os.path.join(HERE, "indexx.npz"), **sorts)


def main_tol8():
    """The ragged-batch workload's Newton tolerance (1e-8) against the reference built with that
    tolerance: single and multiple flow direction, n = 1.5 and 2, periodic and fixed boundaries."""
    cases = {
        "tol8_sfd_1010_n15": dict(ibc=1010, n=1.5, m=0.675),
        "tol8_sfd_1111_n2": dict(ibc=1111, n=2.0, m=0.9),
        "tol8_mfd_1010_n2": dict(ibc=1010, n=2.0, m=0.9, p=1.0),
        "tol8_mfd_1111_n15": dict(ibc=1111, n=1.5, m=0.675, p=1.0),
    }
    for name, kw in cases.items():
        L = landscape(47, 31, seed=len(name) + 100, dx=250.0, kf=2e-6, dt=2e3, sinus=(10.0, 4000.0), kd=3e-2, **kw)
        L["newton_tol"] = 1e-8
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **inputs(L), **run_ref(L, {1, 5, 20}, O.REF8_LIB))
    L = landscape(97, 73, seed=808, dx=250.0, kf=3e-6, dt=2e3, sinus=(10.0, 9000.0), kd=1e-2, ibc=1010, n=2.0,
                  m=0.9)
    L["newton_tol"] = 1e-8
    np.savez_compressed(os.path.join(HERE, "tol8_sfd_1010_n2_97x73.npz"), **inputs(L),
                        **run_ref(L, {1, 30}, O.REF8_LIB))


def main_marine():
    """Landscapes with a sea (the southern half 1000 m down) and the marine transport: single and multiple
    flow, n = 1, 1.5, 2, with and without deposition, several boundary conditions, sea levels, porosities
    and initial silt fractions; and the silt / sand diffusion alone on seeded inputs."""
    base = dict(dx=1000.0, kf=1e-5, dt=1e3, kd=1e-2)
    cases = {
        "marine_mfd_1000_p1_g1": dict(ibc=1000, p=1.0, n=2.0, m=0.8, g=1.0, kfsed=1e-5, kdsed=1e-2),
        "marine_sfd_1000_n1": dict(ibc=1000),
        "marine_sfd_1111_n15_g1": dict(ibc=1111, n=1.5, m=0.6, g=1.0,
                                       marine=(0.0, 0.3, 0.45, 1e3, 2e3, 0.3, 50.0, 2e2, 1e2)),
        "marine_mfd_1010_pneg": dict(ibc=1010, p=-1.0, fmix=np.random.default_rng(9).random(41 * 61)),
        "marine_sfd_0101_n2_sl": dict(ibc=101, n=2.0, m=0.8, marine=(0.5, 0.1, 0.2, 1e3, 1e3, 0.6, 1e2, 2e2, 1e2)),
    }
    for name, kw in cases.items():
        L = shelf(41, 61, seed=len(name) + 300, **dict(base, **kw))
        # multiple flow with deposition fills lakes flat; from the second step the reference then reads its
        # multiple-flow water level at exact flats before setting it (src/StreamPowerLaw.f90:22,98-100), so
        # that run is kept to its first step
        steps = {1} if name == "marine_mfd_1000_p1_g1" else {1, 5, 20}
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **inputs(L), **run_ref(L, steps))
    f = O.FortranRef()
    vec = {}
    for i, (nx, ny, ibc) in enumerate([(11, 13, 1000), (31, 23, 1111), (20, 20, 1010), (40, 30, 0)]):
        rng = np.random.default_rng(40 + i)
        h = rng.random(nx * ny) * 5.0 - 1000.0
        h[: nx * ny // 3] += 1000.0 + rng.random(nx * ny // 3)
        fr = rng.random(nx * ny)
        q1 = np.where(rng.random(nx * ny) < 0.3, rng.random(nx * ny) * 1e-2, 0.0)
        q2 = q1 * 0.7
        args = dict(nx=nx, ny=ny, dx=1e3, dy=1.5e3, dt=1e3, sealevel=0.0, layer=1e2, kdsea1=3e2, kdsea2=1.5e2,
                    ibc=ibc)
        hr, fo, it = f.silt_sand(h, fr, q1, q2, **args)
        vec.update({f"{k}_{i}": np.asarray(v) for k, v in dict(h=h, f=fr, q1=q1, q2=q2, h_out=hr, f_out=fo,
                                                                   niter=it, **args).items()})
    np.savez_compressed(os.path.join(HERE, "silt_sand.npz"), **vec)


def _strati_files(args):
    """Run a landscape with stratigraphy in a scratch directory (one process per run: the reference's
    module state); returns {file name: bytes} of what it wrote into VTK/."""
    import tempfile
    L, sched, steps = args
    d = tempfile.mkdtemp()
    os.chdir(d)
    f = O.FortranRef()
    f.setup(L["nx"], L["ny"])
    f.set_xl_yl(L["xl"], L["yl"])
    f.set_dt(L["dt"])
    f.init_h(L["h"])
    f.set_erosional_parameters(L["kf"], L["kfsed"], L["m"], L["n"], L["kd"], L["kdsed"], L["g1"], L["g2"], L["p"])
    f.set_u(L["u"])
    f.set_bc(L["ibc"])
    if "marine" in L:
        f.set_marine_parameters(*L["marine"])
    f.strati(*sched)
    for _ in range(steps):
        f.execute_step()
    out = {}
    for n in sorted(os.listdir("VTK")):
        with open(os.path.join("VTK", n), "rb") as fh:
            out[n] = fh.read()
    return out


def main_strati():
    """Stratigraphic horizons: the reference's VTK output (Horizon*, Strati-*, CUBE) of two small runs --
    single flow with a sea and the marine transport, and single flow on land -- as byte arrays."""
    import multiprocessing as mp
    cases = {
        "strati_marine_1000": (shelf(17, 13, seed=21, dx=1000.0, kf=1e-5, dt=1e3, kd=1e-2, ibc=1000), (12, 3, 4, 10.0), 12),
        "strati_land_1111": (landscape(15, 12, seed=22, dx=500.0, sinus=(8.0, 3000.0)), (10, 2, 5, -5.0), 10),
    }
    ctx = mp.get_context("spawn")
    for name, (L, sched, steps) in cases.items():
        with ctx.Pool(1) as pool:
            files = pool.map(_strati_files, [(L, sched, steps)])[0]
        rec = {"sched": np.asarray(sched, np.float64), "steps": np.asarray(steps), "names": np.asarray(sorted(files))}
        rec.update({"file_" + n: np.frombuffer(b, np.uint8) for n, b in files.items()})
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **inputs(L), **rec)


if __name__ == "__main__":
    if sys.argv[1:] == ["tol8"]:
        main_tol8()
    elif sys.argv[1:] == ["marine"]:
        main_marine()
    elif sys.argv[1:] == ["strati"]:
        main_strati()
    else:
        main()
