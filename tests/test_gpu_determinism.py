"""With deposition (g > 0) a landscape's results are bitwise the same alone, in a batch with others, in a
lane split and from one run to the next: the lake water volumes (src/StreamPowerLaw.f90:323-324, summed per
lake sill) are added in exact fixed point, and every other sum of the step has a fixed order."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape

pytestmark = pytest.mark.gpu

STEPS = 12


def _cases():
    return [landscape(83, 71, seed=61, g=1.0, sinus=(8.0, 9000.0)),
            landscape(64, 77, seed=62, g=1.0, n=1.5, m=0.675, ibc=1010, sinus=(6.0, 7000.0)),
            landscape(70, 70, seed=63, g=1.0, p=1.0, sinus=(8.0, 8000.0)),
            landscape(59, 81, seed=64, g=0.5, p=1.0, n=2.0, m=0.9, ibc=1010, sinus=(5.0, 6000.0))]


def _fields(b, k):
    return [b.field(k, f) for f in (0, 1, 4)]  # h, b, etot


def _alone(fs, L):
    with fs.Batch([(L["nx"], L["ny"])]) as b:
        apply(b, L, 0)
        b.step(STEPS)
        return _fields(b, 0)


def test_deposition_results_do_not_depend_on_company_or_run():
    import fastscapelib_fortran as fs
    Ls = _cases()
    print(f"determinism: {len(Ls)} landscapes with g > 0, {STEPS} steps", flush=True)
    alone = [_alone(fs, L) for L in Ls]
    again = [_alone(fs, L) for L in Ls]
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        b.step(STEPS)
        batch = [_fields(b, k) for k in range(len(Ls))]
    with fs.Lanes([(L["nx"], L["ny"]) for L in Ls], 2) as b:
        for k, L in enumerate(Ls):
            apply(b, L, k)
        b.step(STEPS)
        lanes = [_fields(b, k) for k in range(len(Ls))]
    for k in range(len(Ls)):
        for name, other in (("again", again), ("batch", batch), ("lanes", lanes)):
            for a, o in zip(alone[k], other[k]):
                assert np.array_equal(a, o), (k, name, float(np.max(np.abs(a - o))))
