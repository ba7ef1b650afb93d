"""Lanes: landscapes split over several contexts stepping on their own streams give the results of
one Batch holding them all (D8 and multiple flow, different sizes, exponents and BCs): bitwise, except
for a landscape with deposition (g > 0), whose lake volumes are summed with device atomics in the
order the hardware runs them (h within the parity bar, 1e-10 relative)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape, rel_err

pytestmark = pytest.mark.gpu


def _set():
    return [landscape(97, 81, seed=51, sinus=(5.0, 8000.0)), landscape(64, 90, seed=52, ibc=1010, n=1.5, m=0.675),
            landscape(50, 50, seed=53, p=1.0), landscape(120, 40, seed=54, g=1.0, kfsed=1e-5),
            landscape(33, 70, seed=55, p=-1.0), landscape(75, 75, seed=56, n=2.0, m=0.8)]


@pytest.mark.parametrize("lanes", [2, 3])
def test_lanes_equal_one_batch(lanes):
    import fastscapelib_fortran as fs
    Ls = _set()
    print(f"lanes={lanes}: {len(Ls)} landscapes, 6 steps", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as a, fs.Lanes([(L["nx"], L["ny"]) for L in Ls], lanes) as b:
        assert len(b.lanes) == lanes
        assert sorted(k for m in b.members for k in m) == list(range(len(Ls)))
        for k, L in enumerate(Ls):
            apply(a, L, k)
            apply(b, L, k)
        a.step(3)
        b.step(3)
        b.step(3, sync=False)
        b.synchronize()
        a.step(3)
        for k, L in enumerate(Ls):
            assert b.get_step(k) == a.get_step(k) == 6
            if L["g1"] > 0.0:  # deposition: atomics order the lake-volume sums (the parity bar for h)
                assert rel_err(b.h(k), a.h(k)) <= 1e-10, k
                continue
            assert np.array_equal(a.h(k), b.h(k)), k
            assert np.array_equal(a.rec(k), b.rec(k)), k
            assert np.array_equal(a.field(k, 6), b.field(k, 6)), k


def test_lanes_split_balances_cells():
    import fastscapelib_fortran as fs
    sizes = [(200, 200), (100, 100), (100, 100), (100, 100), (100, 100), (50, 50)]
    with fs.Lanes(sizes, 2) as b:
        loads = [sum(sizes[k][0] * sizes[k][1] for k in m) for m in b.members]
        assert max(loads) - min(loads) <= 50 * 50
