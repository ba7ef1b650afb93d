"""The launch-shape autotuner picks shapes from its candidate sets, leaves the landscapes as they were,
and steps taken with the tuned shapes (graph-replayed) equal steps with the default ones bitwise."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import apply, landscape

pytestmark = pytest.mark.gpu

CANDIDATES = {"walk_gx": {64, 128, 256, 512}, "jump_blocks": {1024, 2048, 4096, 8192}, "tour_gx": {8, 16, 32},
              "nr_small_gx": {16, 32, 64, 128}, "area_wg": {256, 512, 1024}, "bor_gx": {2, 4, 8, 16},
              "flow_wg": {256, 512, 1024}, "flow_split": {0, 262144, 131072, 65536},
              "walk_big": {64, 256, 1024, 8192}, "mfd_split": {262144, 131072, 65536, 32768},
              "tile_h": {4, 8, 16, 32}, "flow_nk": {4, 8, 16, 32}}


def test_autotune_changes_shapes_not_results():
    import fastscapelib_fortran as fs
    Ls = [landscape(97, 81, seed=31, sinus=(5.0, 8000.0)), landscape(64, 90, seed=32, ibc=1010, n=1.5, m=0.675),
          landscape(50, 50, seed=33, p=1.0)]
    print("autotune: 3 landscapes, 4 steps", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as a, fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(a, L, k)
            apply(b, L, k)
        a.step(2)
        b.step(2)
        h_before = [b.h(k) for k in range(len(Ls))]
        chosen = b.autotune(reps=1)
        assert set(chosen) == set(CANDIDATES)
        for name, v in chosen.items():
            assert v in CANDIDATES[name], (name, v)
        assert b.tuning() == chosen
        for k in range(len(Ls)):
            assert np.array_equal(b.h(k), h_before[k])  # tuning leaves the elevations alone
        a.step(2)
        b.step(2)
        for k in range(len(Ls)):
            assert np.array_equal(a.h(k), b.h(k)), k
            assert np.array_equal(a.rec(k), b.rec(k)), k
            assert np.array_equal(a.field(k, 6), b.field(k, 6)), k


def test_set_tuning_roundtrip_and_bounds():
    import fastscapelib_fortran as fs
    from fastscapelib_fortran import _lib
    Ls = [landscape(70, 66, seed=41), landscape(45, 52, seed=42, ibc=1010)]
    print("set_tuning: 2 landscapes, 3 steps", flush=True)
    with fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as a, fs.Batch([(L["nx"], L["ny"]) for L in Ls]) as b:
        for k, L in enumerate(Ls):
            apply(a, L, k)
            apply(b, L, k)
        a.step(1)
        b.step(1)  # b's default-shape graphs are captured here
        # walk_big 256: every closed basin above 256 nodes takes the node-parallel pass search
        shapes = {"walk_gx": 64, "jump_blocks": 1024, "tour_gx": 32, "nr_small_gx": 16, "area_wg": 256,
                  "bor_gx": 1, "flow_wg": 256, "flow_split": 1000, "walk_big": 256, "mfd_split": 1024,
                  "tile_h": 32, "flow_nk": 9}
        b.set_tuning(shapes)
        assert b.tuning() == shapes
        with pytest.raises(_lib.FastscapeError):
            b.set_tuning({"walk_gx": 0})
        with pytest.raises(_lib.FastscapeError):
            b.set_tuning({"area_wg": 384})
        with pytest.raises(_lib.FastscapeError):
            b.set_tuning({"flow_wg": 128})
        with pytest.raises(_lib.FastscapeError):
            b.set_tuning({"walk_big": 32})
        with pytest.raises(_lib.FastscapeError):
            b.set_tuning({"tile_h": 12})
        with pytest.raises(_lib.FastscapeError):
            b.set_tuning({"flow_nk": 0})
        assert b.tuning() == shapes  # a rejected call changes nothing
        a.step(2)
        b.step(2)
        for k in range(len(Ls)):
            assert np.array_equal(a.h(k), b.h(k)), k
            assert np.array_equal(a.rec(k), b.rec(k)), k
