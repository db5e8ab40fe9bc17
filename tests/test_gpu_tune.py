"""Launch shapes of the force passes and the autotuner (dem_autotune,
dem_batch_autotune).

Every shape (lanes per block, occupancy target) runs the same per-lane
arithmetic and integer counter sums, so a system stepped in any shape must
reach bitwise the state of the default shape, alone and batched, and a run
that autotunes while stepping must follow bitwise the trajectory of plain
stepping."""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import DEM_FP64, DEM_MIXED, NSHAPES, Batch, System, autotune_steps

pytestmark = pytest.mark.gpu

CASES = [
    dict(n_side=(9, 8, 7), nc=2, seed=31, vel=0.4, mu=0.4, e=0.85),
    dict(n_side=(7, 7, 5), nc=3, ratio=1.4, seed=37, walls=False, pbc=(True, True, True),
         vel=0.5, spacing=1.0, mu=0.2, e=0.7),
]


def _make(hip, kw, precision=DEM_FP64):
    wl = workloads.small_box(**kw)
    s = System(hip, wl.red, precision=precision)
    s.upload(wl.comp, wl.x, wl.v, wl.w)
    s.startup()
    return s


def _state(s):
    st = s.download()
    con = s.download_contacts()
    return {k: np.array(st[k]) for k in ("x", "v", "w", "a", "alpha", "ring")} | {
        "keys": con["keys"].copy(), "ft": con["ft"].copy(), "vt": con["vt"].copy(),
        "cells": s.download_cells(), "ncol": np.array(s.stats()["ncol"]),
        "pairs": np.array(s.stats()["pair_contacts"])}


def _same(a, b, tag):
    for k in a:
        assert np.array_equal(a[k], b[k]), (tag, k)


@pytest.mark.parametrize("precision", [DEM_FP64, DEM_MIXED])
def test_every_shape_gives_the_same_trajectory(hip, precision):
    for kw in CASES:
        ref = None
        for shape in range(NSHAPES):
            for block_map in ((0, 1) if shape in (0, 1, 6) else (0,)):
                s = _make(hip, kw, precision)
                s.set_option("pass_shape", shape)
                s.set_option("block_map", block_map)
                s.step(37)
                st = _state(s)
                s.close()
                if ref is None:
                    ref = st
                else:
                    _same(st, ref, (shape, block_map))


def test_autotune_keeps_the_trajectory(hip):
    kw = CASES[0]
    a = _make(hip, kw)
    rep = a.autotune(steps_per_shape=3)
    # every shape in a sweep forward (an untimed first step, 2 timed) and
    # back (1 timed), then the best shape's block orders A B B A (2, 1 + 2,
    # 1, 1 steps)
    assert rep["steps"] == NSHAPES * 4 + 7 == autotune_steps(3)
    assert rep["block_map"] in (0, 1) and all(m > 0 for m in rep["ms_map"])
    assert rep["ms_map"][rep["block_map"]] <= rep["ms_map"][1 - rep["block_map"]]
    assert len(rep["ms"]) == NSHAPES and all(m > 0 for m in rep["ms"])
    assert rep["chosen"] == int(np.argmin(rep["ms"]))
    # lanes per block; variant bit 0 = front rule, bit 1 = staged partner
    # records (a staged shape runs 256-lane blocks, one per tile)
    assert set(rep["block"]) <= {64, 128, 256, 512} and 256 in rep["block"]
    assert set(rep["front"]) <= {0, 1, 2, 3}
    assert {f & 1 for f in rep["front"]} == {0, 1} and any(f & 2 for f in rep["front"])
    assert all(b == 256 for b, f in zip(rep["block"], rep["front"]) if f & 2)
    a.step(11)
    b = _make(hip, kw)
    b.step(rep["steps"] + 11)
    _same(_state(a), _state(b), "autotuned")
    assert a.stats()["next_step"] == b.stats()["next_step"]


def test_batch_shapes_and_autotune_equal_alone(hip):
    systems_kw = CASES + [dict(n_side=(3, 3, 3), nc=1, seed=13, vel=0.1, mu=0.1, e=0.6)]
    alone = []
    for kw in systems_kw:
        s = _make(hip, kw)
        s.step(autotune_steps(2) + 9 + 9)
        alone.append(_state(s))
        s.close()
    systems = [_make(hip, kw) for kw in systems_kw]
    b = Batch(hip)
    for s in systems:
        b.add(s)
    b.set_option("pass_shape", 2)
    b.set_option("block_map", 1)
    b.step(9)
    rep = b.autotune(steps_per_shape=2)
    assert rep["steps"] == NSHAPES * 3 + 5 == autotune_steps(2)
    assert rep["chosen"] == int(np.argmin(rep["ms"]))
    b.step(9)
    for s, ref in zip(systems, alone):
        _same(_state(s), ref, "batch")
    b.close()


def test_bad_shape_is_refused(hip):
    s = _make(hip, CASES[1])
    with pytest.raises(Exception, match="pass_shape"):
        s.set_option("pass_shape", NSHAPES)
    with pytest.raises(Exception, match="pass_shape"):
        s.set_option("pass_shape", 1.5)
    with pytest.raises(Exception, match="block_map"):
        s.set_option("block_map", 2)
    s.close()


def test_staged_shapes_with_few_halo_rows(hip):
    """The LDS-staged shapes read the partners beyond a tile's staged halo
    rows from memory: with 4 rows most of them do, and the trajectory is
    still bitwise the default shape's."""
    kw = CASES[0]
    ref = _make(hip, kw)
    ref.step(12)
    for rows in (4, 512):
        s = _make(hip, kw)
        s.set_option("pass_shape", 4)
        s.set_option("halo_rows", rows)
        s.step(12)
        _same(_state(s), _state(ref), rows)
        st = s.stats()
        assert st["max_halo"] > 4
        assert (st["halo_overflow"] > 0) == (rows == 4)
    with pytest.raises(Exception):
        ref.set_option("halo_rows", 0)
