"""Particle-number changes on the GPU path: dem_pour and dem_remove_escaped
(host-side edits of the resident state, re-uploaded with the cell state of
the last evaluation) against the oracle, itself pinned bit for bit to the
reference by tests/test_dynamic.py."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import RngT, System
from checkers import oracle_system

pytestmark = pytest.mark.gpu


def _same(g, o, rel=1e-8):
    assert g.np == o.np
    sg, so = g.download(), o.download()
    assert np.array_equal(sg["comp"], so["comp"])
    for k in ("x", "v", "w"):
        scale = np.abs(so[k]).max() + 1e-300
        err = np.abs(sg[k] - so[k]).max() / scale
        assert err <= rel, (k, err)
    assert np.array_equal(g.download_cells(), o.download_cells())
    rg, ro = np.zeros(g.np, np.int32), np.zeros(o.np, np.int32)
    ip = C.POINTER(C.c_int32)
    assert g.lib.download_cell_rank(g.h, g.np, rg.ctypes.data_as(ip)) == 0
    assert o.lib.download_cell_rank(o.h, o.np, ro.ctypes.data_as(ip)) == 0
    assert np.array_equal(rg, ro)
    cg, co = g.download_contacts(), o.download_contacts()
    assert np.array_equal(cg["keys"], co["keys"])
    assert np.array_equal(cg["wkeys"], co["wkeys"])
    assert g.stats()["ncol"] == o.stats()["ncol"]


def test_pouring_matches_oracle(hip):
    wl = workloads.ball_bearings(n=0)
    red = wl.red
    g, o = System(hip, red), oracle_system(red)
    for s in (g, o):
        s.upload(wl.comp, wl.x, wl.v, wl.w)
        s.startup()
    rg, ro = RngT.seeded(-20), RngT.seeded(-20)
    size = [0.02 / red.length_scale]
    # 7 pours: the balls bounce on the shaken cone walls, whose device
    # cos/sin differ from glibc's in the last bit, and the bouncing amplifies
    # that ~10x per 1000 steps (1e-18 after the first bounce, 5e-11 after
    # 7000 steps)
    for k in range(7):
        ng = g.pour([39], 2, size, 10, rg, red.length_scale, red.time_scale)
        no = o.pour([39], 2, size, 10, ro, red.length_scale, red.time_scale)
        assert ng == no and g.np == o.np, k
        assert (rg.dummy, rg.iix, rg.iiy, rg.iik, rg.aam) == (ro.dummy, ro.iix, ro.iiy, ro.iik, ro.aam)
        for s in (g, o):
            s.step(1000)
    assert g.np >= 5
    _same(g, o)
    assert g.stats()["time"] == o.stats()["time"]


@pytest.mark.parametrize("seed", [20240015, 20240023])
def test_removal_matches_oracle(hip, seed):
    wl = workloads.discharge(seed=seed)
    red = wl.red
    g, o = System(hip, red), oracle_system(red)
    for s in (g, o):
        s.upload(wl.comp, wl.x, wl.v, wl.w)
        s.startup()
    total, removed_g, removed_o = 400, [], []
    left = total
    while left > 0:
        done, esc = g.step(left, allow_escape=True)
        left -= done
        if esc:
            removed_g += list(g.remove_escaped())
    for s in range(total):
        o.step(1, red.method_order + s, allow_escape=True)
        removed_o += list(o.remove_escaped())
    assert removed_g == removed_o
    assert len(removed_g) >= 20
    _same(g, o)
