"""Fluid drag on the GPU path against the oracle (itself pinned bit for bit
to the reference by tests/test_fluid.py): the layer field after one
evaluation, the accelerations with drag, and a fluidized bed over the
start-up and 300 steps.  The layer sums associate differently (per block,
then over blocks) and the device pow differs from glibc's in the last bit,
hence tolerances instead of equality."""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import System
from checkers import oracle_system
from helpers import accel_tolerance_ok, force_scale

pytestmark = pytest.mark.gpu


def test_fluid_field_and_accelerations(hip):
    wl = workloads.fluidized_bed()
    g, o = System(hip, wl.red), oracle_system(wl.red)
    for s in (g, o):
        s.upload(wl.comp, wl.x, wl.v, wl.w)
        s.compute_accelerations(0, True)
    vg, eg = g.download_fluid_field()
    vo, eo = o.download_fluid_field()
    assert np.abs(eg - eo).max() <= 1e-13
    assert np.abs(vg - vo).max() <= 1e-12 * np.abs(vo).max()
    assert (eo < 1.0).sum() >= 8          # the bed occupies several layers
    sg, so = g.download(), o.download()
    fs, ts = force_scale(o)
    comp = wl.comp - 1
    ok, r = accel_tolerance_ok(sg["a"], so["a"], fs, wl.red.inv_mass[comp], 1e-12)
    assert ok, r
    # the drag is a real part of the force: off, the accelerations change
    h = System(hip, wl.red)
    h.set_fluid(None)
    h.upload(wl.comp, wl.x, wl.v, wl.w)
    h.compute_accelerations(0, True)
    assert np.abs(h.download()["a"] - sg["a"]).max() > 1e3 * np.abs(sg["a"] - so["a"]).max()


def test_fluidized_bed_trajectory(hip):
    wl = workloads.fluidized_bed()
    g, o = System(hip, wl.red), oracle_system(wl.red)
    for s in (g, o):
        s.upload(wl.comp, wl.x, wl.v, wl.w)
        s.startup()
        s.step(300)
    sg, so = g.download(), o.download()
    for k in ("x", "v", "w"):
        scale = np.abs(so[k]).max() + 1e-300
        assert np.abs(sg[k] - so[k]).max() <= 1e-8 * scale, k
    assert np.array_equal(g.download_contacts()["keys"], o.download_contacts()["keys"])
    assert g.stats()["ncol"] == o.stats()["ncol"]
