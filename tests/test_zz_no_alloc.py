"""dem_observables_noalloc (the entry point that allocates no device memory)
through the C ABI against dem_observables on the same states: every field and
the whole profile bitwise equal, on two calls in a row of which the second
has a smaller system and fewer layers, in both call orders; and a layer count
beyond DEM_OBS_MAX_LAYERS refused with DEM_ERR_CAPACITY."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import DemError, ObservablesT, System

pytestmark = pytest.mark.gpu

DEM_ERR_CAPACITY = -3
DEM_OBS_MAX_LAYERS = 256


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def _system(hip, **kw):
    wl = workloads.small_box(**kw)
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    g.step(25)
    return g


def test_noalloc_equals_allocating_observables(hip):
    big = _system(hip, n_side=(14, 13, 12), nc=3, seed=31, vel=0.4)
    small = _system(hip, n_side=(5, 4, 6), nc=2, seed=32, vel=0.2)
    # first call: the larger system, many layers (noalloc first, so its
    # array has held nothing of this run before)
    n1 = big.observables(DEM_OBS_MAX_LAYERS, noalloc=True)
    a1 = big.observables(DEM_OBS_MAX_LAYERS)
    _same(n1, a1)
    assert n1["pair_contacts"] > 0 and n1["profile"][:, 0].max() > 0.0
    # second call: a smaller input on the same array, fewer layers; the array
    # still holds the larger call's partial sums
    n2 = small.observables(16, noalloc=True)
    a2 = small.observables(16)
    _same(n2, a2)
    # no profile, and the other order
    a3 = small.observables(0)
    n3 = small.observables(0, noalloc=True)
    _same(n3, a3)
    assert n3["kinetic_energy"] == n2["kinetic_energy"]
    big.close()
    small.close()


def test_noalloc_refuses_too_many_layers(hip):
    g = _system(hip, n_side=(4, 4, 4), nc=1, seed=33, vel=0.1)
    o = ObservablesT()
    prof = np.zeros((DEM_OBS_MAX_LAYERS + 1, 4))
    rc = g.lib.observables_noalloc(g.h, DEM_OBS_MAX_LAYERS + 1, C.byref(o),
                                   prof.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == DEM_ERR_CAPACITY
    assert b"at most 256 layers" in g.lib.last_error(g.h)
    # the allocating entry point takes them
    d = g.observables(DEM_OBS_MAX_LAYERS + 1)
    assert d["profile"].shape == (DEM_OBS_MAX_LAYERS + 1, 4)
    with pytest.raises(DemError):
        g.observables(-1, noalloc=True)
    g.close()
