"""Persistent batched stepping (dem_batch_set_option "persistent"): one
cooperative launch runs many steps of every system, each system stepped by
its own group of blocks with group barriers between the phases of a step.
The kernel bodies and their order are those of the launch-per-phase path,
so every system must reach bitwise the state it reaches alone: FP64 and
mixed, several launches per call (persist_steps smaller than the call),
rebuilds, moving walls, an escape that stops one system, stale pairs (the
merge of the stale-pair table in scratch memory), and the 256-system
ensemble against the launch-per-phase batch.  A tiny spin limit forced
from the host makes the barrier waits give up: the call must return an
error, and the GPU stays usable."""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import DEM_FP64, DEM_MIXED, Batch, System
from helpers import stale_pair_case
from test_gpu_batch import SYSTEMS, _assert_same, _make, _state

pytestmark = pytest.mark.gpu


def _persistent(hip, steps_per_launch=16):
    b = Batch(hip)
    b.set_option("persistent", 1)
    b.set_option("persist_steps", steps_per_launch)
    return b


@pytest.mark.parametrize("precision", [DEM_FP64, DEM_MIXED])
def test_persistent_equals_alone(hip, precision):
    nsteps = 40
    alone = []
    for kw in SYSTEMS:
        s = _make(hip, kw, precision)
        s.step(nsteps)
        alone.append(_state(s))
        s.close()
    systems = [_make(hip, kw, precision) for kw in SYSTEMS]
    b = _persistent(hip)
    for s in systems:
        b.add(s)
    done, esc = b.step(25)
    assert list(done) == [25] * len(SYSTEMS) and not esc.any()
    done, esc = b.step(nsteps - 25)
    assert list(done) == [nsteps - 25] * len(SYSTEMS)
    for s, ref in zip(systems, alone):
        _assert_same(_state(s), ref)
        assert s.stats()["steps"] == nsteps
    assert sum(s.stats()["rebuilds"] for s in systems) > 0
    b.close()


def test_persistent_escape_and_moving_walls(hip):
    from helpers import shaken_floor
    wl = workloads.small_box(n_side=(5, 5, 4), nc=1, seed=2, walls=True, vel=0.0)
    x = wl.x.copy()
    x[2, 0] = 0.5 * wl.red.lbox[2] - 0.01
    v = wl.v.copy()
    v[:, 0] = 0.0
    v[2, 0] = 1.0

    def esc_system():
        s = System(hip, wl.red)
        s.upload(wl.comp, x, v, wl.w)
        s.startup()
        return s

    def shaken():
        w = shaken_floor()
        s = System(hip, w.red)
        s.upload(w.comp, w.x, w.v, w.w)
        s.startup()
        return s

    ref_esc = esc_system()
    done_ref, esc_ref = ref_esc.step(50, allow_escape=True)
    ref_shaken = shaken()
    ref_shaken.step(50)
    ref_other = _make(hip, SYSTEMS[0], DEM_FP64)
    ref_other.step(50)
    systems = [_make(hip, SYSTEMS[0], DEM_FP64), esc_system(), shaken()]
    b = _persistent(hip, steps_per_launch=7)
    for s in systems:
        b.add(s)
    done, esc = b.step(50, allow_escape=True)
    assert done[0] == 50 and done[2] == 50 and done[1] == done_ref < 50
    assert esc[1] == esc_ref == 1 and esc[0] == 0 and esc[2] == 0
    _assert_same(_state(systems[0]), _state(ref_other))
    _assert_same(_state(systems[1]), _state(ref_esc))
    _assert_same(_state(systems[2]), _state(ref_shaken))
    assert np.array_equal(systems[2].download_walls()["position"],
                          ref_shaken.download_walls()["position"])
    b.close()


def test_persistent_stale_pairs_equal_alone(hip):
    cases = [stale_pair_case(k) for k in ("separate", "return")]
    nsteps = min(c[2] for c in cases)

    def load(wl, con):
        s = System(hip, wl.red)
        s.upload(wl.comp, wl.x, wl.v, wl.w)
        s.upload_contacts(con["keys"], con["ft"], con["vt"], con["wkeys"], con["wft"], con["wvt"])
        return s

    alone = []
    for wl, con, _ in cases:
        s = load(wl, con)
        s.step(nsteps)
        alone.append((s.download(), s.download_contacts()))
        s.close()
    systems = [load(wl, con) for wl, con, _ in cases]
    b = _persistent(hip, steps_per_launch=64)
    for s in systems:
        b.add(s)
    b.step(nsteps)
    for s, (st, c) in zip(systems, alone):
        sb = s.download()
        for q in ("x", "v", "w"):
            assert np.array_equal(sb[q], st[q]), q
        cb = s.download_contacts()
        assert np.array_equal(cb["keys"], c["keys"]) and np.array_equal(cb["ft"], c["ft"])
    b.close()


def test_persistent_ensemble_256_equals_launch_per_phase(hip):
    members = workloads.ensemble_256()

    def build():
        out = []
        for w in members:
            s = System(hip, w.red, precision=DEM_MIXED)
            s.upload(w.comp, w.x, w.v, w.w)
            s.startup()
            out.append(s)
        return out

    def run(persistent):
        systems = build()
        b = _persistent(hip, steps_per_launch=10) if persistent else Batch(hip)
        for s in systems:
            b.add(s)
        done, esc = b.step(30)
        assert np.all(done == 30) and not np.any(esc)
        out = [_state(s) for s in systems]
        b.close()
        for s in systems:
            s.close()
        return out

    ref = run(False)
    got = run(True)
    for k, (a, r) in enumerate(zip(got, ref)):
        for q in a:
            assert np.array_equal(a[q], r[q]), (k, q)


def test_persistent_barrier_wait_gives_up(hip):
    """spin_limit 0: a block that reaches a group barrier before the others
    gives up at once; the call reports the failure instead of waiting."""
    systems = [_make(hip, kw, DEM_FP64) for kw in SYSTEMS]
    b = _persistent(hip)
    b.set_option("spin_limit", 0)
    for s in systems:
        b.add(s)
    with pytest.raises(Exception, match="barrier"):
        b.step(10)
    b.close()
    for s in systems:
        s.close()
    # the GPU still steps
    s = _make(hip, SYSTEMS[0], DEM_FP64)
    s.step(5)
    s.close()
