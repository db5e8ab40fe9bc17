"""Batches: many independent systems stepped with one launch per phase.

Each system of a batch runs the same kernels with the same arguments as when
it runs alone, so its state after N batched steps must be bitwise the one
it reaches alone -- for systems of different sizes, components, friction
and restitution, in FP64 and in the mixed mode, with systems joining and
leaving between calls and one system stopping on an escape."""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import DEM_FP64, DEM_MIXED, Batch, System

pytestmark = pytest.mark.gpu

SYSTEMS = [
    dict(n_side=(6, 6, 6), nc=2, seed=3, vel=0.3, mu=0.5, e=0.9),
    dict(n_side=(8, 8, 5), nc=3, ratio=1.4, seed=5, walls=False, pbc=(True, True, True),
         vel=0.5, spacing=1.0, mu=0.2, e=0.7),
    dict(n_side=(5, 5, 5), nc=1, seed=9, pbc=(False, False, False), walls=True, vel=0.4,
         mu=0.6, e=0.8),
    dict(n_side=(12, 9, 4), nc=2, seed=11, vel=0.2, mu=0.3, e=0.95),
    dict(n_side=(3, 3, 3), nc=1, seed=13, vel=0.1, mu=0.1, e=0.6),
]


def _state(s: System):
    st = s.download()
    con = s.download_contacts()
    return {k: np.array(st[k]) for k in ("x", "v", "w", "a", "alpha", "ring")} | {
        "keys": con["keys"].copy(), "ft": con["ft"].copy(), "cells": s.download_cells()}


def _make(hip, kw, precision):
    wl = workloads.small_box(**kw)
    s = System(hip, wl.red, precision=precision)
    s.upload(wl.comp, wl.x, wl.v, wl.w)
    s.startup()
    return s


def _assert_same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [DEM_FP64, DEM_MIXED])
def test_batch_equals_alone(hip, precision):
    nsteps = 40
    alone = []
    for kw in SYSTEMS:
        s = _make(hip, kw, precision)
        s.step(nsteps)
        alone.append(_state(s))
        s.close()
    systems = [_make(hip, kw, precision) for kw in SYSTEMS]
    b = Batch(hip)
    for s in systems:
        b.add(s)
    done, esc = b.step(nsteps)
    assert list(done) == [nsteps] * len(SYSTEMS) and not esc.any()
    for s, ref in zip(systems, alone):
        _assert_same(_state(s), ref)
        assert s.stats()["steps"] == nsteps
    b.close()


def test_batch_membership_changes_between_calls(hip):
    """Two calls of 15 steps with a system leaving and one joining in between
    give, for every system, the state of the same number of steps alone."""
    def alone(kw, n):
        s = _make(hip, kw, DEM_FP64)
        s.step(n)
        st = _state(s)
        s.close()
        return st

    refs = {0: alone(SYSTEMS[0], 30), 1: alone(SYSTEMS[1], 15), 2: alone(SYSTEMS[2], 30),
            3: alone(SYSTEMS[3], 15)}
    s0, s1, s2 = (_make(hip, SYSTEMS[k], DEM_FP64) for k in (0, 1, 2))
    b = Batch(hip)
    for s in (s0, s1, s2):
        b.add(s)
    b.step(15)
    b.remove(s1)
    s3 = _make(hip, SYSTEMS[3], DEM_FP64)
    b.add(s3)
    b.step(15)
    _assert_same(_state(s0), refs[0])
    _assert_same(_state(s1), refs[1])
    _assert_same(_state(s2), refs[2])
    _assert_same(_state(s3), refs[3])
    # a removed system steps on its own again
    s1.step(5)
    b.close()


def test_batch_escape_stops_only_that_system(hip):
    wl = workloads.small_box(n_side=(5, 5, 4), nc=1, seed=2, walls=True, vel=0.0)
    x = wl.x.copy()
    x[2, 0] = 0.5 * wl.red.lbox[2] - 0.01
    v = wl.v.copy()
    v[:, 0] = 0.0
    v[2, 0] = 1.0
    esc_sys = System(hip, wl.red)
    esc_sys.upload(wl.comp, x, v, wl.w)
    esc_sys.startup()
    ref = System(hip, wl.red)
    ref.upload(wl.comp, x, v, wl.w)
    ref.startup()
    done_ref, esc_ref = ref.step(50, allow_escape=True)
    other = _make(hip, SYSTEMS[0], DEM_FP64)
    other_alone = _make(hip, SYSTEMS[0], DEM_FP64)
    other_alone.step(50)
    b = Batch(hip)
    b.add(other)
    b.add(esc_sys)
    done, esc = b.step(50, allow_escape=True)
    assert done[0] == 50 and esc[0] == 0
    assert done[1] == done_ref < 50 and esc[1] == esc_ref == 1
    _assert_same(_state(other), _state(other_alone))
    _assert_same(_state(esc_sys), _state(ref))
    b.close()


def test_step_on_batched_system_is_refused(hip):
    s = _make(hip, SYSTEMS[4], DEM_FP64)
    b = Batch(hip)
    b.add(s)
    with pytest.raises(Exception):
        s.step(1)
    s.close()          # destroying a batched system takes it out of the batch
    assert b.lib.batch_size(b.b) == 0
    b.close()


def test_batch_with_moving_walls_equals_alone(hip):
    """Systems with moving walls (the shaken example dish, a shaken floor)
    batched with a static one: two calls in a row give each the state, wall
    positions and Time of the same two calls alone."""
    from helpers import shaken_floor

    def make(kind):
        wl = workloads.ball_bearings() if kind == 0 else (
            shaken_floor() if kind == 1 else workloads.small_box(**SYSTEMS[0]))
        s = System(hip, wl.red)
        s.upload(wl.comp, wl.x, wl.v, wl.w)
        s.startup()
        return s

    def snap(s):
        return _state(s) | {"walls": s.download_walls()["position"].copy(),
                            "wvel": s.download_walls()["velocity"].copy(),
                            "time": np.array(s.stats()["time"])}

    refs = []
    for kind in range(3):
        s = make(kind)
        s.step(120)
        s.step(80)
        refs.append(snap(s))
        s.close()
    systems = [make(kind) for kind in range(3)]
    b = Batch(hip)
    for s in systems:
        b.add(s)
    b.step(120)
    b.step(80)
    for s, ref in zip(systems, refs):
        _assert_same(snap(s), ref)
        assert s.stats()["next_step"] == s.red.method_order + 200
    b.close()


@pytest.mark.parametrize("precision", [DEM_FP64, DEM_MIXED])
def test_graph_replays_equal_launches(hip, precision):
    """The batched step replayed from captured HIP graphs (one step per
    graph, and 8 steps per graph with the remainder one by one) gives after
    every call the state the same launches give one by one.  The calls
    interleave the batch's own eager work (table upload, escape reset,
    control read-back, synchronisation) with the replays, a system leaves
    between calls (the graphs are captured again) and one timed call replays
    the one-step graph between events."""
    def run(graph, graph_steps):
        systems = [_make(hip, kw, precision) for kw in SYSTEMS]
        b = Batch(hip)
        b.set_option("graph", graph)
        b.set_option("graph_steps", graph_steps)
        for s in systems:
            b.add(s)
        snaps = []
        for n in (21, 8, 3):
            b.step(n)
            snaps.append([_state(s) for s in systems])
        b.remove(systems[3])
        b.set_timing(True)
        b.step(11)
        b.set_timing(False)
        snaps.append([_state(s) for s in systems])
        b.step(17)
        snaps.append([_state(s) for s in systems])
        b.close()
        for s in systems:
            s.close()
        return snaps

    ref = run(0, 1)
    for graph_steps in (1, 8):
        got = run(1, graph_steps)
        for call, (sa, sb) in enumerate(zip(got, ref)):
            for k, (a, b_) in enumerate(zip(sa, sb)):
                for key in a:
                    assert np.array_equal(a[key], b_[key]), (graph_steps, call, k, key)
