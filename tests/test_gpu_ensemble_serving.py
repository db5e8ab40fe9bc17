"""The batched and served configurations at their own sizes, on the GPU.

* The 256-system ensemble (workloads.ensemble_256: 500..20 000 spheres each,
  1.38 M in all, friction and restitution drawn per system), mixed precision
  as configured: 30 batched steps leave every system bitwise in the state it
  reaches alone; the same in FP64 for 200 steps, and a sample of members (the
  smallest, the largest and two drawn ones) then equals the oracle's 200 steps
  from the same start within the trajectory bars of the parity tests
  (positions 1e-11, velocities 1e-9 of their largest value, contact sets
  identical); and the mixed forces of members after 2 000 batched steps (the
  lattices have closed their gaps into contacts by then) against the
  oracle's FP64 forces of the same states
  within the mixed bar (1e-5 of each particle's force scale, torques 1e-5 of
  its sum of r |F_T|).
* The Poisson serving stream (workloads.serving_trace, sizes log-uniform in
  [1 000, 500 000]): the first 12 requests of the trace, step counts capped
  at 60, served in real time with continuous batching under a footprint
  budget that makes requests wait; each request's final state is bitwise
  the state it reaches alone.
"""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import serve, workloads
from dem3d_amd.native import DEM_FP64, DEM_MIXED, Batch, System
from checkers import oracle_system
from helpers import accel_tolerance_ok, force_scale, upload_same

pytestmark = pytest.mark.gpu


def _members(steps=30):
    return workloads.ensemble_256(steps=steps)


def _build(hip, members, precision):
    out = []
    for w in members:
        s = System(hip, w.red, precision=precision)
        s.upload(w.comp, w.x, w.v, w.w)
        s.startup()
        out.append(s)
    return out


def _xvw(s):
    st = s.download()
    return st["x"].copy(), st["v"].copy(), st["w"].copy()


@pytest.mark.parametrize("precision,nsteps", [(DEM_MIXED, 30), (DEM_FP64, 200)])
def test_ensemble_256_batched_equals_alone(hip, precision, nsteps):
    members = _members()
    assert len(members) == 256 and sum(w.np for w in members) > 1_000_000
    systems = _build(hip, members, precision)
    b = Batch(hip)
    for s in systems:
        b.add(s)
    done, esc = b.step(nsteps)
    assert np.all(done == nsteps) and not np.any(esc)
    batched = [_xvw(s) for s in systems]
    keys = [s.download_contacts()["keys"] for s in systems]
    b.close()
    for s in systems:
        s.close()
    alone = _build(hip, members, precision)
    for k, (s, bs, bk) in enumerate(zip(alone, batched, keys)):
        s.step(nsteps)
        a = _xvw(s)
        for q in range(3):
            assert np.array_equal(a[q], bs[q]), (k, q)
        assert np.array_equal(s.download_contacts()["keys"], bk), k
        s.close()
    if precision != DEM_FP64:
        return
    # a sample of members against the oracle, from the same start
    sizes = np.array([w.np for w in members])
    rng = np.random.default_rng(4)
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)),
                     *map(int, rng.choice(len(members), 2, replace=False))})
    for k in sample:
        w = members[k]
        o = oracle_system(w.red)
        o.upload(w.comp, w.x, w.v, w.w)
        o.startup()
        o.step(nsteps)
        so = o.download()
        for q, (name, rel) in enumerate((("x", 1e-11), ("v", 1e-9), ("w", 1e-9))):
            scale = np.abs(so[name]).max() + 1e-300
            assert np.abs(batched[k][q] - so[name]).max() <= rel * scale, (k, name)
        assert np.array_equal(o.download_contacts()["keys"], keys[k]), k
        o.close()


def test_ensemble_members_mixed_forces_against_oracle(hip):
    """Mixed-precision forces at the batched states of a sample of members,
    against the oracle's FP64 forces of the same states (mixed bar)."""
    members = _members()
    sizes = np.array([w.np for w in members])
    sample = [int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 101]
    sub = [members[k] for k in sample]
    systems = _build(hip, sub, DEM_MIXED)
    b = Batch(hip)
    for s in systems:
        b.add(s)
    b.step(2000)
    b.close()
    stored = 0
    for w, s in zip(sub, systems):
        st = s.download()
        st["cell_rank"] = None
        con = s.download_contacts()
        stored += len(con["keys"])
        m, o = System(hip, w.red, precision=DEM_MIXED), oracle_system(w.red)
        upload_same([m, o], w, state={k: st[k] for k in ("comp", "x", "v", "w")},
                    contacts=con)
        for q in (m, o):
            q.compute_accelerations(0, False)
        sm, so = m.download(), o.download()
        fs, ts = force_scale(o)
        c = w.comp - 1
        ok_a, ra = accel_tolerance_ok(sm["a"], so["a"], fs, w.red.inv_mass[c], 1e-5)
        ok_b, rb = accel_tolerance_ok(sm["alpha"], so["alpha"], ts + 1e-300,
                                      w.red.inv_inertia[c], 1e-5)
        print(f"member of {w.np} with {len(con['keys'])} stored pairs: force error {ra:.3g}, "
              f"torque error {rb:.3g} of the bar", flush=True)
        assert ok_a and ok_b, (w.np, ra, rb)
        for q in (m, o, s):
            q.close()
    assert stored > 100


def test_serving_trace_prefix_same_alone_and_in_the_mix(hip):
    trace = workloads.serving_trace()[:12]
    for q in trace:
        q.steps = min(q.steps, 60)
    assert max(q.n for q in trace) > 100_000

    def keep(r):
        st = r.system.download()
        return {k: st[k].copy() for k in ("x", "v", "w")}

    # room for about a third of the trace's footprint at once
    budget = sum(serve.request_bytes(hip, q, DEM_MIXED) for q in trace) // 3
    rep = serve.serve(hip, trace, precision=DEM_MIXED, budget_bytes=budget, chunk=25,
                      realtime=True, keep=keep)
    assert rep.max_resident >= 2
    assert any(r.admitted - r.request.arrival > 0.0 for r in rep.records)
    for r in rep.records:
        s = serve.build_system(hip, r.request, DEM_MIXED)
        s.step(r.request.steps)
        st = s.download()
        s.close()
        for k in ("x", "v", "w"):
            assert np.array_equal(r.result[k], st[k]), (r.index, k)
        assert r.steps_done == r.request.steps
    summ = rep.summary()
    assert summ["particle_steps"] == sum(q.n * q.steps for q in trace)
