"""The serving loop: requests of different sizes, step counts and contact
parameters admitted while others run, stepped together in chunks, retired
when done.  Each request's final state is bitwise the state it reaches when
run alone."""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import serve, workloads
from dem3d_amd.native import DEM_FP64, DEM_MIXED

pytestmark = pytest.mark.gpu


def _trace():
    rng = np.random.default_rng(7)
    out = []
    for k in range(6):
        out.append(workloads.Request(arrival=0.01 * k, n=int(rng.integers(300, 2500)),
                                     steps=int(rng.integers(30, 140)),
                                     mu=float(rng.uniform(0.1, 0.7)),
                                     e=float(rng.uniform(0.6, 0.95)),
                                     seed=int(rng.integers(0, 2 ** 31 - 1))))
    return out


def _keep(r):
    st = r.system.download()
    return {"x": st["x"].copy(), "v": st["v"].copy(), "w": st["w"].copy()}


@pytest.mark.parametrize("precision", [DEM_FP64, DEM_MIXED])
def test_each_request_same_alone_and_in_the_mix(hip, precision):
    trace = _trace()
    # a small footprint budget (the two largest requests) forces some
    # requests to wait for others
    need = [serve.request_bytes(hip, q, precision) for q in trace]
    budget = sum(sorted(need)[-2:])
    rep = serve.serve(hip, trace, precision=precision, budget_bytes=budget, chunk=25,
                      realtime=False, keep=_keep)
    assert rep.max_resident >= 2
    assert rep.peak_bytes <= budget
    for r in rep.records:
        s = serve.build_system(hip, r.request, precision)
        # admission counted what the library allocates for the system
        assert s.stats()["device_bytes"] == r.bytes == serve.request_bytes(hip, r.request, precision)
        s.step(r.request.steps)
        st = s.download()
        s.close()
        for k in ("x", "v", "w"):
            assert np.array_equal(r.result[k], st[k]), (r.index, k)
        assert r.finished >= r.started >= r.admitted
    summ = rep.summary()
    assert summ["requests"] == len(trace)
    assert summ["particle_steps"] == sum(q.n * q.steps for q in trace)
