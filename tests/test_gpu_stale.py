"""Stale pairs on the GPU (the stale-pair table of the neighbour lists)
against the oracle, whose semantics tests/test_stale.py pins to the
reference's (src/MForces.f90:262-270): the two scenarios of
helpers.stale_pair_case stepped one step at a time, and after every step the
stored pairs with their (FT, VT) bit for bit, Ncol, the stored-pair count,
positions and velocities within 1e-13 of their scale, accelerations within
1e-12 of the force scale.  Then the same two systems stepped as one batch
equal the systems stepped alone, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd.native import Batch, System
from checkers import oracle_system
from helpers import accel_tolerance_ok, compare_contacts, force_scale, stale_pair_case

pytestmark = pytest.mark.gpu


def _load(s, wl, con):
    s.upload(wl.comp, wl.x, wl.v, wl.w)
    s.upload_contacts(con["keys"], con["ft"], con["vt"], con["wkeys"], con["wft"], con["wvt"])


@pytest.mark.parametrize("kind", ["separate", "return"])
def test_stale_pairs_match_oracle_every_step(hip, kind):
    wl, con, steps = stale_pair_case(kind)
    g, o = System(hip, wl.red), oracle_system(wl.red)
    for s in (g, o):
        _load(s, wl, con)
    compare_contacts(g.download_contacts(), o.download_contacts())
    stale_steps = 0
    for k in range(steps):
        for s in (g, o):
            s.step(1, k + 1)
        cg, co = g.download_contacts(), o.download_contacts()
        compare_contacts(cg, co)
        sg, so = g.stats(), o.stats()
        assert sg["ncol"] == so["ncol"], k
        assert sg["stored_pairs"] == so["stored_pairs"] == len(co["keys"]), k
        assert sg["stale_dropped"] == 0, k
        if len(co["keys"]) and so["pair_contacts"] == 0:
            stale_steps += 1
        xg, xo = g.download(), o.download()
        for q in ("x", "v"):
            scale = np.abs(xo[q]).max()
            assert np.abs(xg[q] - xo[q]).max() <= 1e-13 * scale, (k, q)
        fs, _ = force_scale(o)
        ok, r = accel_tolerance_ok(xg["a"], xo["a"], fs + 1e-300,
                                   wl.red.inv_mass[wl.comp - 1], 1e-12)
        assert ok, (k, r)
    assert stale_steps > 20
    assert g.stats()["rebuilds"] > 5
    g.close()
    o.close()


def test_stale_pairs_batched_equal_alone(hip):
    cases = [stale_pair_case(k) for k in ("separate", "return")]
    nsteps = min(c[2] for c in cases)
    alone = []
    for wl, con, _ in cases:
        s = System(hip, wl.red)
        _load(s, wl, con)
        s.step(nsteps)
        alone.append((s.download(), s.download_contacts()))
        s.close()
    systems = []
    b = Batch(hip)
    for wl, con, _ in cases:
        s = System(hip, wl.red)
        _load(s, wl, con)
        b.add(s)
        systems.append(s)
    done, _ = b.step(nsteps)
    assert np.all(done == nsteps)
    b.close()
    for s, (st, c) in zip(systems, alone):
        sb = s.download()
        for q in ("x", "v", "w"):
            assert np.array_equal(sb[q], st[q]), q
        compare_contacts(s.download_contacts(), c)
        s.close()
