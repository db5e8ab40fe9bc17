"""Stale pairs (reference semantics): a stored pair stays stored until an
updating ComputeAccelerations visits it -- reference cells the same or
neighbours -- and finds no overlap (RemovePairContact, src/MForces.f90:262-270,
src/MContacts.f90:427-437).

The two scenarios of helpers.stale_pair_case, on the CPU:

* the oracle against the reference modules compiled from their own sources
  (oracle/_ref), loaded with the same planted pair: positions, velocities,
  accelerations, the stored pairs with their (FT, VT) and Ncol bit for bit
  after every step;
* what the scenarios exercise, on the oracle: a pair kept while neither
  overlapping nor adjacent and then forgotten when its cells become adjacent,
  and a contact that starts from a stale history without a new collision.
The GPU side is tests/test_gpu_stale.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from checkers import contacts_to_keys, oracle_system, ref_available, run_reference
from helpers import stale_pair_case

live = pytest.mark.skipif(not ref_available(), reason="reference build not present")


def _oracle(kind):
    wl, con, steps = stale_pair_case(kind)
    o = oracle_system(wl.red)
    o.upload(wl.comp, wl.x, wl.v, wl.w)
    o.upload_contacts(con["keys"], con["ft"], con["vt"], con["wkeys"], con["wft"], con["wvt"])
    return wl, con, steps, o


def _separation(wl, st):
    d = 2.0 * float(wl.red.radius[0])
    Lx = wl.red.lbox[0]
    dx = st["x"][0, 1] - st["x"][0, 0]
    return abs(dx - Lx * np.round(dx / Lx)) / d


def _adjacent_x(cells, nx):
    """Reference cells (1-based linear, x fastest) adjacent along x (the
    scenarios keep both spheres in one cell across)."""
    cx = (np.asarray(cells) - 1) % nx
    off = abs(int(cx[0]) - int(cx[1]))
    return min(off, nx - off) <= 1


@live
@pytest.mark.parametrize("kind", ["separate", "return"])
def test_oracle_stale_pairs_match_live_reference(kind):
    wl, con, steps, o = _oracle(kind)
    ncell = int(np.prod(o.stats()["ref_ncell"]))
    cells = [[] for _ in range(ncell)]
    for i, c in enumerate(o.download_cells()):
        cells[c - 1].append(i + 1)
    pairs = [(1, 2, con["ft"][0], con["vt"][0])]
    state = dict(cells=cells, ipc=[1, 2, 3], ring=np.zeros((3, 9, 2)), pairs=pairs, walls=[])
    actions = []
    for k in range(steps):
        actions += [("step", k + 1, 1), ("snapshot", str(k))]
    res = run_reference(dict(red=wl.red, comp=wl.comp, x=wl.x, v=wl.v, w=wl.w, state=state,
                             actions=actions))
    for k, snap in enumerate(res["snapshots"]):
        o.step(1, k + 1)
        st = o.download()
        for q in ("x", "v", "w", "a", "alpha"):
            assert np.array_equal(st[q], snap[q]), (k, q)
        keys, ft, vt = contacts_to_keys(snap["state"]["pairs"])
        c = o.download_contacts()
        assert np.array_equal(c["keys"], keys), k
        assert np.array_equal(c["ft"].reshape(ft.shape), ft), k
        assert np.array_equal(c["vt"].reshape(vt.shape), vt), k
        assert o.stats()["ncol"] == snap["ncol"], k
        assert np.array_equal(o.download_cells(), snap["cell_index"]), k


def test_separating_pair_is_kept_stale_then_forgotten():
    wl, con, steps, o = _oracle("separate")
    nx = o.stats()["ref_ncell"][0]
    kept_far, forgotten, recontact = 0, None, None
    adj_before = True   # cells of the positions the last updating pass saw
    for k in range(steps):
        o.step(1, k + 1)
        s, c = o.stats(), o.download_contacts()
        sep = _separation(wl, o.download())
        if len(c["keys"]) and sep > 1.2:
            assert not adj_before, k
            kept_far += 1       # beyond any neighbour list of skin 0.1 d
        if forgotten is None and kept_far and not len(c["keys"]):
            forgotten = k
            assert adj_before and sep > 1.0, k
        if forgotten is not None and s["ncol"] > 0 and recontact is None:
            recontact = k
        adj_before = _adjacent_x(o.download_cells(), nx)
    assert kept_far > 50 and forgotten is not None and recontact is not None, (
        kept_far, forgotten, recontact)


def test_returning_pair_resumes_its_stale_history():
    wl, con, steps, o = _oracle("return")
    nx = o.stats()["ref_ncell"][0]
    first = None
    for k in range(steps):
        o.step(1, k + 1)
        s, c = o.stats(), o.download_contacts()
        if first is None and s["pair_contacts"] > 0:
            first = k
            # the step before: apart and in cells two apart, history stored
        assert len(c["keys"]) == 1 and s["ncol"] == 0, k
    assert first is not None
    # the contact changed the planted history: it was read, not restarted
    assert not np.array_equal(o.download_contacts()["ft"], con["ft"])
    o2 = _oracle("return")[3]
    o2.step(first, 1)
    assert not _adjacent_x(o2.download_cells(), nx)
    assert _separation(wl, o2.download()) > 1.0
