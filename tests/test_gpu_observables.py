"""dem_observables on the GPU against the oracle and the numpy restatement
(tests/obs_ref.py), on the same states:

* KineticEnergy and PotentialEnergy (src/MDEM.f90:183-213) against the
  oracle's, which the reference pins bit for bit (tests/test_oracle.py):
  within 1e-13 of the sum of the magnitudes of their terms (the GPU adds the
  terms in another order);
* overlapping pairs and particle-wall contacts equal to the oracle's counts
  after an evaluation of the same state, and to a k-d tree search;
* the bulk region bit for bit, its packing fraction and the vertical profile
  within 1e-12 relative;
* the same after 40 GPU steps, the checkers fed the GPU's own state.
"""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import System
from checkers import oracle_system
from obs_ref import overlapping_pairs, packing_fraction, profile

pytestmark = pytest.mark.gpu

CASES = [
    dict(n_side=(6, 6, 6), nc=2, seed=3, vel=0.3),
    dict(n_side=(7, 5, 6), nc=3, seed=8, vel=0.5, pbc=(True, True, True), walls=False),
    dict(n_side=(5, 6, 7), nc=1, seed=9, vel=0.2, pbc=(False, False, False)),
]


def _check(g: System, red, st, nl=24):
    ob = g.observables(nl)
    o = oracle_system(red)
    o.upload(st["comp"], st["x"], st["v"], st["w"])
    o.compute_accelerations(0, update=False)
    os_ = o.stats()
    m = np.asarray(red.mass)[st["comp"] - 1]
    I = np.asarray(red.inertia)[st["comp"] - 1]
    ke_terms = m * (st["v"] ** 2).sum(axis=0) + I * (st["w"] ** 2).sum(axis=0)
    assert abs(ob["kinetic_energy"] - os_["kinetic_energy"]) <= 1e-13 * 0.5 * ke_terms.sum()
    pe_terms = np.abs(m * (st["x"] * np.asarray(red.gravity)[:, None]).sum(axis=0))
    pe = o.lib.potential_energy(o.h)
    assert abs(ob["potential_energy"] - pe) <= 1e-13 * pe_terms.sum() + 1e-300
    assert ob["pair_contacts"] == os_["pair_contacts"]
    assert ob["pair_contacts"] == overlapping_pairs(red, st["x"], st["comp"])
    assert ob["wall_contacts"] == os_["wall_contacts"]
    n = st["x"].shape[1]
    assert ob["coordination"] == 2.0 * ob["pair_contacts"] / n
    phi, zlo, zhi, vol = packing_fraction(red, st["x"], st["comp"])
    assert ob["region"] == [zlo, zhi]
    assert abs(ob["solid_volume"] - vol) <= 1e-12 * vol
    assert abs(ob["packing_fraction"] - phi) <= 1e-12 * phi
    pr = profile(red, st["x"], st["v"], st["comp"], nl)
    for c in range(4):
        scale = np.abs(pr[:, c]).max() + 1e-300
        assert np.abs(ob["profile"][:, c] - pr[:, c]).max() <= 1e-12 * scale, c
    o.close()
    return ob


@pytest.mark.parametrize("kw", CASES)
def test_observables_match_oracle_and_checker(hip, kw):
    wl = workloads.small_box(**kw)
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    st = g.download()
    ob0 = _check(g, wl.red, st)
    assert ob0["pair_contacts"] > 0
    g.startup()
    g.step(40)
    st = g.download()
    _check(g, wl.red, st)
    # bitwise reproducible
    a, b = g.observables(24), g.observables(24)
    assert all(np.array_equal(np.asarray(a[k]), np.asarray(b[k])) for k in a)
    g.close()


def test_observables_of_a_packed_bed(hip):
    """A bed grown to phi 0.60 (the headline protocol at 12 000 spheres):
    contact counts and bulk packing fraction against the checkers, and the
    global bond order Q6 of a random packing."""
    wl = workloads.packed_bed(hip, n=12000, settle_steps=100)
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    st = g.download()
    ob = _check(g, wl.red, st, nl=64)
    assert 0.57 < ob["packing_fraction"] < 0.66
    assert ob["coordination"] > 1.0
    # disordered: the global bond order of a random packing (FCC 0.575, the
    # rock-salt crystal of the second leg 0.354)
    from bond_order import global_q6
    dmean = 2.0 * float(np.mean(np.asarray(wl.red.radius)[wl.comp - 1]))
    q6 = global_q6(st["x"], 1.2 * dmean, box=None)
    print(f"packed bed: phi {ob['packing_fraction']:.4f}, Z {ob['coordination']:.3f}, Q6 {q6:.4f}",
          flush=True)
    assert q6 < 0.1, q6
    g.close()
