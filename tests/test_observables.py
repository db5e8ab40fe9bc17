"""Known answers of the observables' definitions (tests/obs_ref.py), the
checker of dem_observables: what the mathematics fixes, so that a wrong
formula in the checker cannot pass the GPU parity tests with it.

* the sphere-slab volume integrates the cross-sections pi (r^2 - s^2):
  a slab holding the whole sphere gives 4/3 pi r^3, half-spaces give the
  cap formula pi h^2 (3 r - h) / 3, and the layers of a profile add up to the
  spheres' volume;
* a simple cubic lattice of spheres at spacing a has packing fraction
  (4/3 pi r^3) / a^3 in a region that holds whole lattice layers;
* the pair count equals the brute-force count over all pairs.
"""
from __future__ import annotations

import math

import numpy as np

from dem3d_amd import workloads
from obs_ref import overlapping_pairs, packing_fraction, profile, slab_volume


def test_slab_volume_whole_sphere_and_caps():
    r = 0.7
    assert abs(float(slab_volume(0.0, r, -1.0, 1.0)) - 4.0 / 3.0 * math.pi * r ** 3) < 1e-14
    for h in (0.1, 0.35, 0.7, 1.2):
        cap = math.pi * h * h * (3.0 * r - h) / 3.0
        assert abs(float(slab_volume(0.0, r, r - h, 5.0)) - cap) < 1e-14
    # numerical integration of the cross-sections over an interior slab
    a, b = -0.2, 0.45
    s = np.linspace(a, b, 200001)
    num = np.trapezoid(math.pi * (r * r - s * s), s)
    assert abs(float(slab_volume(0.0, r, a, b)) - num) < 1e-9


def test_profile_layers_add_up_to_the_spheres():
    wl = workloads.small_box(n_side=(4, 4, 5), nc=2, seed=5, vel=0.3)
    red = wl.red
    # lifted so that every sphere lies inside the box height
    wl.x[2] += 0.2
    r = np.asarray(red.radius)[wl.comp - 1]
    assert np.all(wl.x[2] - r > -0.5 * red.lbox[2]) and np.all(wl.x[2] + r < 0.5 * red.lbox[2])
    nl = 37
    p = profile(red, wl.x, wl.v, wl.comp, nl)
    lx, ly, lz = (float(q) for q in red.lbox)
    assert abs(p[:, 0].sum() * lx * ly * lz / nl - float(np.sum(4.0 / 3.0 * math.pi * r ** 3))) \
        < 1e-12 * float(np.sum(r ** 3))
    # a uniform velocity field is its own layer mean
    v = np.tile(np.array([[0.3], [-0.2], [0.05]]), (1, wl.np))
    q = profile(red, wl.x, v, wl.comp, nl)
    full = q[:, 0] > 0
    assert np.allclose(q[full, 1:], [0.3, -0.2, 0.05], rtol=0, atol=1e-15)


def test_packing_fraction_of_a_cubic_lattice():
    wl = workloads.small_box(n_side=(5, 5, 8), nc=1, seed=0, jitter=0.0, spacing=1.05,
                             walls=False, pbc=(True, True, True))
    red = wl.red
    a = float(red.lbox[0]) / 5.0
    r = float(red.radius[0])
    phi, zlo, zhi, _ = packing_fraction(red, wl.x, wl.comp)
    assert (zlo, zhi) == (-0.5 * float(red.lbox[2]), 0.5 * float(red.lbox[2]))
    assert abs(phi - 4.0 / 3.0 * math.pi * r ** 3 / a ** 3) < 1e-12


def test_pair_count_equals_brute_force():
    for seed in (3, 4):
        wl = workloads.small_box(n_side=(5, 4, 5), nc=3, seed=seed, spacing=0.98, jitter=0.08)
        red = wl.red
        x, c = wl.x, wl.comp - 1
        L = np.asarray(red.lbox)
        mdsq = np.asarray(red.min_dij_sq).reshape(red.nc, red.nc)
        n = wl.np
        count = 0
        for i in range(n):
            d = x[:, i + 1:] - x[:, [i]]
            for k in range(3):
                if red.pbc[k]:
                    d[k] -= L[k] * np.where(d[k] / L[k] >= 0, np.floor(d[k] / L[k] + 0.5),
                                            -np.floor(-d[k] / L[k] + 0.5))
            dsq = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
            count += int(np.count_nonzero(mdsq[c[i], c[i + 1:]] - dsq > 0.0))
        assert count > 0
        assert overlapping_pairs(red, x, wl.comp) == count


def test_global_q6_of_lattices_and_a_gas():
    """tests/bond_order.py global_q6 against its published values: FCC
    0.57452, simple cubic 1/(2 sqrt 2), an ideal gas nearly 0."""
    from bond_order import global_q6
    n = 5
    g = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)], float)
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    fcc = (g[:, None, :] + base[None, :, :]).reshape(-1, 3).T
    assert abs(global_q6(fcc, 0.75, box=[n] * 3) - 0.57452) < 1e-4
    assert abs(global_q6(g.T, 1.1, box=[n] * 3) - 1.0 / (2.0 * np.sqrt(2.0))) < 1e-12
    rng = np.random.default_rng(1)
    assert global_q6(rng.uniform(0, 10, (3, 4000)), 1.0, box=[10] * 3) < 0.02
