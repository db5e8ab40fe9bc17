"""Global bond-orientational order Q6 (Steinhardt, Nelson and Ronchetti,
Phys. Rev. B 28, 784 (1983)): over all neighbour bonds r_ij of a packing,
Q6 = sqrt(4 pi / 13 * sum_m |<Y6m(r_ij)>|^2).  An FCC crystal gives 0.5745,
a simple cubic one 0.3536, a random packing nearly 0.  Test helper (numpy,
scipy); tests/test_observables.py pins it on lattices."""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def _y6(theta, phi):
    from scipy.special import sph_harm_y
    return np.stack([sph_harm_y(6, m, theta, phi) for m in range(-6, 7)])


def global_q6(x, cutoff, box=None):
    """x[3, n]; bonds are the pairs closer than cutoff (minimum image in a
    periodic box of side lengths `box` when given)."""
    pts = np.asarray(x, np.float64).T
    if box is not None:
        box = np.asarray(box, np.float64)
        tree = cKDTree(np.mod(pts + 0.5 * box, box), boxsize=box)
    else:
        tree = cKDTree(pts)
    pairs = tree.query_pairs(cutoff, output_type="ndarray")
    if len(pairs) == 0:
        return 0.0
    d = pts[pairs[:, 1]] - pts[pairs[:, 0]]
    if box is not None:
        d -= box * np.round(d / box)
    r = np.linalg.norm(d, axis=1)
    theta = np.arccos(np.clip(d[:, 2] / r, -1.0, 1.0))
    phi = np.arctan2(d[:, 1], d[:, 0])
    q = _y6(theta, phi).mean(axis=1)
    return float(np.sqrt(4.0 * np.pi / 13.0 * np.sum(np.abs(q) ** 2)))
