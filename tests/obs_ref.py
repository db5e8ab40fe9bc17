"""Test-side restatement of the observables that the reference does not
define (dem_observables, dem3d_amd/csrc/observe.hpp): the overlapping pairs
and the coordination number, the bulk packing fraction and the vertical
profile, in numpy.  Used only as the checker of the GPU results.

Definitions:
* a pair overlaps when MinDijSq - DijSq > 0 with the minimum image of the
  periodic axes (src/MForces.f90:209-224); the count is over pairs, each once;
* the bulk region is the whole box when z is periodic, else from one largest
  diameter above -Lz/2 to one largest diameter below max(z + r); the packing
  fraction is the volume of the spheres whose centre lies in [zlo, zhi) over
  Lx Ly (zhi - zlo);
* the profile divides the box height into nl layers from -Lz/2; for each it
  gives the sphere volume inside the layer (exact sphere-slab intersection)
  over the layer volume and the volume-weighted mean velocity.
"""
from __future__ import annotations

import math

import numpy as np


def _nint(x):
    """Fortran nint (half away from zero)."""
    return np.where(x >= 0.0, np.floor(x + 0.5), -np.floor(-x + 0.5))


def overlapping_pairs(red, x, comp) -> int:
    """Pairs with MinDijSq - DijSq > 0, by a k-d tree search for candidates
    and the reference's arithmetic for the test (dot product from zero in
    index order, minimum image by nint)."""
    from scipy.spatial import cKDTree
    x = np.asarray(x, np.float64)
    n = x.shape[1]
    c = np.asarray(comp) - 1
    L = np.asarray(red.lbox, np.float64)
    pts = x.T.copy()
    box = []
    for d in range(3):
        if red.pbc[d]:
            pts[:, d] = np.mod(pts[:, d] + 0.5 * L[d], L[d])
            box.append(L[d])
        else:
            pts[:, d] = pts[:, d] - pts[:, d].min()
            box.append(max(1e9, 4.0 * pts[:, d].max() + 1.0))
    rmax = float(np.max(red.radius))
    pairs = cKDTree(pts, boxsize=box).query_pairs(2.0 * rmax * 1.0000001, output_type="ndarray")
    if len(pairs) == 0:
        return 0
    i, j = pairs[:, 0], pairs[:, 1]
    dsq = np.zeros(len(i))
    for d in range(3):
        r = x[d, j] - x[d, i]
        if red.pbc[d]:
            r = r - L[d] * _nint(r * (1.0 / L[d]))
        dsq = dsq + r * r
    mdsq = np.asarray(red.min_dij_sq, np.float64).reshape(red.nc, red.nc)[c[i], c[j]]
    return int(np.count_nonzero(mdsq - dsq > 0.0))


def bulk_region(red, x, comp):
    r = np.asarray(red.radius, np.float64)[np.asarray(comp) - 1]
    lz = float(red.lbox[2])
    zlo, zhi = -0.5 * lz, 0.5 * lz
    if not red.pbc[2] and x.shape[1] > 0:
        dmax = 2.0 * float(np.max(red.radius))
        a, b = zlo + dmax, float(np.max(x[2] + r)) - dmax
        if b > a:
            zlo, zhi = a, b
    return zlo, zhi


def packing_fraction(red, x, comp):
    """(phi, zlo, zhi, solid volume) of the bulk region."""
    r = np.asarray(red.radius, np.float64)[np.asarray(comp) - 1]
    zlo, zhi = bulk_region(red, x, comp)
    sel = (x[2] >= zlo) & (x[2] < zhi)
    vol = float(np.sum(4.0 / 3.0 * math.pi * r[sel] ** 3))
    lx, ly = float(red.lbox[0]), float(red.lbox[1])
    return vol / (lx * ly * (zhi - zlo)), zlo, zhi, vol


def slab_volume(z, r, a, b):
    """Volume of the sphere (centre height z, radius r) between heights a < b."""
    lo = np.maximum(a, z - r)
    hi = np.minimum(b, z + r)
    u, v = lo - z, hi - z
    vol = math.pi * (r * r * (v - u) - (v ** 3 - u ** 3) / 3.0)
    return np.where(hi > lo, vol, 0.0)


def profile(red, x, v, comp, nl):
    """[nl, 4]: solid fraction and mean velocity of the material in each layer."""
    r = np.asarray(red.radius, np.float64)[np.asarray(comp) - 1]
    lx, ly, lz = (float(q) for q in red.lbox)
    dz = lz / nl
    out = np.zeros((nl, 4))
    for k in range(nl):
        a = -0.5 * lz + k * dz
        vol = slab_volume(x[2], r, a, a + dz)
        s = vol.sum()
        out[k, 0] = s / (lx * ly * dz)
        if s > 0.0:
            out[k, 1:] = (vol[None, :] * v).sum(axis=1) / s
    return out
