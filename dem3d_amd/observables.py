"""Bulk observables of a granular state that the reference does not compute
itself, defined here (host side, numpy) for the accuracy statements of the
mixed-precision mode.

* Coordination number Z = 2 P / N, with P the overlapping pairs
  (dem_stats().pair_contacts); wall contacts are reported separately.
* Packing fraction phi = sum of sphere volumes / region volume, counting a
  sphere when its centre lies in the region.  The region is the whole box
  when z is periodic; otherwise the slab that starts one largest diameter
  above the floor (z = -Lz/2) and ends one largest diameter below the top of
  the packing (max z_i + r_i), so the layers ordered by a flat wall or
  thinned at the free surface are left out.
"""
from __future__ import annotations

import numpy as np

from .spec import Reduced


def coordination_number(pair_contacts: int, np_: int) -> float:
    return 2.0 * pair_contacts / max(np_, 1)


def packing_fraction(red: Reduced, x: np.ndarray, comp: np.ndarray) -> float:
    """x: positions [3, np] in reduced units; comp: 1-based components."""
    r = np.asarray(red.radius, np.float64)[np.asarray(comp) - 1]
    vol = 4.0 / 3.0 * np.pi * r ** 3
    lx, ly, lz = (float(v) for v in red.lbox)
    if red.pbc[2]:
        return float(vol.sum() / (lx * ly * lz))
    dmax = 2.0 * float(np.max(red.radius))
    z = x[2]
    zlo = -0.5 * lz + dmax
    zhi = float(np.max(z + r)) - dmax
    if zhi <= zlo:
        zlo, zhi = -0.5 * lz, 0.5 * lz
    sel = (z >= zlo) & (z < zhi)
    return float(vol[sel].sum() / (lx * ly * (zhi - zlo)))
