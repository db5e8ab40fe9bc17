"""Bulk observables of a resident system, computed on the GPU
(dem_observables, dem3d_amd/csrc/observe.hpp):

* KineticEnergy and PotentialEnergy (src/MDEM.f90:183-213);
* coordination number Z = 2 P / N, with P the overlapping pairs at the
  current positions; wall contacts are reported separately;
* packing fraction phi = volume of the spheres whose centre lies in a bulk
  region / region volume.  The region is the whole box when z is periodic;
  otherwise the slab that starts one largest diameter above the bottom of the
  box (z = -Lz/2) and ends one largest diameter below the top of the packing
  (max z_i + r_i), so the layers ordered by a flat wall or thinned at the
  free surface are left out;
* optionally a vertical profile of solid fraction and mean velocity.
"""
from __future__ import annotations

from .native import System


def coordination_number(pair_contacts: int, np_: int) -> float:
    return 2.0 * pair_contacts / max(np_, 1)


def observables(system: System, nlayers: int = 0) -> dict:
    """dem_observables of the system's current state."""
    return system.observables(nlayers)


def packing_fraction(system: System) -> float:
    return system.observables()["packing_fraction"]
