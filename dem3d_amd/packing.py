"""Seeded disordered packings, grown on the GPU with the library's own path.

A dense random bed cannot be placed directly: random sequential addition
jams near volume fraction 0.38, and a lattice is ordered.  The protocol here
is the one molecular-dynamics codes use for hard- and soft-sphere glasses
(radius growth, after Lubachevsky and Stillinger): start from a sparse,
randomly occupied lattice with shrunken radii, let the spheres move as a
granular gas to erase the lattice, then grow every radius in small stages
and let the contact forces push the overlaps apart after each stage, until
the radii reach their true values and the box holds the target volume
fraction.  Every stage is an ordinary system of this library (dem_create,
dem_upload_state, the Euler start-up, dem_step), so the packing comes from
the same contact law as the runs that use it: linear spring-dashpot with
Coulomb friction (src/MForces.f90:339-387).

The box is periodic in x and y and closed by a floor and a lid (Rectangle
walls), and gravity is off while the bed grows.  The stages are:

1. lattice: `sites_per_sphere` times as many simple-cubic sites as spheres
   fill the box; a seeded choice of sites holds the spheres (species shuffled
   with the same seed) with radii scaled by `s0`, jittered inside the gaps,
   with random velocities of `melt_speed` (in units of the smallest grown
   diameter per collision time);
2. melt: `melt_steps` steps nearly elastic (e = 0.95) and frictionless, so the
   gas loses the lattice;
3. growth: radii scaled by s = s0 ... 1 in steps of `ds_loose` below volume
   fraction `phi_switch` and of `ds_dense` above it, `stage_steps` steps after
   each, with the target contact law (its restitution and friction);
4. the result: positions, velocities and spins at the true radii.

Deterministic: the GPU path is bitwise reproducible run to run, so one seed
gives one bed on one machine.
"""
from __future__ import annotations

import copy
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from .spec import SystemSpec, reduce_spec


@dataclass
class Growth:
    """Parameters of the growth protocol (see the module docstring)."""

    phi: float                       # target volume fraction of the box between floor and lid
    s0: float = 0.70                 # initial radius scale
    sites_per_sphere: float = 1.5
    melt_steps: int = 1500
    melt_speed: float = 0.1
    ds_loose: float = 0.02
    ds_dense: float = 0.01
    phi_switch: float = 0.45
    stage_steps: int = 150
    dense_steps: int = 250
    seed: int = 0
    log: Optional[Callable[[str], None]] = field(default=None, repr=False)

    def scales(self):
        """The radius scales of the growth stages, ending at 1."""
        out = []
        s = self.s0
        while s < 1.0 - 1e-12:
            ds = self.ds_loose if self.phi * s ** 3 < self.phi_switch else self.ds_dense
            s = min(1.0, s + ds)
            out.append(s)
        return out


def scaled_spec(spec: SystemSpec, s: float, e: Optional[float] = None,
                mu: Optional[float] = None) -> SystemSpec:
    """spec with every diameter times s, no gravity, walls at rest, and (when
    given) restitution e and friction mu for every pair and wall."""
    out = copy.deepcopy(spec)
    for c in out.components:
        c.diameter = c.diameter * s
    out.gravity = [0.0, 0.0, 0.0]
    for w in out.walls:
        w.velocity = (0.0, 0.0, 0.0)
    if e is not None:
        for k in ("en", "et", "enw", "etw"):
            setattr(out, k, np.full_like(np.asarray(getattr(out, k), float), e))
    if mu is not None:
        out.mu = np.full_like(np.asarray(out.mu, float), mu)
        out.muw = np.full_like(np.asarray(out.muw, float), mu)
    out.walls_in_movement = False
    return out


def lattice_start(spec: SystemSpec, n: int, floor_z: float, height: float, g: Growth,
                  comp: np.ndarray):
    """Step 1: sparse simple-cubic lattice in the bed region [floor_z,
    floor_z + height] (SI), spheres on a seeded choice of its sites with radii
    scaled by g.s0.  Returns positions [3, n] in metres."""
    rng = np.random.default_rng(g.seed)
    lx, ly = float(spec.lbox[0]), float(spec.lbox[1])
    a = (lx * ly * height / (g.sites_per_sphere * n)) ** (1.0 / 3.0)
    nx, ny = max(1, int(round(lx / a))), max(1, int(round(ly / a)))
    ax, ay = lx / nx, ly / ny
    nz = max(1, int(math.ceil(g.sites_per_sphere * n / (nx * ny))))
    az = height / nz
    # the rounding of the site counts can leave a spacing below the scaled
    # diameter (small boxes): start smaller then
    dtrue = max(c.diameter for c in spec.components)
    g.s0 = min(g.s0, 0.92 * min(ax, ay, az) / dtrue)
    dmax = dtrue * g.s0
    gap = min(ax, ay, az) - dmax
    nsite = nx * ny * nz
    if nsite < n:
        raise ValueError("lattice has fewer sites than spheres")
    site = np.sort(rng.choice(nsite, size=n, replace=False))
    ix, rem = site % nx, site // nx
    iy, iz = rem % ny, rem // ny
    x = np.stack([-0.5 * lx + (ix + 0.5) * ax, -0.5 * ly + (iy + 0.5) * ay,
                  floor_z + (iz + 0.5) * az]).astype(np.float64)
    x += rng.uniform(-0.45 * gap, 0.45 * gap, x.shape) * 0.5
    return x


def grow(lib, spec: SystemSpec, comp: np.ndarray, x_si: np.ndarray, g: Growth, e: float,
         mu: float):
    """Steps 2-4 on the GPU.  Returns (x, v, w) in SI units at the true radii."""
    from .native import System

    rng = np.random.default_rng(g.seed + 1)
    n = x_si.shape[1]
    log = g.log or (lambda m: None)

    def run(s, e_, mu_, x, v, w, steps, tag):
        red = reduce_spec(scaled_spec(spec, s, e_, mu_))
        L, T = red.length_scale, red.time_scale
        t0 = time.perf_counter()
        sysm = System(lib, red)
        sysm.upload(comp, x / L, v * (T / L), w * T)
        sysm.startup()
        sysm.step(steps)
        st = sysm.download()
        stats = sysm.stats()
        sysm.close()
        log(f"packing {tag}: s={s:.3f} phi={g.phi * s ** 3:.3f} steps={steps} "
            f"pairs={stats['pair_contacts']} walls={stats['wall_contacts']} "
            f"{time.perf_counter() - t0:.2f} s")
        return st["x"] * L, st["v"] * (L / T), st["w"] / T, red

    # melt: random directions at melt_speed smallest-diameters per collision time
    red0 = reduce_spec(scaled_spec(spec, 1.0))
    vdir = rng.normal(size=(3, n))
    vdir /= np.linalg.norm(vdir, axis=0, keepdims=True)
    v = vdir * (g.melt_speed * red0.length_scale / red0.time_scale)
    w = np.zeros_like(v)
    x, v, w, _ = run(g.s0, 0.95, 0.0, x_si, v, w, g.melt_steps, "melt")
    for s in g.scales():
        steps = g.stage_steps if g.phi * s ** 3 < g.phi_switch else g.dense_steps
        x, v, w, _ = run(s, e, mu, x, v, w, steps, "grow")
    return x, v, w
