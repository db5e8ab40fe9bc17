"""Seeded synthetic granular systems for parity tests and benchmarks.

All systems use the reference's input mapping (src/DEM.f90:360-739): the pair
spring kn enters as component Stiffness = 2 kn (series springs,
src/DEM.f90:273,278-280), ShearStiffness = 2/7 Stiffness, walls have
Stiffness 2 kn (so Knw = kn), tangential restitution equals normal
restitution, MinStepsPerCollision = 50, MethodOrder = 4,
NumberOfEulerIntervals = 10, no fluid.  The box is centred on the origin
(src/MConfig.f90:3); the floor is a Rectangle wall at z = -Lz/2 whose
half-lengths exceed the periodic box so every particle sees its face.

Positions are returned in reduced units (divided by the length scale) as an
array x[3, np]; components are 1-based.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .spec import CONE, RECTANGLE, Component, Reduced, SystemSpec, Wall, reduce_spec


@dataclass
class Workload:
    name: str
    spec: SystemSpec
    red: Reduced
    comp: np.ndarray  # int32 [np], 1-based
    x: np.ndarray     # [3, np] reduced
    v: np.ndarray     # [3, np] reduced
    w: np.ndarray     # [3, np] reduced
    steps: int
    note: str = ""

    @property
    def np(self) -> int:
        return int(self.comp.shape[0])


def _spec(lbox, diameters, density, kn, e, mu, pbc=(True, True, False), walls=None,
          total_steps=1000, gravity=(0.0, 0.0, -9.81)):
    nc = len(diameters)
    comps = [Component(d, density, 2.0 * kn, 2.0 / 7.0 * 2.0 * kn) for d in diameters]
    walls = walls or []
    nw = len(walls)
    spec = SystemSpec(
        lbox=list(lbox), components=comps, walls=walls,
        en=np.full((nc, nc), e), et=np.full((nc, nc), e), mu=np.full((nc, nc), mu),
        enw=np.full((nc, nw), e), etw=np.full((nc, nw), e), muw=np.full((nc, nw), mu),
        pbc=list(pbc), gravity=list(gravity), total_time=1.0,
    )
    # choose TotalTime so that NumberOfSteps = total_steps at dt = Tc/50
    probe = reduce_spec(spec)
    tc = probe.time_scale
    spec.total_time = total_steps * tc / spec.min_steps_per_collision
    return spec


def _floor(lx, ly, lz, kn, top: Optional[float] = None, top_velocity=(0.0, 0.0, 0.0)):
    walls = [Wall(RECTANGLE, 4.0 * lx, 4.0 * ly, 0.0, 2.0 * kn, 2.0 / 7.0 * 2.0 * kn,
                  (0.0, 0.0, -0.5 * lz), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))]
    if top is not None:
        walls.append(Wall(RECTANGLE, 4.0 * lx, 4.0 * ly, 0.0, 2.0 * kn, 2.0 / 7.0 * 2.0 * kn,
                          (0.0, 0.0, top), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                          velocity=top_velocity))
    return walls


def mono_settling_2k(seed: int = 20240001, steps: int = 20000) -> Workload:
    """2 000 monodisperse spheres (d = 1 mm, 2500 kg/m^3) on a 10x10x20 simple
    cubic lattice at spacing 1.2 d with +-0.05 d jitter, box 12 x 12 x 48 mm,
    periodic x,y, floor at z = -Lz/2, velocities U(+-0.05 m/s), kn = 1e3 N/m,
    e = 0.9, mu = 0.5."""
    d, kn = 1e-3, 1e3
    lx, ly, lz = 12e-3, 12e-3, 48e-3
    spec = _spec((lx, ly, lz), [d], 2500.0, kn, 0.9, 0.5, walls=_floor(lx, ly, lz, kn),
                 total_steps=steps)
    red = reduce_spec(spec)
    rng = np.random.default_rng(seed)
    a = 1.2 * d
    ix, iy, iz = np.meshgrid(np.arange(10), np.arange(10), np.arange(20), indexing="ij")
    pos = np.stack([-0.5 * lx + 0.5 * a + a * ix.ravel(),
                    -0.5 * ly + 0.5 * a + a * iy.ravel(),
                    -0.5 * lz + 0.5 * a + a * iz.ravel()])
    pos = pos + rng.uniform(-0.05 * d, 0.05 * d, pos.shape)
    vel = rng.uniform(-0.05, 0.05, pos.shape)
    L, T = red.length_scale, red.time_scale
    n = pos.shape[1]
    return Workload("mono-settling-2k", spec, red, np.ones(n, np.int32), pos / L, vel * T / L,
                    np.zeros_like(pos), steps)


def poly_settling_200k(seed: int = 20240002, steps: int = 5000, n: int = 200_000, nbins: int = 16) -> Workload:
    """200 000 polydisperse spheres, radii U[0.4, 0.6] mm discretised into 16
    equal-width components, on a jittered FCC lattice, base 40 x 40 mm,
    periodic x,y, floor, settling under gravity.  The lattice aims at volume
    fraction 0.45, but its nearest-neighbour distance a/sqrt(2) must exceed
    the largest diameter (1.2 mm) so that no two spheres start in contact;
    that bound sets a = 1.74 mm and phi ~ 0.41."""
    kn = 1e3
    rng = np.random.default_rng(seed)
    edges = np.linspace(0.4e-3, 0.6e-3, nbins + 1)
    centres = 0.5 * (edges[1:] + edges[:-1])
    r = rng.uniform(0.4e-3, 0.6e-3, n)
    comp = np.clip(np.searchsorted(edges, r) - 1, 0, nbins - 1)
    rr = centres[comp]
    mean_vol = float(np.mean(4.0 / 3.0 * math.pi * rr ** 3))
    lx = ly = 40e-3
    # FCC lattice constant for phi = 0.45, widened so the largest spheres do
    # not touch (nearest neighbours at a/sqrt(2) > 1.2 mm, 2% margin)
    dens = 0.45 / mean_vol
    a = max((4.0 / dens) ** (1.0 / 3.0), math.sqrt(2.0) * 1.2e-3 * 1.02)
    nxy = int(lx // a)
    a = lx / nxy
    nz = int(math.ceil(n / (4 * nxy * nxy)))
    lz = 2.0 * nz * a + 10e-3
    pos = _fcc(nxy, nxy, nz, a)[:, :n]
    pos[0] -= 0.5 * lx
    pos[1] -= 0.5 * ly
    pos[2] += -0.5 * lz + 0.6e-3
    gap = a / math.sqrt(2.0) - 1.2e-3
    pos = pos + rng.uniform(-0.45 * gap, 0.45 * gap, pos.shape)
    spec = _spec((lx, ly, lz), list(2.0 * centres), 2500.0, kn, 0.9, 0.5,
                 walls=_floor(lx, ly, lz, kn), total_steps=steps)
    red = reduce_spec(spec)
    L, T = red.length_scale, red.time_scale
    vel = rng.uniform(-0.05, 0.05, pos.shape)
    return Workload("poly-settling-200k", spec, red, (comp + 1).astype(np.int32), pos / L,
                    vel * T / L, np.zeros_like(pos), steps,
                    note=f"FCC a={a*1e3:.4f} mm, phi~{4.0 * mean_vol / a ** 3:.3f}, "
                         f"Lz={lz*1e3:.2f} mm")


def poly_loose_packed(lib, seed: int = 20240002, steps: int = 5000, n: int = 200_000,
                      nbins: int = 16, phi: float = 0.45, lxy: float = 40e-3,
                      log=None, state=None) -> Workload:
    """The settling bed: n polydisperse spheres, radii U[0.4, 0.6] mm (seeded)
    discretised into nbins equal-width components, 2500 kg/m^3, kn = 1e3 N/m,
    e = 0.9, mu = 0.5, in a box periodic in x and y (side lxy) with a floor.
    The spheres are first grown on the GPU (dem3d_amd/packing.py) into a
    random loose packing at volume fraction phi, held between the floor and
    a lid at the height that gives phi and without gravity; the returned
    workload has no lid and gravity on, so the packing settles.  state =
    (x, v, w) in reduced units, from an earlier call with the same arguments,
    skips the growth (lib may then be None)."""
    from .packing import Growth, grow, lattice_start

    kn = 1e3
    rng = np.random.default_rng(seed)
    edges = np.linspace(0.4e-3, 0.6e-3, nbins + 1)
    centres = 0.5 * (edges[1:] + edges[:-1])
    comp = np.clip(np.searchsorted(edges, rng.uniform(0.4e-3, 0.6e-3, n)) - 1, 0, nbins - 1)
    vol = float(np.sum(4.0 / 3.0 * math.pi * centres[comp] ** 3))
    h = vol / (phi * lxy * lxy)
    lz = 2.0 * h + 4.0e-3
    floor_z = -0.5 * lz
    # growth: the bed region between the floor and a lid
    walls = _floor(lxy, lxy, lz, kn, top=floor_z + h)
    walls[1].u, walls[1].w = (0.0, 1.0, 0.0), (1.0, 0.0, 0.0)
    gspec = _spec((lxy, lxy, lz), list(2.0 * centres), 2500.0, kn, 0.9, 0.5, walls=walls,
                  total_steps=steps)
    g = Growth(phi=phi, seed=seed, log=log)
    c1 = (comp + 1).astype(np.int32)
    spec = _spec((lxy, lxy, lz), list(2.0 * centres), 2500.0, kn, 0.9, 0.5,
                 walls=_floor(lxy, lxy, lz, kn), total_steps=steps)
    red = reduce_spec(spec)
    L, T = red.length_scale, red.time_scale
    if state is None:
        x0 = lattice_start(gspec, n, floor_z, h, g, c1)
        x, v, w = grow(lib, gspec, c1, x0, g, e=0.9, mu=0.5)
    else:
        x, v, w = state[0] * L, state[1] * (L / T), state[2] / T
    return Workload(f"poly-loose-{n}-phi{phi:.2f}", spec, red, c1, x / L, v * (T / L), w * T,
                    steps, note=(f"random loose packing grown to phi={phi} (seed {seed}, "
                                 f"dem3d_amd/packing.py) in a {lxy * 1e3:.0f} mm box, bed "
                                 f"{h * 1e3:.2f} mm, lid removed, settling under gravity"))


def _fcc(nx, ny, nz, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1)
    # order layer by layer in z so that truncation removes the top
    order = np.lexsort((cells[:, 0], cells[:, 1], cells[:, 2]))
    cells = cells[order]
    pts = (cells[:, None, :] + base[None, :, :]).reshape(-1, 3)
    pts = pts[np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))]
    return (pts * a + 0.5 * a * 0.5).T.copy()


def bidisperse_sheared_4m(seed: int = 20240003, steps: int = 2000, cells=(100, 100, 50),
            shear_rate: float = 10.0) -> Workload:
    """4 000 000 bidisperse spheres (d = 1.0 / 1.4 mm, 50/50 by number) in a
    dense bed: rock-salt arrangement (large spheres on an FCC lattice, small
    ones on its octahedral sites) with lattice constant 2(r_s + r_L), so every
    sphere starts with 6 touching neighbours of the other size (Z = 6,
    phi = 0.567).  Periodic x,y; floor at z = -Lz/2; a top Rectangle wall in
    contact with the top layer moves in +x at shear_rate * H.  mu = 0.3,
    e = 0.8, kn = 1e3 N/m.  cells = rock-salt unit cells per direction
    (8 spheres each): the default gives 4 000 000 spheres."""
    ds, dl, kn = 1.0e-3, 1.4e-3, 1e3
    a = (ds + dl) * (1.0 - 1e-5)  # L-S distance a/2 = r_s + r_l, slightly compressed
    nx, ny, nz = cells
    lx, ly = nx * a, ny * a
    h = nz * a
    lz = h + 2.0e-3
    rng = np.random.default_rng(seed)
    ix, iy, iz = np.meshgrid(np.arange(2 * nx), np.arange(2 * ny), np.arange(2 * nz), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    large = ((ix + iy + iz) % 2) == 0
    pos = np.stack([ix * 0.5 * a, iy * 0.5 * a, iz * 0.5 * a]).astype(np.float64)
    pos[0] -= 0.5 * lx
    pos[1] -= 0.5 * ly
    z0 = -0.5 * lz + 0.5 * dl * (1.0 - 1e-5)
    pos[2] += z0
    pos = pos + rng.uniform(-1e-9, 1e-9, pos.shape)
    ztop = float(pos[2].max()) + 0.5 * dl * (1.0 - 1e-5)
    vtop = shear_rate * (ztop - (-0.5 * lz))
    walls = _floor(lx, ly, lz, kn, top=ztop, top_velocity=(vtop, 0.0, 0.0))
    # walls[1] normal must point into the bed: W x U flips N = U x W
    walls[1].u, walls[1].w = (0.0, 1.0, 0.0), (1.0, 0.0, 0.0)
    spec = _spec((lx, ly, lz), [ds, dl], 2500.0, kn, 0.8, 0.3, walls=walls, total_steps=steps)
    red = reduce_spec(spec)
    comp = np.where(large, 2, 1).astype(np.int32)
    L = red.length_scale
    n = comp.shape[0]
    # the 4 M bed under its own name; a cut-down bed is named by its cells
    name = ("bidisperse-sheared-4M" if (nx, ny, nz) == (100, 100, 50)
            else f"bidisperse-sheared-{nx}x{ny}x{nz}")
    return Workload(name, spec, red, comp, pos / L,
                    np.zeros((3, n)), np.zeros((3, n)), steps,
                    note=f"rock-salt a={a*1e3:.5f} mm, {nx}x{ny}x{nz} cells, phi=0.567, "
                         f"top wall v_x={vtop:.4f} m/s")


BED_ASPECT = 0.4726  # bed height / side of the 4 M bed (0.1134 m over 0.24 m)


def bidisperse_sheared_bed(n: int = 4_000_000, seed: int = 20240003, phi: float = 0.60,
                           steps: int = 2000, shear_rate: float = 10.0,
                           lxy: Optional[float] = None):
    """The dense sheared bed of bidisperse spheres: n spheres of d = 1.0 and
    1.4 mm, half of each (species assigned by a seeded shuffle), 2500 kg/m^3,
    kn = 1e3 N/m, e = 0.8, mu = 0.3 (pairs and walls).  The box is periodic in
    x and y (side lxy, 0.24 m by default from 1 M spheres up, else chosen so
    that the bed is 0.47 times as high as wide), with a floor at z = -Lz/2 and
    a rough top wall (mu = 0.3) at the height that holds the spheres at
    volume fraction phi; the top wall moves in +x at shear_rate * H.  Gravity
    -9.81 z.

    Returns (spec, comp, x, growth): the specification, the components, the
    seeded sparse-lattice start (metres, radii to be grown) and the growth
    protocol that packs it (dem3d_amd.packing) -- the packed bed is the
    product of running that protocol on the GPU, see packed_bed()."""
    from .packing import Growth, lattice_start

    ds, dl, kn = 1.0e-3, 1.4e-3, 1e3
    vmean = 0.5 * (math.pi / 6.0) * (ds ** 3 + dl ** 3)
    if lxy is None:
        lxy = 0.24 if n >= 1_000_000 else (n * vmean / (phi * BED_ASPECT)) ** (1.0 / 3.0)
    nl = n // 2
    rng = np.random.default_rng(seed)
    comp = np.where(rng.permutation(n) < nl, 1, 2).astype(np.int32)
    vol = nl * (math.pi / 6.0) * ds ** 3 + (n - nl) * (math.pi / 6.0) * dl ** 3
    h = vol / (phi * lxy * lxy)
    lz = h + 2.0e-3
    floor_z = -0.5 * lz
    ztop = floor_z + h
    vtop = shear_rate * h
    walls = _floor(lxy, lxy, lz, kn, top=ztop, top_velocity=(vtop, 0.0, 0.0))
    # walls[1] normal must point into the bed: W x U flips N = U x W
    walls[1].u, walls[1].w = (0.0, 1.0, 0.0), (1.0, 0.0, 0.0)
    spec = _spec((lxy, lxy, lz), [ds, dl], 2500.0, kn, 0.8, 0.3, walls=walls, total_steps=steps)
    g = Growth(phi=phi, seed=seed)
    x = lattice_start(spec, n, floor_z, h, g, comp)
    return spec, comp, x, g


def packed_bed(lib, n: int = 4_000_000, seed: int = 20240003, phi: float = 0.60,
               steps: int = 2000, shear_rate: float = 10.0, lxy: Optional[float] = None,
               settle_steps: int = 1000, log=None, state=None, **growth) -> Workload:
    """bidisperse_sheared_bed grown on the GPU (library `lib`) into the dense
    disordered bed, then run settle_steps steps with its own contact law,
    gravity and the top wall shearing: the Workload of the sheared bed at
    that state.  state = (x, v, w) in reduced units, from an earlier call
    with the same arguments, skips the growth and the settling."""
    from .native import System
    from .packing import grow

    spec, comp, x0, g = bidisperse_sheared_bed(n, seed, phi, steps, shear_rate, lxy)
    red = reduce_spec(spec)
    L, T = red.length_scale, red.time_scale
    if state is None:
        for k, v in growth.items():
            setattr(g, k, v)
        g.log = log
        x, v, w = grow(lib, spec, comp, x0, g, e=0.8, mu=0.3)
    else:
        x, v, w = state[0] * L, state[1] * (L / T), state[2] / T
    if settle_steps > 0 and state is None:
        sysm = System(lib, red)
        sysm.upload(comp, x / L, v * (T / L), w * T)
        sysm.startup()
        sysm.step(settle_steps)
        st = sysm.download()
        sysm.close()
        x, v, w = st["x"] * L, st["v"] * (L / T), st["w"] / T
        if log:
            log(f"packing settle: {settle_steps} steps with gravity and the shearing top wall")
    name = (f"bidisperse-sheared-{n // 1_000_000}M-phi{phi:.2f}" if n % 1_000_000 == 0
            else f"bidisperse-sheared-{n}-phi{phi:.2f}")
    return Workload(name, spec, red, comp, x / L, v * (T / L), w * T, steps,
                    note=(f"disordered bed grown to phi={phi} from a seeded sparse lattice "
                          f"(seed {seed}, dem3d_amd/packing.py), side {spec.lbox[0] * 1e3:.1f} mm, "
                          f"bed {(spec.walls[1].position[2] - spec.walls[0].position[2]) * 1e3:.2f}"
                          f" mm, top wall v_x={spec.walls[1].velocity[0]:.4f} m/s, then "
                          f"{settle_steps} steps sheared under gravity"))


def small_box(n_side=(6, 6, 6), seed=1, jitter=0.05, spacing=1.02, nc=2, ratio=1.3,
              walls=True, pbc=(True, True, False), vel=0.5, steps=1000, gravity=-9.81,
              mu=0.5, e=0.9) -> Workload:
    """Small test system: jittered cubic lattice of spheres in a periodic box
    (a few hundred particles), dense enough (spacing ~ max diameter) to have
    many contacts and strong shear at the first steps."""
    kn = 1e3
    rng = np.random.default_rng(seed)
    d0 = 1e-3
    diams = [d0 * (1.0 + (ratio - 1.0) * k / max(nc - 1, 1)) for k in range(nc)]
    dmax = max(diams)
    a = spacing * dmax
    nx, ny, nz = n_side
    lx, ly = nx * a, ny * a
    lz = nz * a + (0.0 if pbc[2] else 2.0 * dmax)
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = np.stack([-0.5 * lx + 0.5 * a + a * ix.ravel(),
                    -0.5 * ly + 0.5 * a + a * iy.ravel(),
                    -0.5 * lz + 0.5 * a + a * iz.ravel()]).astype(np.float64)
    pos += rng.uniform(-jitter * d0, jitter * d0, pos.shape)
    n = pos.shape[1]
    comp = rng.integers(1, nc + 1, n).astype(np.int32)
    wl = _floor(lx, ly, lz, kn) if walls else []
    spec = _spec((lx, ly, lz), diams, 2500.0, kn, e, mu, pbc=pbc, walls=wl, total_steps=steps,
                 gravity=(0.0, 0.0, gravity))
    red = reduce_spec(spec)
    L, T = red.length_scale, red.time_scale
    v = rng.uniform(-vel, vel, pos.shape) * T / L
    w = rng.uniform(-vel, vel, pos.shape) * T / L / 0.5
    return Workload(f"small-{n}", spec, red, comp, pos / L, v, w, steps)


def ball_bearings(seed: int = 20240006, steps: int = 2000, n: int = 39,
                  shake: bool = True) -> Workload:
    """The reference's example input (examples/ball-bearings/ball-bearings.inp):
    39 steel balls (10 mm, 7800 kg/m^3, stiffness 1e5 N/m) in a dish made of
    five Cone walls -- an inner cylinder of radius 35 mm and its bottom disc,
    an outer cylinder of radius 36 mm and its disc, and the annulus that joins
    their rims -- shaken on a 10 mm circle at 1.5 Hz (amplitude (10, 10, 0)
    mm, phases (0, -1/4, 0) of a period), e = 0.7, mu = 0.3, 10 steps per
    collision, 100 Euler intervals, closed box 0.11 x 0.11 x 0.05 m.  The
    example pours the balls (ball-bearings.prg); here they start at rest on two
    square layers (10.5 mm pitch, jittered by up to 0.1 mm) inside the inner
    cylinder."""
    rng = np.random.default_rng(seed)
    k = 1.0e5
    up, side = (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)
    amp, freq, phase = (0.01, 0.01, 0.0), (1.5, 1.5, 1.5), (0.0, -0.25, 0.0)
    mv = dict(amplitude=amp, frequency=freq, phase=phase) if shake else {}
    geo = [(0.015, 0.070, 0.070, -0.0165), (0.000, 0.000, 0.070, -0.0240),
           (0.016, 0.072, 0.072, -0.0170), (0.000, 0.000, 0.072, -0.0250),
           (0.000, 0.070, 0.072, -0.0090)]
    walls = [Wall(CONE, l1, l2, l3, k, k, (0.01, 0.0, z), up, side, **mv) for l1, l2, l3, z in geo]
    nw = len(walls)
    spec = SystemSpec(
        lbox=[0.11, 0.11, 0.05], components=[Component(0.01, 7.8e3, k, k)], walls=walls,
        en=np.full((1, 1), 0.7), et=np.full((1, 1), 0.7), mu=np.full((1, 1), 0.3),
        enw=np.full((1, nw), 0.7), etw=np.full((1, nw), 0.7), muw=np.full((1, nw), 0.3),
        pbc=[False, False, False], gravity=[0.0, 0.0, -9.8], total_time=30.0,
        min_steps_per_collision=10.0, method_order=4, n_euler=100, walls_in_movement=shake)
    red = reduce_spec(spec)
    L = red.length_scale
    a, rmax = 0.0105, 0.035 - 0.005 - 0.0005
    cx, cy, z0 = 0.01 + 3.7e-4, 2.1e-4, -0.024 + 0.005 + 2e-4
    pts = []
    for layer in range(3):
        off = 0.5 * a * (layer % 2)
        for i in range(-4, 5):
            for j in range(-4, 5):
                px, py = i * a + off, j * a + off
                if math.hypot(px, py) <= rmax:
                    pts.append((cx + px, cy + py, z0 + layer * a))
    if len(pts) < n:
        raise ValueError(f"at most {len(pts)} balls fit the start layers")
    pos = np.array(pts[:n], dtype=np.float64).reshape(n, 3).T + rng.uniform(-1e-4, 1e-4, (3, n))
    zero = np.zeros((3, n))
    return Workload(f"ball-bearings-{n}", spec, red, np.ones(n, np.int32), pos / L, zero.copy(),
                    zero.copy(), steps, note="example input, balls placed instead of poured")


def discharge(seed: int = 20240007, n_side=(8, 8, 6), gravity: float = -20000.0,
              steps: int = 600) -> Workload:
    """A bed with an open bottom (no floor, periodic in x and y) falling out of
    the box under strong gravity, for the removal of escaped particles
    (UpdateParticleList).  The particles of the two lowest layers, which
    leave first, hold the last indices in random order, so the last particle
    -- the one that takes a removed particle's place -- is usually a
    neighbour of the one leaving, with contacts to rename."""
    wl = small_box(n_side=n_side, nc=2, ratio=1.2, seed=seed, walls=False,
                   pbc=(True, True, False), vel=0.2, spacing=1.01, gravity=gravity, e=0.5,
                   steps=steps)
    rng = np.random.default_rng(seed)
    # the two lowest layers take the last indices, in random order; the rest
    # are shuffled below them
    z = wl.x[2]
    low = np.nonzero(z < np.sort(np.unique(np.round(z, 3)))[0] + 1.5)[0]
    high = np.setdiff1d(np.arange(wl.np), low)
    perm = np.concatenate([rng.permutation(high), rng.permutation(low)])
    return Workload(f"discharge-{wl.np}", wl.spec, wl.red, wl.comp[perm], wl.x[:, perm],
                    wl.v[:, perm], wl.w[:, perm], steps, note="open bottom, shuffled indices")


def fluidized_bed(seed: int = 20240008, n_side=(6, 6, 6), velocity: float = 0.02,
                  layers: int = 40, steps: int = 1000) -> Workload:
    """Glass beads (1.0 / 1.2 mm, 2500 kg/m^3) in water (1000 kg/m^3,
    1e-3 Pa s) flowing upward at `velocity` m/s superficial speed, in a box
    periodic in x and y between a floor and a lid, with NL = `layers` fluid
    layers: mFluid's porosity field and drag at every evaluation."""
    wl = small_box(n_side=n_side, nc=2, ratio=1.2, seed=seed, vel=0.1, spacing=1.15, e=0.8)
    spec = wl.spec
    lx, ly, lz = spec.lbox
    lz = 2.0 * lz
    spec.lbox = [lx, ly, lz]
    spec.walls = _floor(lx, ly, lz, 1e3, top=0.5 * lz)
    nc, nw = len(spec.components), len(spec.walls)
    spec.enw = np.full((nc, nw), 0.8)
    spec.etw = np.full((nc, nw), 0.8)
    spec.muw = np.full((nc, nw), 0.5)
    spec.fluid_density = 1000.0
    spec.fluid_flow = True
    spec.fluid_viscosity = 1.0e-3
    spec.fluid_flow_rate = velocity * lx * ly
    spec.fluid_layers = layers
    red = reduce_spec(spec)
    L = red.length_scale
    x = wl.x * wl.red.length_scale / L
    x[2] = x[2] - 0.5 * (wl.spec.lbox[2] / L) * 0.5   # the bed in the lower half
    return Workload(f"fluidized-{wl.np}", spec, red, wl.comp, x, wl.v, wl.w, steps,
                    note="upward water flow, drag of src/mFluid.f90")


def _member_layout(n, mu, e, seed, nbins, r0, steps, phi):
    """The radii, box and constants of an ensemble member (the first draws of
    its generator, which the member's positions and velocities continue)."""
    kn = 1e3
    rng = np.random.default_rng(seed)
    edges = np.linspace(0.9 * r0, 1.1 * r0, nbins + 1)
    centres = 0.5 * (edges[1:] + edges[:-1])
    r = rng.uniform(0.9 * r0, 1.1 * r0, n)
    comp = np.clip(np.searchsorted(edges, r) - 1, 0, nbins - 1)
    rr = centres[comp]
    mean_vol = float(np.mean(4.0 / 3.0 * math.pi * rr ** 3))
    a = (4.0 * mean_vol / phi) ** (1.0 / 3.0)          # FCC constant for phi
    nxy = max(3, int(round((n / 4.0) ** (1.0 / 3.0))))  # roughly cubic
    nz = int(math.ceil(n / (4 * nxy * nxy)))
    lx = ly = nxy * a
    lz = nz * a + 6.0 * r0 + 0.5 * a
    spec = _spec((lx, ly, lz), list(2.0 * centres), 2500.0, kn, e, mu,
                 walls=_floor(lx, ly, lz, kn), total_steps=steps)
    return rng, comp, a, (nxy, nz), spec, reduce_spec(spec)


def ensemble_member_reduced(n: int, mu: float, e: float, seed: int, nbins: int = 8,
                            r0: float = 0.5e-3, steps: int = 10000, phi: float = 0.5):
    """The reduced constants of ensemble_member(n, mu, e, seed, ...) without
    building its particles (what admission by device footprint needs)."""
    return _member_layout(n, mu, e, seed, nbins, r0, steps, phi)[5]


def ensemble_member(n: int, mu: float, e: float, seed: int, nbins: int = 8, r0: float = 0.5e-3,
                    steps: int = 10000, phi: float = 0.5) -> Workload:
    """One system of an ensemble: n spheres with radii U[0.9, 1.1] r0
    discretised into nbins equal-width components, on a jittered FCC lattice
    at volume fraction ~phi in a box periodic in x,y with a floor, settling
    under gravity; friction mu and restitution e for every pair and the
    floor, kn = 1e3 N/m."""
    rng, comp, a, (nxy, nz), spec, red = _member_layout(n, mu, e, seed, nbins, r0, steps, phi)
    lx, ly, lz = spec.lbox
    pos = _fcc(nxy, nxy, nz, a)[:, :n]
    pos[0] -= 0.5 * lx
    pos[1] -= 0.5 * ly
    pos[2] += -0.5 * lz + 1.1 * r0
    gap = a / math.sqrt(2.0) - 2.2 * r0
    pos = pos + rng.uniform(-0.45 * gap, 0.45 * gap, pos.shape)
    L, T = red.length_scale, red.time_scale
    vel = rng.uniform(-0.05, 0.05, pos.shape)
    return Workload(f"member-{n}", spec, red, (comp + 1).astype(np.int32), pos / L,
                    vel * T / L, np.zeros_like(pos), steps,
                    note=f"n={n} mu={mu:.3f} e={e:.3f} FCC a={a*1e3:.4f} mm phi~{phi}")


def ensemble_256(seed: int = 20240004, n_systems: int = 256, steps: int = 10000):
    """The ensemble of 256 independent systems: sizes log-uniform in
    [500, 20 000] spheres, mu ~ U[0.1, 0.7], e ~ U[0.6, 0.95], radii
    U[0.9, 1.1] r0 in 8 components, each in its own periodic box with a floor
    at phi ~ 0.5.  Returns the list of member workloads."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(500), math.log(20000), n_systems))).astype(int)
    mus = rng.uniform(0.1, 0.7, n_systems)
    es = rng.uniform(0.6, 0.95, n_systems)
    seeds = rng.integers(0, 2 ** 31 - 1, n_systems)
    return [ensemble_member(int(n), float(m), float(e), int(sd), steps=steps)
            for n, m, e, sd in zip(sizes, mus, es, seeds)]


@dataclass
class Request:
    arrival: float      # seconds after the start of the trace
    n: int
    steps: int
    mu: float
    e: float
    seed: int


def serving_trace(seed: int = 20240005, n_requests: int = 200, rate: float = 50.0,
                  n_range=(1000, 500000), step_range=(1000, 20000)):
    """Requests of the serving stream: Poisson arrivals at `rate` per second,
    sizes log-uniform in n_range, step counts uniform in step_range, contact
    parameters as in the ensemble."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    ns = np.round(np.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1]),
                                     n_requests))).astype(int)
    st = rng.integers(step_range[0], step_range[1] + 1, n_requests)
    mus = rng.uniform(0.1, 0.7, n_requests)
    es = rng.uniform(0.6, 0.95, n_requests)
    seeds = rng.integers(0, 2 ** 31 - 1, n_requests)
    return [Request(float(a), int(n), int(s), float(m), float(e), int(sd))
            for a, n, s, m, e, sd in zip(t, ns, st, mus, es, seeds)]
