"""Physical specification of a granular system and its reduction to the
constants the hot path consumes.

The Fortran driver derives these constants before the time loop; this module
restates that derivation so that a caller can build `dem_params_t` without the
driver:

* ``CalculateProperties``  src/DEM.f90:245-356  (masses, inertia, series
  springs, damping from restitution, collision time, time step)
* ``ScaleVariables``       src/DEM.f90:807-866  (reduced units: length = min
  diameter, mass = min mass, time = min collision time)
* ``InitializeForce``      src/MForces.f90:54-87 (MinDij, MinDijSq, InvMass,
  InvInertia, EffectiveMass, MuSq, MuwSq)
* wall parsing             src/DEM.f90:472-509 (half lengths, Gram-Schmidt,
  N = U x W)

Every expression keeps the reference's association so that IEEE double results
are the same bits (x**2 -> x*x, x**3 -> (x*x)*x, left-to-right products).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

PI = 3.141592653589793238  # src/DEM.f90:249 and src/MUtil.f90
TWO_PI = 2.0 * PI

RECTANGLE = 1
CONE = 2


@dataclass
class Component:
    """One line of the component block of the specification (src/DEM.f90:395-424)."""

    diameter: float
    density: float
    stiffness: float
    shear_stiffness: float


@dataclass
class Wall:
    """One wall line (src/DEM.f90:474-507): full lengths as read, before halving."""

    shape: int
    l1: float
    l2: float
    l3: float
    stiffness: float
    shear_stiffness: float
    position: Sequence[float]
    u: Sequence[float]
    w: Sequence[float]
    velocity: Sequence[float] = (0.0, 0.0, 0.0)
    omega: Sequence[float] = (0.0, 0.0, 0.0)
    # movement line (src/DEM.f90:722): amplitude (m), frequency (Hz) and phase
    # (fraction of a period) per axis; used when the system's walls move
    amplitude: Sequence[float] = (0.0, 0.0, 0.0)
    frequency: Sequence[float] = (0.0, 0.0, 0.0)
    phase: Sequence[float] = (0.0, 0.0, 0.0)


@dataclass
class SystemSpec:
    """Physical (SI) description, the content of a DEM3D `.inp` file."""

    lbox: Sequence[float]
    components: List[Component]
    walls: List[Wall]
    en: np.ndarray  # NC x NC
    et: np.ndarray
    mu: np.ndarray
    enw: np.ndarray  # NC x NW
    etw: np.ndarray
    muw: np.ndarray
    pbc: Sequence[bool]
    gravity: Sequence[float]
    total_time: float
    min_steps_per_collision: float = 50.0
    method_order: int = 4
    n_euler: int = 10
    fluid_density: float = 0.0
    walls_in_movement: bool = False  # WallsInMovement (src/DEM.f90:716)
    # fluid block (src/mFluid.f90:29-57): FluidFlow, viscosity (Pa s), volume
    # flow rate upward (m^3/s), number of layers NL
    fluid_flow: bool = False
    fluid_viscosity: float = 0.0
    fluid_flow_rate: float = 0.0
    fluid_layers: int = 0


@dataclass
class Reduced:
    """Constants in reduced units, as the hot path sees them."""

    nc: int
    nw: int
    pbc: List[int]
    method_order: int
    n_euler: int
    lbox: np.ndarray
    gravity: np.ndarray
    time_step: float
    number_of_steps: int
    radius: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    density: np.ndarray
    inv_mass: np.ndarray
    inv_inertia: np.ndarray
    eff_mass: np.ndarray
    min_dij: np.ndarray
    min_dij_sq: np.ndarray
    kn: np.ndarray
    yn: np.ndarray
    kt: np.ndarray
    yt: np.ndarray
    mu: np.ndarray
    musq: np.ndarray
    knw: np.ndarray
    ynw: np.ndarray
    ktw: np.ndarray
    ytw: np.ndarray
    muw: np.ndarray
    muwsq: np.ndarray
    walls: list  # list of dicts (reduced TWall fields)
    length_scale: float
    mass_scale: float
    time_scale: float
    max_ncell: int = 30000
    ref_grid: Optional[Sequence[int]] = None
    wall_motion: Optional[list] = None  # WM(1:NW), reduced (list of dicts) or None
    fluid: Optional[dict] = None  # FluidDensity, viscosity, flow_rate (reduced), nl


def _gram_schmidt(u, w):
    """GramSchmidt (src/MUtil.f90:163-171)."""
    u = [float(x) for x in u]
    w = [float(x) for x in w]
    s = 0.0
    for x in u:
        s = s + x * x
    n = math.sqrt(s)
    u = [x / n for x in u]
    d = 0.0
    for a, b in zip(u, w):
        d = d + a * b
    w = [b - d * a for a, b in zip(u, w)]
    s = 0.0
    for x in w:
        s = s + x * x
    n = math.sqrt(s)
    w = [x / n for x in w]
    return u, w


def _cross(a, b):
    """CrossProduct (src/MDimension3D.f90:83-95)."""
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def reduce_spec(s: SystemSpec, max_ncell: int = 30000) -> Reduced:
    nc = len(s.components)
    nw = len(s.walls)
    radius = np.array([0.5 * c.diameter for c in s.components])  # src/DEM.f90:402
    density = np.array([c.density for c in s.components])
    stiff = np.array([c.stiffness for c in s.components])
    shear = np.array([c.shear_stiffness for c in s.components])
    # CalculateProperties (src/DEM.f90:259-263)
    mass = np.array([4.0 * PI / 3.0 * ((r * r) * r) * rho for r, rho in zip(radius, density)])
    inertia = np.array([2.0 / 5.0 * m * (r * r) for m, r in zip(mass, radius)])
    meff = np.zeros((nc, nc))
    kn = np.zeros((nc, nc))
    kt = np.zeros((nc, nc))
    for i in range(nc):
        meff[i, i] = 0.5 * mass[i]
        kn[i, i] = 0.5 * stiff[i]
        kt[i, i] = 0.5 * shear[i]
        for j in range(i + 1, nc):
            meff[i, j] = meff[j, i] = mass[i] * mass[j] / (mass[i] + mass[j])
            kn[i, j] = kn[j, i] = stiff[i] * stiff[j] / (stiff[i] + stiff[j])
            kt[i, j] = kt[j, i] = shear[i] * shear[j] / (shear[i] + shear[j])

    def damping(e, m, k):
        out = np.zeros_like(m)
        for idx in np.ndindex(m.shape):
            loge = math.log(float(e[idx]))
            gamma = -2.0 * loge / math.sqrt(loge * loge + PI * PI)
            out[idx] = gamma * math.sqrt(m[idx] * k[idx])
        return out

    en = np.asarray(s.en, float).reshape(nc, nc)
    et = np.asarray(s.et, float).reshape(nc, nc)
    yn = damping(en, meff, kn)
    yt = damping(et, meff, kt)
    meffw = np.zeros((nc, nw))
    knw = np.zeros((nc, nw))
    ktw = np.zeros((nc, nw))
    for i in range(nc):
        for j in range(nw):
            wl = s.walls[j]
            meffw[i, j] = mass[i]
            knw[i, j] = stiff[i] * wl.stiffness / (stiff[i] + wl.stiffness)
            ktw[i, j] = shear[i] * wl.shear_stiffness / (shear[i] + wl.shear_stiffness)
    enw = np.asarray(s.enw, float).reshape(nc, nw)
    etw = np.asarray(s.etw, float).reshape(nc, nw)
    ynw = damping(enw, meffw, knw)
    ytw = damping(etw, meffw, ktw)
    # MinCollisionTime (src/DEM.f90:330-331)
    tmin = math.inf
    for idx in np.ndindex(meff.shape):
        tmin = min(tmin, meff[idx] / math.sqrt(4.0 * meff[idx] * kn[idx] - yn[idx] * yn[idx]))
    tminw = math.inf
    for idx in np.ndindex(meffw.shape):
        tminw = min(tminw, meffw[idx] / math.sqrt(4.0 * meffw[idx] * knw[idx] - ynw[idx] * ynw[idx]))
    tc = TWO_PI * min(tmin, tminw)
    dt = tc / s.min_steps_per_collision
    nsteps = int(round_half_away(s.total_time / dt))
    dt = s.total_time / float(nsteps)
    # ScaleVariables (src/DEM.f90:814-860)
    L = 2.0 * float(radius.min())
    M = float(mass.min())
    T = tc
    t2 = T * T
    lbox = np.array([x / L for x in s.lbox])
    radius_r = radius / L
    mass_r = mass / M
    density_r = np.array([(rho * ((L * L) * L)) / M for rho in density])
    inertia_r = np.array([i / ((L * L) * M) for i in inertia])
    kn_r = kn * t2 / M
    yn_r = yn * T / M
    knw_r = knw * t2 / M
    ynw_r = ynw * T / M
    kt_r = kt * t2 / M
    yt_r = yt * T / M
    ktw_r = ktw * t2 / M
    ytw_r = ytw * T / M
    gravity = np.array([g * t2 / L for g in s.gravity])
    dt_r = dt / T
    wall_motion = None
    if s.walls_in_movement and nw:
        wall_motion = []
        for wl in s.walls:
            # src/DEM.f90:722-732 in SI, then ScaleVariables (src/DEM.f90:841-846)
            for f in wl.frequency:
                if f < 0.0:
                    raise ValueError("wall movement frequencies cannot be negative")
            tpf = [TWO_PI * f for f in wl.frequency]
            phase = [TWO_PI * p for p in wl.phase]
            vamp = [-t * a for t, a in zip(tpf, wl.amplitude)]
            cf = [p - a * math.cos(ph) for p, a, ph in zip(wl.position, wl.amplitude, phase)]
            wall_motion.append(dict(
                constant_factor=np.array([c / L for c in cf]),
                amplitude=np.array([a / L for a in wl.amplitude]),
                velocity_amplitude=np.array([v * T / L for v in vamp]),
                two_pi_frequency=np.array([t * T for t in tpf]),
                phase_angle=np.array(phase),
            ))
    walls = []
    for wl in s.walls:
        u, w = _gram_schmidt(wl.u, wl.w)
        walls.append(dict(
            shape=wl.shape,
            l1=0.5 * wl.l1 / L, l2=0.5 * wl.l2 / L, l3=0.5 * wl.l3 / L,
            stiffness=wl.stiffness * t2 / M, shear_stiffness=wl.shear_stiffness * t2 / M,
            u=np.array(u), w=np.array(w), n=np.array(_cross(u, w)),
            position=np.array([p / L for p in wl.position]),
            velocity=np.array([v * T / L for v in wl.velocity]),
            omega=np.array([float(o) for o in wl.omega]),
        ))
    r = Reduced(
        nc=nc, nw=nw, pbc=[int(bool(p)) for p in s.pbc], method_order=s.method_order,
        n_euler=s.n_euler, lbox=lbox, gravity=gravity, time_step=dt_r,
        number_of_steps=nsteps, radius=radius_r, mass=mass_r, inertia=inertia_r,
        density=density_r, inv_mass=None, inv_inertia=None, eff_mass=None,
        min_dij=None, min_dij_sq=None, kn=kn_r, yn=yn_r, kt=kt_r, yt=yt_r,
        mu=np.asarray(s.mu, float).reshape(nc, nc), musq=None, knw=knw_r, ynw=ynw_r,
        ktw=ktw_r, ytw=ytw_r, muw=np.asarray(s.muw, float).reshape(nc, nw), muwsq=None,
        walls=walls, length_scale=L, mass_scale=M, time_scale=T, max_ncell=max_ncell,
        wall_motion=wall_motion,
    )
    if s.fluid_flow:
        # ScaleVariables and ScaleFluidProperties (src/DEM.f90:858,
        # src/mFluid.f90:61-66)
        r.fluid = dict(density=s.fluid_density * ((L * L) * L) / M,
                       viscosity=s.fluid_viscosity * L * T / M,
                       flow_rate=s.fluid_flow_rate * T / ((L * L) * L),
                       nl=int(s.fluid_layers))
    initialize_force(r, s.fluid_density * ((L * L) * L) / M)
    return r


def initialize_force(r: Reduced, fluid_density: float = 0.0) -> Reduced:
    """InitializeForce (src/MForces.f90:54-87) on reduced constants."""
    nc = r.nc
    md = np.zeros((nc, nc))
    mdsq = np.zeros((nc, nc))
    for i in range(nc):
        md[i, i] = 2.0 * r.radius[i]
        mdsq[i, i] = md[i, i] * md[i, i]
        for j in range(i + 1, nc):
            md[i, j] = md[j, i] = r.radius[i] + r.radius[j]
            mdsq[i, j] = mdsq[j, i] = md[i, j] * md[i, j]
    r.min_dij = md
    r.min_dij_sq = mdsq
    r.inv_mass = np.array([1.0 / m for m in r.mass])
    r.inv_inertia = np.array([1.0 / i for i in r.inertia])
    r.eff_mass = np.array([m * (rho - fluid_density) / rho for m, rho in zip(r.mass, r.density)])
    r.musq = r.mu * r.mu
    r.muwsq = r.muw * r.muw
    return r


def round_half_away(x: float) -> float:
    """Fortran nint."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def reference_grid(r: Reduced, max_ncell: Optional[int] = None):
    """Grid of InitCellStructure3D (src/MCells.f90:155-227) via InitializeCellList
    (src/MForces.f90:91-106)."""
    if r.ref_grid is not None:
        return tuple(int(x) for x in r.ref_grid)
    cap = r.max_ncell if max_ncell is None else max_ncell
    cutoff = 2.0 * float(np.max(r.radius))
    lx, ly, lz = (float(x) for x in r.lbox)
    n4 = [int(lx / cutoff), int(ly / cutoff), int(lz / cutoff)]
    if n4[0] * n4[1] * n4[2] > cap:
        lcell = (lx * ly * lz / float(cap)) ** (1.0 / 3.0)
        return int(lx / lcell), int(ly / lcell), int(lz / lcell)
    return tuple(n4)
