"""The host restatement of the reference's set-up arithmetic -- CalculateProperties
and ScaleVariables (src/DEM.f90:245-356, 807-866), dem3d_amd/spec.py
reduce_spec -- pinned to the reference's own program.

The reference's program DEM (src/DEM.f90 and its modules, unchanged, built
into oracle/_ref by oracle/build_ref.sh) runs a few steps of a two-component
specification with a wall and writes its final configuration.
SaveConfiguration (src/MConfig.f90:109-240) stores the module variables as
they stand after ScaleVariables: the three scale factors, Lbox, each
component's radius, density, mass and inertia, and each wall's half lengths,
position and stiffnesses, all in reduced units and as raw binary doubles.
reduce_spec of the same specification must give every one of them bit for
bit.  No GPU is needed (the configuration reader is host code of the product
library).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from dem3d_amd import files, spec
from dem3d_amd.native import load_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG_REF = os.path.join(ROOT, "oracle", "_ref", "dem3d_ref")

# (number, diameter, density, stiffness, shear stiffness) as the .inp holds them
COMPONENTS = [("140", "0.0010", "2500", "2.e3", "5.7142857142857e2"),
              ("80", "0.0013", "2700", "3.e3", "8.5714285714286e2")]
WALL = ("1", "0.032", "0.032", "0.0", "2.5e3", "7.1428571428571e2", "0.0 0.0 -0.005",
        "1. 0. 0.", "0. 1. 0.")
BOX = ("0.008", "0.008", "0.010")
TOTAL = "0.0002"


def _inp() -> str:
    comp = "\n".join(" ".join(c) for c in COMPONENTS)
    wall = "  ".join(WALL)
    return f"""[Box dimensions (m)]
{' '.join(BOX)}
[Number of components - NC]
2
[Component properties (NC rows): Initial Number, Diameter(m), Density(Kg/m^3), StiffNess(N/m), Shear StiffNess(N/m)]
{comp}
[Read the initial configuration? (If yes, just type the file name)]
no
[Promote particle pouring? (If yes, just type the pouring file name)]
no
[Number of Walls - NW]
1
[Wall properties (NW rows): Shape*, L1(m), L2(m), L3(m), Stiffness(N/m), Shear-StiffNess(N/m), Position(m), U, W]
{wall}
[Normal coefficients of restitution of the component-component pairs (NC by NC matrix)]
0.3 0.4
0.4 0.5
[Tangential coefficients of restitution of the component-component pairs (NC by NC matrix)]
0.3 0.4
0.4 0.5
[Coefficients of friction of the component-component pairs (NC by NC matrix)]
0.5 0.5
0.5 0.5
[Normal coefficients of restitution of the component-wall pairs (NC by NW matrix)]
0.3
0.6
[Tangential coefficients of restitution of the component-wall pairs (NC by NW matrix)]
0.3
0.6
[Coefficients of friction of the component-wall pairs (NC by NW matrix)]
0.5
0.5
[Periodical Boundary Conditions]
T T F
[Gravity Acceletarion Vector (m/s^2)]
0.0 0.0 -9.81
[Fluid density (Kg/m^3)]
0.0
[Seed for random number generation (a negative integer)]
-7
[Base for file names (with no extention; maximum 100 characters)]
case
[Total time to be simulated (s)]
{TOTAL}
[Minimum number of time steps per collision]
50.
[Order of the predictor-corrector integration method (1-4)]
4
[Number of intervals in each initial euler time step]
10
[Time interval for saving the state of the simulation (s)]
{TOTAL}
[Time interval for saving snapshots of the simulation (s)]
{TOTAL}
[Time interval for reporting the progress of the simulation (s)]
{TOTAL}
[Resume simulation (if it was interrupted)?]
no
[Move walls?]
no
[Vectors A (m), f (1/s), and phase for wall velocities (NW rows)]
[Simulate fluid flow?]
no
"""


def _spec() -> spec.SystemSpec:
    f = float
    vec = lambda s: [f(x) for x in s.split()]
    return spec.SystemSpec(
        lbox=[f(x) for x in BOX],
        components=[spec.Component(diameter=f(c[1]), density=f(c[2]), stiffness=f(c[3]),
                                   shear_stiffness=f(c[4])) for c in COMPONENTS],
        walls=[spec.Wall(shape=int(WALL[0]), l1=f(WALL[1]), l2=f(WALL[2]), l3=f(WALL[3]),
                         stiffness=f(WALL[4]), shear_stiffness=f(WALL[5]),
                         position=vec(WALL[6]), u=vec(WALL[7]), w=vec(WALL[8]))],
        en=np.array([[0.3, 0.4], [0.4, 0.5]]), et=np.array([[0.3, 0.4], [0.4, 0.5]]),
        mu=np.array([[0.5, 0.5], [0.5, 0.5]]),
        enw=np.array([[0.3], [0.6]]), etw=np.array([[0.3], [0.6]]),
        muw=np.array([[0.5], [0.5]]),
        pbc=[True, True, False], gravity=[0.0, 0.0, -9.81], total_time=f(TOTAL),
        min_steps_per_collision=50.0, method_order=4, n_euler=10)


@pytest.mark.skipif(not os.path.exists(PROG_REF),
                    reason="the reference program is not built (oracle/build_ref.sh)")
def test_reduce_spec_equals_the_reference_scaling(tmp_path):
    with open(tmp_path / "case.inp", "w") as fh:
        fh.write(_inp())
    with open(tmp_path / "case.inp") as inp:
        r = subprocess.run([PROG_REF], stdin=inp, cwd=tmp_path, capture_output=True, text=True,
                           timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    cfg = files.read_config(load_hip(), str(tmp_path / "case.cfg"))
    red = spec.reduce_spec(_spec())
    # LengthScale, MassScale, TimeScale (src/DEM.f90:814-816)
    assert cfg["length_scale"] == red.length_scale
    assert cfg["mass_scale"] == red.mass_scale
    assert cfg["time_scale"] == red.time_scale
    assert np.array_equal(cfg["lbox"], red.lbox)
    # CalculateProperties then ScaleVariables, per component
    for k in ("radius", "density", "mass", "inertia"):
        assert np.array_equal(np.asarray(cfg[k]), np.asarray(getattr(red, k))), k
    # the wall: ConfigureWalls halves the lengths; ScaleVariables scales them
    w, wr = cfg["walls"][0], red.walls[0]
    for k in ("l1", "l2", "l3", "stiffness", "shear_stiffness"):
        assert w[k] == wr[k], k
    for k in ("position", "u", "w", "n"):
        assert np.array_equal(w[k], wr[k]), k
