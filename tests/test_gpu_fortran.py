"""The reference driver's time loop in Fortran (dem3d_amd/fortran/drive_steps.f90)
calling the GPU path through the iso_c_binding module dem3d_gpu.f90, as the
DEM3D driver would (MoveParticlesEuler start-up, then MoveParticles steps):
the state it reaches equals bitwise the state the same calls reach through
ctypes, and matches the oracle within the parity tolerances."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.build import DRIVER
from dem3d_amd.native import System, make_params, make_walls
from checkers import oracle_system

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")]

NSTEPS = 40


def write_input(path, wl, nsteps):
    red = wl.red
    p = make_params(red)
    nc, nw, n = red.nc, red.nw, wl.np
    hdr = np.array([nc, nw, n, red.method_order, red.n_euler, p.max_ncell, *p.pbc,
                    *p.ref_grid, p.precision, p.history_mode, nsteps, 0], np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([*p.lbox, *p.gravity, p.time_step, p.skin], np.float64).tobytes())
        for name in ("radius", "mass", "inertia", "inv_mass", "inv_inertia", "eff_mass",
                     "min_dij", "min_dij_sq", "kn", "yn", "kt", "yt", "musq",
                     "knw", "ynw", "ktw", "ytw", "muwsq"):
            f.write(p._keep[name].tobytes())
        walls = make_walls(red)
        f.write(bytes(walls))
        f.write(np.ascontiguousarray(wl.comp, np.int32).tobytes())
        for q in (wl.x, wl.v, wl.w):
            f.write(np.ascontiguousarray(np.asarray(q, np.float64).reshape(3, n)).tobytes())


def read_output(path, n):
    raw = open(path, "rb").read()
    head = np.frombuffer(raw[:40], np.int64)
    arr = np.frombuffer(raw[40:], np.float64).reshape(5, 3, n)
    return dict(rc=int(head[0]), done=int(head[1]), ncol=int(head[2]), pairs=int(head[3]),
                walls=int(head[4]), x=arr[0], v=arr[1], w=arr[2], a=arr[3], alpha=arr[4])


@pytest.mark.parametrize("case", [
    dict(n_side=(7, 6, 5), nc=2, seed=61, vel=0.4),
    dict(n_side=(6, 6, 6), nc=3, ratio=1.4, seed=67, walls=False, pbc=(True, True, True),
         spacing=1.0, vel=0.5),
])
def test_fortran_driver_equals_ctypes_and_oracle(hip, tmp_path, case):
    wl = workloads.small_box(**case)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(inp, wl, NSTEPS)
    print(f"drive_steps {case}", flush=True)
    r = subprocess.run([DRIVER, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f = read_output(out, wl.np)
    assert f["rc"] == 0 and f["done"] == NSTEPS
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    g.step(NSTEPS)
    st, stats = g.download(), g.stats()
    for k in ("x", "v", "w", "a", "alpha"):
        assert np.array_equal(f[k], st[k]), k
    assert (f["ncol"], f["pairs"], f["walls"]) == (stats["ncol"], stats["pair_contacts"],
                                                   stats["wall_contacts"])
    o = oracle_system(wl.red)
    o.upload(wl.comp, wl.x, wl.v, wl.w)
    o.startup()
    o.step(NSTEPS)
    so = o.download()
    for k, tol in (("x", 1e-11), ("v", 1e-9)):
        scale = np.abs(so[k]).max() + 1e-300
        assert np.abs(f[k] - so[k]).max() <= tol * scale, k
    assert f["ncol"] == o.stats()["ncol"]
    g.close()
