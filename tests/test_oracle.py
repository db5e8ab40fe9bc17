"""Pins the oracle (oracle/dem3d_oracle.c) to the reference.

1. Golden fixtures written by the reference Fortran itself
   (tests/golden/make_golden.py): the oracle must reproduce every output bit
   for bit -- accelerations, positions, velocities, the ABM ring, contact
   lists with their history vectors, CellIndex, the cell-list order, Ncol and
   the kinetic energy.
2. Where the reference build exists (this container), the same comparison
   against freshly run reference code on further cases.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from dem3d_amd.spec import Reduced
from checkers import (cells_to_rank, contacts_to_keys, oracle_system, ref_available,
                      run_reference)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture(name):
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return {k: z[k] for k in z.files}


def reduced_from_fixture(f) -> Reduced:
    ints = f["red_ints"]
    nc, nw, order, neuler = (int(x) for x in ints[:4])
    pbc = [int(x) for x in ints[4:7]]
    sc = f["red_scalars"]
    walls = []
    for j in range(nw):
        s = f[f"wall{j}_scalars"]
        w = dict(shape=int(s[0]), l1=s[1], l2=s[2], l3=s[3], stiffness=s[4],
                 shear_stiffness=s[5])
        for k in ("u", "w", "n", "position", "velocity", "omega"):
            w[k] = f[f"wall{j}_{k}"]
        walls.append(w)
    wm = None
    if nw and "wm0_amplitude" in f:
        wm = [{k: f[f"wm{j}_{k}"] for k in ("constant_factor", "amplitude", "velocity_amplitude",
                                            "two_pi_frequency", "phase_angle")}
              for j in range(nw)]
    g = lambda k: f["red_" + k]
    return Reduced(nc=nc, nw=nw, pbc=pbc, method_order=order, n_euler=neuler, lbox=g("lbox"),
                   gravity=g("gravity"), time_step=float(sc[0]), number_of_steps=int(ints[7]),
                   radius=g("radius"), mass=g("mass"), inertia=g("inertia"), density=g("density"),
                   inv_mass=g("inv_mass"), inv_inertia=g("inv_inertia"), eff_mass=g("eff_mass"),
                   min_dij=g("min_dij"), min_dij_sq=g("min_dij_sq"), kn=g("kn"), yn=g("yn"),
                   kt=g("kt"), yt=g("yt"), mu=g("mu"), musq=g("musq"), knw=g("knw"), ynw=g("ynw"),
                   ktw=g("ktw"), ytw=g("ytw"), muw=g("muw"), muwsq=g("muwsq"), walls=walls,
                   length_scale=float(sc[1]), mass_scale=float(sc[2]), time_scale=float(sc[3]),
                   wall_motion=wm)


def assert_snapshot(o, f, prefix):
    st = o.download()
    for k in ("x", "v", "w", "a", "alpha"):
        assert np.array_equal(st[k], f[f"{prefix}_{k}"]), (prefix, k)
    assert np.array_equal(o.download_cells(), f[f"{prefix}_cell_index"]), prefix
    con = o.download_contacts()
    for k in ("keys", "ft", "vt", "wkeys", "wft", "wvt"):
        a, b = con[k], f[f"{prefix}_{k}"]
        assert a.shape[0] == b.shape[0], (prefix, k, a.shape, b.shape)
        assert np.array_equal(a.reshape(b.shape), b), (prefix, k)
    nprev = o.red.method_order - 1
    assert np.array_equal(st["ring"], f[f"{prefix}_ring"][:nprev]), prefix
    assert np.array_equal(st["ipc"], f[f"{prefix}_ipc"]), prefix
    s = o.stats()
    assert s["ncol"] == int(f[f"{prefix}_ncol"]), prefix
    assert s["kinetic_energy"] == float(f[f"{prefix}_ke"]), prefix
    if f"{prefix}_time" in f:
        assert s["time"] == float(f[f"{prefix}_time"]), prefix
        if o.red.nw:
            wl = o.download_walls()
            assert np.array_equal(wl["position"], f[f"{prefix}_wall_pos"]), prefix
            assert np.array_equal(wl["velocity"], f[f"{prefix}_wall_vel"]), prefix
    # cell-list order (the reference's head-pushed lists)
    import ctypes as C
    r = np.zeros(o.np, np.int32)
    assert o.lib.download_cell_rank(o.h, o.np, r.ctypes.data_as(C.POINTER(C.c_int32))) == 0
    lens, mem = f[f"{prefix}_cell_len"], f[f"{prefix}_cell_members"]
    cells, k = [], 0
    for n in lens:
        cells.append(list(mem[k:k + n]))
        k += n
    assert np.array_equal(r, cells_to_rank(cells, o.np)), prefix


def run_actions(o, actions):
    """Apply the fixture's action list to the oracle and yield at snapshots."""
    for act in actions:
        if act[0] == "compute":
            o.compute_accelerations(0, bool(act[1]))
        elif act[0] == "euler":
            o.euler_startup()
        elif act[0] == "step":
            o.step(act[2], act[1])
        elif act[0] == "snapshot":
            yield act[1]


FIXTURES = {
    "ref_small.npz": [("compute", 1), ("snapshot", "c0"), ("euler",), ("euler",), ("euler",),
                      ("snapshot", "e"), ("step", 4, 20), ("snapshot", "s")],
    "ref_mono_settling_2k.npz": [("euler",), ("euler",), ("euler",), ("snapshot", "e"), ("step", 4, 60),
                        ("snapshot", "s")],
    "ref_capped.npz": [("compute", 1), ("snapshot", "c0")],
    "ref_cone.npz": [("euler",), ("euler",), ("euler",), ("snapshot", "e"), ("step", 4, 30),
                     ("snapshot", "s")],
    "ref_ball_bearings.npz": [("euler",), ("euler",), ("euler",), ("snapshot", "e"),
                              ("step", 4, 1500), ("snapshot", "s")],
}


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_oracle_reproduces_reference_fixture(name):
    f = load_fixture(name)
    red = reduced_from_fixture(f)
    o = oracle_system(red)
    o.red = red
    o.upload(f["comp"], f["x0"], f["v0"], f["w0"])
    assert list(o.stats()["ref_ncell"]) == list(f["grid"])
    for prefix in run_actions(o, FIXTURES[name]):
        assert_snapshot(o, f, prefix)


def test_fixtures_exercise_their_features():
    """The fixtures cover what they claim: contacts with history, wall
    contacts, the MaxNCell cap, and contacts with the cone wall."""
    small = load_fixture("ref_small.npz")
    assert len(small["s_keys"]) > 20 and len(small["s_wkeys"]) > 0
    cap = load_fixture("ref_capped.npz")
    red = reduced_from_fixture(cap)
    natural = [int(l / (2 * red.radius.max())) for l in red.lbox]
    assert np.prod(natural) > 30000 and np.prod(cap["grid"]) <= 30000
    cone = load_fixture("ref_cone.npz")
    assert np.any((cone["s_wkeys"] & 0xffffffff) == 2), "no contact with the cone"
    bb = load_fixture("ref_ball_bearings.npz")
    red = reduced_from_fixture(bb)
    assert red.wall_motion is not None and len(bb["s_keys"]) > 10
    touched = set(int(k) for k in (bb["s_wkeys"] & 0xffffffff))
    assert {1, 2} <= touched, touched          # the inner cylinder and its disc
    moved = bb["s_wall_pos"] - np.array([w["position"] for w in red.walls])
    assert np.abs(moved[:, :2]).max() > 0.1    # shaken by more than 1 mm
    assert np.abs(bb["s_wall_vel"]).max() > 0.0


def test_derived_tables_match_initialize_force():
    """spec.initialize_force restates InitializeForce (src/MForces.f90:54-87)
    bit for bit: compare with the tables the reference computed."""
    for name in FIXTURES:
        f = load_fixture(name)
        red = reduced_from_fixture(f)
        for k in ("min_dij", "min_dij_sq", "inv_mass", "inv_inertia", "eff_mass", "musq", "muwsq"):
            a = np.asarray(getattr(red, k)).reshape(f["tab_" + k].shape)
            assert np.array_equal(a, f["tab_" + k]), (name, k)


# ---------------------------------------------------------------------------
# live reference (only where the reference build exists)

live = pytest.mark.skipif(not ref_available(), reason="reference build not present")


@live
@pytest.mark.parametrize("seed,nc,pbc,walls", [
    (21, 1, (True, True, False), True),
    (22, 3, (True, True, True), False),
    (23, 2, (False, False, False), True),
    (24, 4, (True, False, True), True),
])
def test_oracle_matches_live_reference(seed, nc, pbc, walls):
    from dem3d_amd import workloads
    wl = workloads.small_box(n_side=(6, 5, 7), nc=nc, ratio=1.35, seed=seed, pbc=pbc, walls=walls,
                             vel=0.6)
    actions = [("compute", 1), ("snapshot", "c0"), ("compute", 0), ("snapshot", "c1"),
               ("euler",), ("euler",), ("euler",), ("snapshot", "e"), ("step", 4, 40),
               ("snapshot", "s")]
    res = run_reference(dict(red=wl.red, comp=wl.comp, x=wl.x, v=wl.v, w=wl.w, actions=actions))
    o = oracle_system(wl.red)
    o.red = wl.red
    o.upload(wl.comp, wl.x, wl.v, wl.w)
    snaps = iter(res["snapshots"])
    for _ in run_actions(o, actions):
        snap = next(snaps)
        st = o.download()
        for k in ("x", "v", "w", "a", "alpha"):
            assert np.array_equal(st[k], snap[k]), k
        keys, ft, vt = contacts_to_keys(snap["state"]["pairs"])
        con = o.download_contacts()
        assert np.array_equal(con["keys"], keys)
        assert np.array_equal(con["ft"].reshape(ft.shape), ft)
        assert np.array_equal(o.download_cells(), snap["cell_index"])
        assert o.stats()["ncol"] == snap["ncol"]
        # KineticEnergy and PotentialEnergy (src/MDEM.f90:183-213), bit for bit
        assert o.stats()["kinetic_energy"] == snap["ke"]
        assert o.lib.potential_energy(o.h) == snap["pe"]


@live
def test_resume_from_state_file_matches():
    """Load a mid-run state (cell lists in the reference's order, ABM ring,
    contact lists) into both the reference (through its own ReadState
    routines) and the oracle, then step both."""
    from dem3d_amd import workloads
    wl = workloads.small_box(n_side=(6, 6, 6), nc=2, seed=31, vel=0.4)
    pre = run_reference(dict(red=wl.red, comp=wl.comp, x=wl.x, v=wl.v, w=wl.w,
                             actions=[("euler",), ("euler",), ("euler",), ("step", 4, 25),
                                      ("snapshot", "m")]))["snapshots"][0]
    st = pre["state"]
    job = dict(red=wl.red, comp=wl.comp, x=pre["x"], v=pre["v"], w=pre["w"], a=pre["a"],
               alpha=pre["alpha"], state=st, actions=[("step", 29, 15), ("snapshot", "r")])
    post = run_reference(job)["snapshots"][0]
    o = oracle_system(wl.red)
    o.red = wl.red
    o.upload(wl.comp, pre["x"], pre["v"], pre["w"], pre["a"], pre["alpha"], st["ring"],
             st["ipc"], cells_to_rank(st["cells"], wl.np))
    keys, ft, vt = contacts_to_keys(st["pairs"])
    wk, wft, wvt = contacts_to_keys(st["walls"])
    o.upload_contacts(keys, ft, vt, wk, wft, wvt)
    o.step(15, 29)
    s = o.download()
    for k in ("x", "v", "w", "a", "alpha"):
        assert np.array_equal(s[k], post[k]), k
    assert np.array_equal(o.download_contacts()["keys"], contacts_to_keys(post["state"]["pairs"])[0])


@live
def test_oracle_matches_live_reference_moving_floor():
    """MoveWalls on a Rectangle floor shaken along all three axes with
    different frequencies and phases (one negative): positions, velocities,
    wall state and Time stay bitwise equal through the Euler start-up and the
    steps."""
    from helpers import shaken_floor
    wl = shaken_floor()
    red = wl.red
    actions = [("euler",), ("euler",), ("euler",), ("snapshot", "e"), ("step", 4, 600),
               ("snapshot", "s")]
    res = run_reference(dict(red=red, comp=wl.comp, x=wl.x, v=wl.v, w=wl.w, actions=actions))
    o = oracle_system(red)
    o.red = red
    o.upload(wl.comp, wl.x, wl.v, wl.w)
    snaps = iter(res["snapshots"])
    for _ in run_actions(o, actions):
        snap = next(snaps)
        st = o.download()
        for k in ("x", "v", "w", "a", "alpha"):
            assert np.array_equal(st[k], snap[k]), k
        walls = o.download_walls()
        assert np.array_equal(walls["position"], snap["wall_pos"])
        assert np.array_equal(walls["velocity"], snap["wall_vel"])
        assert o.stats()["time"] == snap["time"]
        wk, wft, _ = contacts_to_keys(snap["state"]["walls"])
        con = o.download_contacts()
        assert np.array_equal(con["wkeys"], wk)
        assert np.array_equal(con["wft"].reshape(wft.shape), wft)
    assert len(wk) > 10
