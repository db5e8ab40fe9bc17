"""Shared comparison helpers of the parity tests."""
from __future__ import annotations

import numpy as np

from dem3d_amd.native import System


def upload_same(systems, wl, state=None, contacts=None):
    """Upload the same initial state (fresh cell lists unless `state` carries
    an explicit order) to every system."""
    for s in systems:
        if state is None:
            s.upload(wl.comp, wl.x, wl.v, wl.w)
        else:
            s.upload(state["comp"], state["x"], state["v"], state["w"], state.get("a"),
                     state.get("alpha"), state.get("ring"), state.get("ipc"),
                     state.get("cell_rank"))
        if contacts is not None:
            s.upload_contacts(contacts["keys"], contacts["ft"], contacts["vt"], contacts["wkeys"],
                              contacts["wft"], contacts["wvt"])


def full_state(s: System, with_rank=False):
    st = s.download()
    if with_rank:
        import ctypes as C
        r = np.zeros(s.np, np.int32)
        rc = s.lib.download_cell_rank(s.h, s.np, r.ctypes.data_as(C.POINTER(C.c_int32)))
        assert rc == 0
        st["cell_rank"] = r
    return st


def force_scale(o: System):
    """|m_eff g| + sum |F_N| + |F_T| per particle (and the torque analogue) from
    the oracle's last evaluation, converted to accelerations."""
    import ctypes as C
    n = o.np
    sc = np.zeros(2 * n)
    rc = o.lib.download_force_scale(o.h, n, sc.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    return sc[:n], sc[n:]


def accel_tolerance_ok(a_gpu, a_ref, scale, inv, rel=1e-12):
    """|delta a_i| <= rel * scale_i * inv_mass_i, per particle."""
    err = np.abs(a_gpu - a_ref).max(axis=0)
    bound = rel * scale * inv + 1e-300
    return np.all(err <= bound), float(np.max(err / bound))


def compare_contacts(cg, co, exact=True, rtol=0.0):
    assert np.array_equal(cg["keys"], co["keys"]), (len(cg["keys"]), len(co["keys"]))
    assert np.array_equal(cg["wkeys"], co["wkeys"]), (len(cg["wkeys"]), len(co["wkeys"]))
    for k in ("ft", "vt", "wft", "wvt"):
        if exact:
            assert np.array_equal(cg[k], co[k]), (k, np.abs(cg[k] - co[k]).max())
        else:
            scale = np.abs(co[k]).max() + 1e-300
            assert np.abs(cg[k] - co[k]).max() <= rtol * scale, k


def shaken_floor(seed=41):
    """A bed of two components settled by strong gravity (2000 m/s^2) on a
    floor that MoveWalls shakes along x, y and z at 40, 55 and 70 Hz with
    phases 0, -0.3 and 0.6 of a period."""
    from dem3d_amd import workloads
    wl = workloads.small_box(n_side=(5, 5, 3), nc=2, ratio=1.2, seed=seed, vel=0.05, e=0.3,
                             gravity=-2000.0)
    red = wl.red
    T, L = red.time_scale, red.length_scale
    amp = np.array([0.05e-3, 0.03e-3, 0.02e-3]) / L
    tpf = 2.0 * np.pi * np.array([40.0, 55.0, 70.0]) * T
    phase = 2.0 * np.pi * np.array([0.0, -0.3, 0.6])
    pos = np.asarray(red.walls[0]["position"])
    red.wall_motion = [dict(constant_factor=pos - amp * np.cos(phase), amplitude=amp,
                            velocity_amplitude=-tpf * amp, two_pi_frequency=tpf,
                            phase_angle=phase)]
    return wl


def stale_pair_case(kind: str):
    """Two spheres (d = 1 mm) in a box periodic in x, y, z of 5.01 x 4.08 x
    4.08 d, no gravity: 5 reference cells of 1.002 d along x, 4 of 1.02 d
    across.  Sphere 1 rests just below the top of reference cell 1 along x;
    sphere 2 moves along x at 0.004 d per time step, a displacement larger
    than the 0.002 d by which a cell exceeds a diameter, so the pair can
    pass between overlapping and non-adjacent reference cells in one step.

    kind = "separate": sphere 2 overlaps sphere 1 by 0.001 d with a stored
    (FT, VT), moving away: one step later the pair is neither overlapping nor
    in adjacent cells, so the reference keeps its history (stale); it then
    leaves the neighbour list, travels round the periodic box and becomes
    adjacent again from the other side, where the reference forgets it.

    kind = "return": the pair starts 2.2 d apart, in cells two apart, with a
    stored (FT, VT) (a stale pair, as a state file of the reference can hold
    one); sphere 2 approaches, and one step takes it from cell 3 (not
    adjacent) into an overlap, so the contact starts from the stale history
    (no new collision counted).

    Returns (workload, contacts, steps)."""
    from dem3d_amd import workloads
    from dem3d_amd.spec import reduce_spec
    d = 1e-3
    lx, ly = 5.01 * d, 4.08 * d
    spec = workloads._spec((lx, ly, ly), [d], 2500.0, 1e3, 0.9, 0.5, pbc=(True, True, True),
                           gravity=(0.0, 0.0, 0.0))
    red = reduce_spec(spec)
    L = red.length_scale
    dr, Lx, Ly = d / L, lx / L, ly / L
    w = Lx / 5
    step = 0.004 * dr
    xa = -0.5 * Lx + 2 * w - 0.0005 * dr
    yz = -0.5 * Ly + 1.5 * (Ly / 4)
    if kind == "separate":
        xb, vb, steps = xa + 0.999 * dr, step, 800
    elif kind == "return":
        xb, vb, steps = -0.5 * Lx + 3 * w + 0.001 * dr + 300 * step, -step, 330
    else:
        raise ValueError(kind)
    x = np.array([[xa, xb], [yz, yz], [yz, yz]])
    v = np.zeros((3, 2))
    v[0, 1] = vb / red.time_step
    kn = float(np.asarray(red.kn)[0, 0])
    f = 1e-4 * kn * dr
    contacts = dict(keys=np.array([(1 << 32) | 2], np.uint64),
                    ft=np.array([[0.0, 0.3 * f, -0.2 * f]]),
                    vt=np.array([[0.0, 2e-3 * dr, 1e-3 * dr]]),
                    wkeys=np.zeros(0, np.uint64), wft=np.zeros((0, 3)), wvt=np.zeros((0, 3)))
    wl = workloads.Workload(f"stale-{kind}", spec, red, np.ones(2, np.int32), x, v,
                            np.zeros((3, 2)), steps)
    return wl, contacts, steps
