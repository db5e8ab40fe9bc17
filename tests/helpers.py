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
