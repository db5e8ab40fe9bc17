"""The benchmark configurations at their own sizes, on the GPU.

* 2 000 monodisperse spheres settling (mono_settling_2k): the GPU against
  the reference's own output for the start-up and 60 steps
  (tests/golden/ref_mono_settling_2k.npz, written by the reference).
* 200 000 polydisperse spheres in 16 components (poly_settling_200k): from a
  state with contacts and history that the GPU reached, one evaluation and
  three ABM steps against the oracle (FP64 bar), and the mixed mode by
  shadow evaluation -- at each of three successive FP64 states, the mixed
  forces of the same state against the oracle's (mixed bar).
* 4 000 000 spheres of the sheared bed (bidisperse_sheared_4m), too large
  for the oracle: the pair contacts the force pass counts at the current
  positions equal the overlapping pairs an independent k-d tree search finds
  on the downloaded positions with the same overlap test
  (MinDijSq - DijSq > 0, minimum image in x and y).
"""
from __future__ import annotations

import numpy as np
import pytest

from dem3d_amd import workloads
from dem3d_amd.native import DEM_MIXED, System
from checkers import oracle_system
from helpers import accel_tolerance_ok, compare_contacts, force_scale, full_state, upload_same
from test_oracle import load_fixture, reduced_from_fixture

pytestmark = pytest.mark.gpu


def _close(a, b, rel):
    scale = np.abs(b).max() + 1e-300
    return np.abs(a - b).max() <= rel * scale


def _inv(wl_comp, red):
    c = wl_comp - 1
    return red.inv_mass[c], red.inv_inertia[c]


def test_config0_mono_2k_against_reference_fixture(hip):
    f = load_fixture("ref_mono_settling_2k.npz")
    red = reduced_from_fixture(f)
    g = System(hip, red)
    g.upload(f["comp"], f["x0"], f["v0"], f["w0"])
    g.startup()
    for prefix, steps in (("e", 0), ("s", 60)):
        if steps:
            g.step(steps)
        st = g.download()
        for k, rel in (("x", 1e-12), ("v", 1e-10), ("w", 1e-10), ("a", 1e-10),
                       ("alpha", 1e-10)):
            assert _close(st[k], f[f"{prefix}_{k}"], rel), (prefix, k)
        assert np.array_equal(g.download_cells(), f[f"{prefix}_cell_index"]), prefix
        nprev = red.method_order - 1
        assert _close(st["ring"], f[f"{prefix}_ring"][:nprev], 1e-10), prefix
        assert np.array_equal(st["ipc"], f[f"{prefix}_ipc"]), prefix
        s = g.stats()
        assert s["ncol"] == int(f[f"{prefix}_ncol"]), prefix
        assert abs(s["kinetic_energy"] - float(f[f"{prefix}_ke"])) <= 1e-10 * float(
            f[f"{prefix}_ke"]), prefix
        con = g.download_contacts()
        assert np.array_equal(con["keys"], f[f"{prefix}_keys"]), prefix
        assert np.array_equal(con["wkeys"], f[f"{prefix}_wkeys"]), prefix


@pytest.fixture(scope="module")
def poly200k(hip):
    wl = workloads.poly_settling_200k()
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    # settle until the bed has a few thousand contacts with history (within
    # the configuration's 5 000 steps)
    for _ in range(5):
        g.step(1000)
        if g.stats()["pair_contacts"] > 2000:
            break
    return wl, g


def test_config1_poly_200k_fp64_against_oracle(hip, poly200k):
    wl, g0 = poly200k
    print(f"poly 200k: {g0.stats()['pair_contacts']} pairs", flush=True)
    st = full_state(g0, with_rank=True)
    con = g0.download_contacts()
    assert len(con["keys"]) > 500
    g, o = System(hip, wl.red), oracle_system(wl.red)
    upload_same([g, o], wl, state=st, contacts=con)
    for s in (g, o):
        s.compute_accelerations(0, True)
    sg, so = g.download(), o.download()
    fs, ts = force_scale(o)
    im, ii = _inv(wl.comp, wl.red)
    ok_a, ra = accel_tolerance_ok(sg["a"], so["a"], fs, im, 1e-12)
    ok_b, rb = accel_tolerance_ok(sg["alpha"], so["alpha"], ts + 1e-300, ii, 1e-12)
    assert ok_a, f"acceleration error {ra:.3g} x tolerance"
    assert ok_b, f"angular acceleration error {rb:.3g} x tolerance"
    assert np.array_equal(g.download_cells(), o.download_cells())
    compare_contacts(g.download_contacts(), o.download_contacts(), exact=True, rtol=1e-12)
    # three ABM steps from the same state
    upload_same([g, o], wl, state=st, contacts=con)
    ncol0 = [s.stats()["ncol"] for s in (g, o)]
    for s in (g, o):
        s.step(3)
    sg, so = g.download(), o.download()
    for k, rel in (("x", 1e-12), ("v", 1e-10), ("w", 1e-10)):
        assert _close(sg[k], so[k], rel), k
    assert np.array_equal(g.download_contacts()["keys"], o.download_contacts()["keys"])
    # new contacts of the three steps (Ncol is cumulative)
    assert g.stats()["ncol"] - ncol0[0] == o.stats()["ncol"] - ncol0[1]
    g.close()


def test_config1_poly_200k_mixed_shadow_evaluation(hip, poly200k):
    wl, g0 = poly200k
    im, ii = _inv(wl.comp, wl.red)
    rad = np.asarray(wl.red.radius)[wl.comp - 1]
    m = System(hip, wl.red, precision=DEM_MIXED)
    o = oracle_system(wl.red)
    for k in range(3):
        st = full_state(g0, with_rank=True)
        con = g0.download_contacts()
        upload_same([m, o], wl, state=st, contacts=con)
        for s in (m, o):
            s.compute_accelerations(0, False)
        sm, so = m.download(), o.download()
        fs, ts = force_scale(o)
        ok_a, ra = accel_tolerance_ok(sm["a"], so["a"], fs, im, 1e-5)
        # torques against the particle's radius times its force scale (a
        # contact's torque is r x F_T): the mixed bar of forces and torques.
        # Against sum r |F_T| alone a nearly head-on new contact, whose tiny
        # tangential velocity is a difference of FP32 vectors, sets the scale.
        ok_b, rb = accel_tolerance_ok(sm["alpha"], so["alpha"], rad * fs, ii, 1e-5)
        _, rb_t = accel_tolerance_ok(sm["alpha"], so["alpha"], ts + 1e-300, ii, 1e-5)
        print(f"state {k}: force error {ra:.3g}, torque error {rb:.3g} of the bar "
              f"({rb_t:.3g} against sum r|F_T|)", flush=True)
        assert ok_a, f"state {k}: acceleration error {ra:.3g} x tolerance"
        assert ok_b, f"state {k}: angular acceleration error {rb:.3g} x tolerance"
        assert m.stats()["pair_contacts"] == o.stats()["pair_contacts"], k
        g0.step(1)
    m.close()


def test_config2_sheared_4m_pair_count_against_kdtree(hip):
    from scipy.spatial import cKDTree
    wl = workloads.bidisperse_sheared_4m()
    red = wl.red
    g = System(hip, red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    g.step(20)
    g.compute_accelerations(0, False)
    p_gpu = g.stats()["pair_contacts"]
    x = g.download()["x"]
    print(f"sheared 4M: {p_gpu} pairs on the GPU", flush=True)
    L = np.asarray(red.lbox, np.float64)
    box = np.array([L[0], L[1], 4.0 * L[2]])
    pts = np.stack([np.mod(x[0] + 0.5 * L[0], L[0]), np.mod(x[1] + 0.5 * L[1], L[1]),
                    x[2] + 1.5 * L[2]], axis=1)
    pts = np.minimum(pts, np.nextafter(box, 0))
    rmax = float(np.max(red.radius))
    pairs = cKDTree(pts, boxsize=box).query_pairs(2.0 * rmax * 1.001, output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    r = x[:, j] - x[:, i]
    for d in range(2):  # minimum image, periodic x and y (invl = 1/L as the kernels use)
        if red.pbc[d]:
            r[d] = r[d] - L[d] * np.trunc(r[d] * (1.0 / L[d]) + np.copysign(0.5, r[d]))
    dsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2]
    c = wl.comp - 1
    mindijsq = np.asarray(red.min_dij_sq).reshape(red.nc, red.nc)[c[i], c[j]]
    p_tree = int(np.count_nonzero(mindijsq - dsq > 0.0))
    assert p_gpu == p_tree, (p_gpu, p_tree)
    assert p_gpu > 3 * wl.np // 2
    g.close()
