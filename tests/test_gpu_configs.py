"""The benchmark configurations at their own sizes, on the GPU.

* 2 000 monodisperse spheres settling (mono_settling_2k): the GPU against
  the reference's own output for the start-up and 60 steps
  (tests/golden/ref_mono_settling_2k.npz, written by the reference).
* 200 000 polydisperse spheres in 16 components, grown into a random loose
  packing at volume fraction 0.45 and settling under gravity
  (poly_loose_packed): from a state with contacts and history that the GPU
  reached, one evaluation and three ABM steps against the oracle (FP64 bar);
  300 steps compared with the oracle after every step (accelerations and
  kinetic energy), the oracle running six 50-step windows from the GPU's
  checkpoints in parallel threads; the configuration's 5 000 steps sampled
  by eleven such windows, one every 500 steps and the last 50; and the mixed mode by shadow evaluation --
  at each of three successive FP64 states, the mixed forces of the same state
  against the oracle's (mixed bar).
* the dense sheared bed, packed disordered at volume fraction 0.60
  (packed_bed): at 960 000 spheres, one updating evaluation and three ABM
  steps against the oracle from the same state -- reference cells, contact
  keys, Ncol and history bit for bit, accelerations within 1e-12 of each
  particle's force scale; at the 4 000 000 spheres of the benchmark, too
  large for the oracle, the pair contacts of the force pass and of
  dem_observables equal the overlapping pairs an independent k-d tree search
  finds on the downloaded positions with the same overlap test.
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
    wl = workloads.poly_loose_packed(hip)
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    # some settling, so that contacts carry history (within the
    # configuration's 5 000 steps)
    g.step(200)
    ob = g.observables()
    print(f"poly 200k: phi {ob['packing_fraction']:.4f}, {ob['pair_contacts']} pairs", flush=True)
    assert 0.40 < ob["packing_fraction"] < 0.50
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


def _oracle_window(wl, nstep, st, con, a, al, ke, keys):
    """The oracle's nstep steps from a GPU checkpoint (state st, contacts
    con) against the GPU's accelerations a, al and kinetic energies ke of the
    same steps and its contact keys at the end.  Returns the worst relative
    errors (first-step a, later a, first-step ke, later ke) and whether the
    contact sets are identical."""
    im, ii = _inv(wl.comp, wl.red)
    o = oracle_system(wl.red)
    upload_same([o], wl, state=st, contacts=con)
    worst = [0.0, 0.0, 0.0, 0.0]
    for k in range(nstep):
        o.step(1)
        so = o.download()
        fs, ts = force_scale(o)
        if k == 0:
            _, ra = accel_tolerance_ok(a[k], so["a"], fs, im, 1.0)
            _, rb = accel_tolerance_ok(al[k], so["alpha"], ts + 1e-300, ii, 1.0)
        else:
            ra = np.abs(a[k] - so["a"]).max() / (fs * im).max()
            rb = np.abs(al[k] - so["alpha"]).max() / (ts * ii).max()
        ko = o.stats()["kinetic_energy"]
        q = 0 if k == 0 else 1
        worst[q] = max(worst[q], ra, rb)
        worst[2 + q] = max(worst[2 + q], abs(ke[k] - ko) / ko)
    same = bool(np.array_equal(o.download_contacts()["keys"], keys))
    o.close()
    return worst, same


def _windows_against_oracle(wl, g, starts, nstep):
    """The GPU system g steps to each window's first step, then one step at a
    time for nstep steps, keeping the accelerations of every step's first
    evaluation and the kinetic energy; from each window's starting state the
    oracle takes the same steps in a thread of its own, while the GPU goes
    on.  Asserts the bars of the window test on every window."""
    import threading
    out = [None] * len(starts)

    def run(w, *args):
        try:
            out[w] = _oracle_window(wl, nstep, *args)
        except BaseException as exc:  # reported by the main thread
            out[w] = exc

    threads = []
    done = 0
    for w, s0 in enumerate(starts):
        if s0 > done:
            g.step(s0 - done)
        st = full_state(g, with_rank=True)
        con = g.download_contacts()
        a, al, ke = [], [], []
        for _ in range(nstep):
            g.step(1)
            s = g.download()
            a.append(s["a"])
            al.append(s["alpha"])
            ke.append(g.stats()["kinetic_energy"])
        done = s0 + nstep
        keys = g.download_contacts()["keys"]
        t = threading.Thread(target=run, args=(w, st, con, a, al, ke, keys))
        t.start()
        threads.append(t)
        print(f"window {w} (steps {s0}-{s0 + nstep}): GPU done, oracle started", flush=True)
    for t in threads:
        t.join()
    for w, r in enumerate(out):
        assert not isinstance(r, BaseException), r
        worst, same = r
        print(f"window {w}: accelerations {worst[0]:.3g} / {worst[1]:.3g}, kinetic energy "
              f"{worst[2]:.3g} / {worst[3]:.3g} (first step / later steps)", flush=True)
        assert worst[0] <= 1e-12 and worst[1] <= 1e-10, (w, worst)
        assert worst[2] <= 1e-12 and worst[3] <= 1e-10, (w, worst)
        assert same, w


def test_config1_every_step_for_300_steps(hip, poly200k):
    """Six consecutive windows of 50 steps.  The GPU steps one step at a
    time, keeping the accelerations of every step's first evaluation and the
    kinetic energy; from each window's starting state the oracle takes the
    same 50 steps in a thread of its own.  After every step: on a window's
    first step (identical input) accelerations within 1e-12 of each
    particle's force scale; on the later ones, where the two trajectories
    separate by rounding, within 1e-10 of the largest force scale of the
    system; the kinetic energy within 1e-12 and 1e-10 relative; at a window's
    end the contact sets are identical."""
    wl, g0 = poly200k
    _windows_against_oracle(wl, g0, [50 * w for w in range(6)], 50)


def test_config1_windows_across_5000_steps(hip, poly200k):
    """The configuration's 5 000 steps from the Euler start-up of the loose
    packing, sampled: eleven windows of 50 steps, one every 500 steps and the
    last 50 (the settling from the first steps to the end), each checked against the
    oracle from the GPU's state at its start with the bars of the 300-step
    test.  The first window starts after the first MoveParticles step: the
    state of the handle right after the start-up holds particles that the
    last Euler sub-step moved into another cell in their old cell list until
    the next evaluation, which a state upload (or a state file the handle
    writes at that point) does not represent -- from such a checkpoint the
    GPU and the oracle both step alike, and both differ from the run that
    goes on in one pair of this packing."""
    wl, _ = poly200k
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    _windows_against_oracle(wl, g, [1] + [500 * w for w in range(1, 10)] + [4950], 50)
    ob = g.observables()
    print(f"after {g.stats()['steps']} steps: phi {ob['packing_fraction']:.4f}, "
          f"{ob['pair_contacts']} pairs", flush=True)
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
        # torques against sum r |F_T| of the particle's contacts (the oracle's
        # torque scale), and for information against r times its force scale
        ok_b, rb = accel_tolerance_ok(sm["alpha"], so["alpha"], ts + 1e-300, ii, 1e-5)
        _, rb_f = accel_tolerance_ok(sm["alpha"], so["alpha"], rad * fs, ii, 1e-5)
        print(f"state {k}: force error {ra:.3g}, torque error {rb:.3g} of the bar "
              f"({rb_f:.3g} against r times the force scale)", flush=True)
        assert ok_a, f"state {k}: acceleration error {ra:.3g} x tolerance"
        assert ok_b, f"state {k}: angular acceleration error {rb:.3g} x tolerance"
        assert m.stats()["pair_contacts"] == o.stats()["pair_contacts"], k
        g0.step(1)
    m.close()


def test_config2_disordered_960k_against_oracle(hip):
    wl = workloads.packed_bed(hip, n=960_000, settle_steps=200)
    g0 = System(hip, wl.red)
    g0.upload(wl.comp, wl.x, wl.v, wl.w)
    g0.startup()
    g0.step(10)
    st = full_state(g0, with_rank=True)
    con = g0.download_contacts()
    g0.close()
    print(f"packed 960k: {len(con['keys'])} stored pairs, {len(con['wkeys'])} wall contacts",
          flush=True)
    assert len(con["keys"]) > wl.np
    g, o = System(hip, wl.red), oracle_system(wl.red)
    upload_same([g, o], wl, state=st, contacts=con)
    for s in (g, o):
        s.compute_accelerations(0, True)
    sg, so = g.download(), o.download()
    fs, ts = force_scale(o)
    im, ii = _inv(wl.comp, wl.red)
    ok_a, ra = accel_tolerance_ok(sg["a"], so["a"], fs, im, 1e-12)
    ok_b, rb = accel_tolerance_ok(sg["alpha"], so["alpha"], ts + 1e-300, ii, 1e-12)
    print(f"960k evaluation: acceleration error {ra:.3g}, angular {rb:.3g} of the bar", flush=True)
    assert ok_a and ok_b, (ra, rb)
    assert np.array_equal(g.download_cells(), o.download_cells())
    compare_contacts(g.download_contacts(), o.download_contacts(), exact=True)
    assert g.stats()["ncol"] == o.stats()["ncol"]
    # three ABM steps from the same state
    upload_same([g, o], wl, state=st, contacts=con)
    ncol0 = [s.stats()["ncol"] for s in (g, o)]
    for s in (g, o):
        s.step(3)
    sg, so = g.download(), o.download()
    for k, rel in (("x", 1e-12), ("v", 1e-10), ("w", 1e-10)):
        assert _close(sg[k], so[k], rel), k
    assert np.array_equal(g.download_cells(), o.download_cells())
    cg, co = g.download_contacts(), o.download_contacts()
    assert np.array_equal(cg["keys"], co["keys"])
    assert np.array_equal(cg["wkeys"], co["wkeys"])
    for k in ("ft", "vt"):
        assert _close(cg[k], co[k], 1e-9), k
    assert g.stats()["ncol"] - ncol0[0] == o.stats()["ncol"] - ncol0[1]
    g.close()
    o.close()
    # the launch shapes that stage each tile and its halo in LDS (with all
    # halo rows, and with 8, so that most partners overflow to memory) give
    # the global-gather shape's steps bit for bit on this bed
    runs = []
    for shape, rows in ((7, 512), (5, 512), (4, 512), (4, 8)):
        gs = System(hip, wl.red)
        upload_same([gs], wl, state=st, contacts=con)
        gs.set_option("pass_shape", shape)
        gs.set_option("halo_rows", rows)
        gs.step(3)
        d, c, stt = gs.download(), gs.download_contacts(), gs.stats()
        runs.append((d, c, stt))
        gs.close()
        print(f"shape {shape}, {rows} halo rows: largest halo {stt['max_halo']}, "
              f"{stt['halo_overflow']} tile listings over the rows", flush=True)
    for d, c, stt in runs[1:]:
        for k in ("x", "v", "w", "a", "alpha"):
            assert np.array_equal(d[k], runs[0][0][k]), k
        for k in ("keys", "ft", "vt", "wkeys", "wft", "wvt"):
            assert np.array_equal(c[k], runs[0][1][k]), k
    assert runs[1][2]["max_halo"] > 0 and runs[1][2]["halo_overflow"] == 0
    assert runs[3][2]["halo_overflow"] > 0


def test_config2_packed_4m_pair_count_against_kdtree(hip):
    from obs_ref import overlapping_pairs
    wl = workloads.packed_bed(hip, n=4_000_000, settle_steps=200)
    g = System(hip, wl.red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    g.step(20)
    g.compute_accelerations(0, False)
    p_gpu = g.stats()["pair_contacts"]
    p_obs = g.observables()["pair_contacts"]
    st = g.download()
    g.close()
    print(f"packed 4M: {p_gpu} pairs on the GPU", flush=True)
    p_tree = overlapping_pairs(wl.red, st["x"], st["comp"])
    assert p_gpu == p_tree == p_obs, (p_gpu, p_tree, p_obs)
    assert p_gpu > wl.np // 2
