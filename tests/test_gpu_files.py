"""State files on the GPU path: dem_load_state_file resumes from a file the
reference wrote, dem_save_state_file writes what the oracle (and, through
the format tests on the CPU, the reference) resumes from, and the
reference's cell-list order survives on the device (dem_download_cell_rank
equals the oracle's)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from dem3d_amd import files, workloads
from dem3d_amd.native import System
from checkers import oracle_system
from test_oracle import reduced_from_fixture

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STT = os.path.join(GOLDEN, "ref_resume.stt")


def _fixture():
    f = np.load(os.path.join(GOLDEN, "ref_resume.npz"), allow_pickle=False)
    return f, reduced_from_fixture(f)


def _close(a, b, rel):
    scale = np.abs(b).max() + 1e-300
    return np.abs(a - b).max() <= rel * scale


def test_resume_from_reference_state_file(hip):
    f, red = _fixture()
    g = System(hip, red)
    run, meta = files.load_state(g, STT)
    assert run["next_step"] == int(f["next_step"]) and run["conf"] == 3
    assert g.stats()["next_step"] == int(f["next_step"])
    g.step(int(f["resume_steps"]))
    st = g.download()
    for k in ("x", "v", "w"):
        assert _close(st[k], f[f"r_{k}"], 1e-9), k
    assert np.array_equal(g.download_contacts()["keys"], f["r_keys"])
    assert g.stats()["next_step"] == int(f["next_step"]) + int(f["resume_steps"])


def test_save_after_load_reproduces_the_reference_file(hip, tmp_path):
    f, red = _fixture()
    g = System(hip, red)
    run, meta = files.load_state(g, STT)
    out = str(tmp_path / "again.stt")
    files.save_state(g, out, meta, run)
    assert open(out, "rb").read() == open(STT, "rb").read()


def test_gpu_state_file_resumed_by_oracle(hip, tmp_path):
    wl = workloads.small_box(n_side=(6, 6, 5), nc=2, seed=71, vel=0.4)
    red = wl.red
    g = System(hip, red)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    g.step(30)
    # the device keeps the reference's cell-list order
    o = oracle_system(red)
    o.upload(wl.comp, wl.x, wl.v, wl.w)
    o.startup()
    o.step(30)
    import ctypes as C
    rg, ro = np.zeros(g.np, np.int32), np.zeros(o.np, np.int32)
    ip = C.POINTER(C.c_int32)
    assert hip.download_cell_rank(g.h, g.np, rg.ctypes.data_as(ip)) == 0
    assert o.lib.download_cell_rank(o.h, o.np, ro.ctypes.data_as(ip)) == 0
    assert np.array_equal(g.download_cells(), o.download_cells())
    assert np.array_equal(rg, ro)
    # save on the GPU, resume in the oracle from the file
    meta = dict(length_scale=red.length_scale, time_scale=red.time_scale,
                mass_scale=red.mass_scale, density=red.density,
                stiffness=np.ones(red.nc), shear_stiffness=np.ones(red.nc))
    path = str(tmp_path / "gpu.stt")
    files.save_state(g, path, meta, dict(conf=5))
    r, cfg, sec = files.read_state_file(hip, path, red.method_order - 1)
    assert r["next_step"] == red.method_order + 30 and r["conf"] == 5
    assert r["ncol"] == g.stats()["ncol"]
    st = g.download()
    for k in ("x", "v", "w"):
        assert np.array_equal(cfg[k], st[k]), k
    assert np.array_equal(sec["ring"], st["ring"])
    assert files.cell_rank_from_lists(sec["cells"], g.np).tolist() == rg.tolist()
    o2 = oracle_system(red)
    o2.upload(cfg["comp"], cfg["x"], cfg["v"], cfg["w"], None, None, sec["ring"], sec["ipc"],
              files.cell_rank_from_lists(sec["cells"], g.np))
    o2.upload_contacts(sec["keys"], sec["ft"], sec["vt"], sec["wkeys"], sec["wft"], sec["wvt"])
    o2.step(15, r["next_step"])
    g.step(15)
    sg, so = g.download(), o2.download()
    for k in ("x", "v", "w"):
        assert _close(sg[k], so[k], 1e-9), k
    assert np.array_equal(g.download_contacts()["keys"], o2.download_contacts()["keys"])
