"""The C-ABI boundary without a GPU: the library exports every entry point
include/dem3d.h declares, the struct layouts agree between the C header, the
ctypes view (dem3d_amd/native.py) and the Fortran bindings
(dem3d_amd/fortran/dem3d_gpu.f90), and a Fortran program can call through
the bindings (dem_create's error path)."""
from __future__ import annotations

import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dem3d.h")
LIB = os.path.join(ROOT, "dem3d_amd", "libdem3d_hip.so")
FDIR = os.path.join(ROOT, "dem3d_amd", "fortran")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dem_[a-z_0-9]+)\s*\(", text)))


def c_sizes(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "dem3d.h"\nint main(void){\n'
                   + "".join(f'printf("sizeof {t} %zu\\n", sizeof({t}));\n'
                             for t in ("dem_params_t", "dem_wall_t", "dem_particle_t",
                                       "dem_stats_t", "dem_wall_motion_t", "dem_config_t",
                                       "dem_run_t", "dem_state_sections_t", "dem_rng_t",
                                       "dem_pour_t", "dem_fluid_t", "dem_tune_t", "dem_observables_t"))
                   + "return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {l.split()[1]: int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.skipif(not os.path.exists(LIB), reason="libdem3d_hip.so not built")
def test_library_exports_every_declared_entry_point():
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 17, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_ctypes(tmp_path):
    from dem3d_amd import native
    cs = c_sizes(tmp_path)
    assert cs["dem_particle_t"] == 128  # the reference's TParticle record
    assert C.sizeof(native.Params) == cs["dem_params_t"]
    assert C.sizeof(native.WallT) == cs["dem_wall_t"]
    assert C.sizeof(native.Stats) == cs["dem_stats_t"]
    assert C.sizeof(native.WallMotionT) == cs["dem_wall_motion_t"]
    assert C.sizeof(native.ConfigT) == cs["dem_config_t"]
    assert C.sizeof(native.RunT) == cs["dem_run_t"]
    assert C.sizeof(native.SectionsT) == cs["dem_state_sections_t"]
    assert C.sizeof(native.RngT) == cs["dem_rng_t"]
    assert C.sizeof(native.PourT) == cs["dem_pour_t"]
    assert C.sizeof(native.FluidT) == cs["dem_fluid_t"]
    assert C.sizeof(native.TuneT) == cs["dem_tune_t"]
    assert C.sizeof(native.ObservablesT) == cs["dem_observables_t"]


@pytest.mark.skipif(shutil.which("amdflang") is None or not os.path.exists(LIB),
                    reason="needs amdflang and libdem3d_hip.so")
def test_fortran_bindings_compile_link_and_match_layout(tmp_path):
    exe = tmp_path / "check_bindings"
    subprocess.run(["amdflang", "-O1", "-J", str(tmp_path),
                    os.path.join(FDIR, "dem3d_gpu.f90"), os.path.join(os.path.dirname(os.path.abspath(__file__)), "check_bindings.f90"),
                    "-o", str(exe), LIB, f"-Wl,-rpath,{os.path.dirname(LIB)}"],
                   check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    fs = {l.split()[1]: int(l.split()[2]) for l in out.splitlines() if l.startswith("sizeof")}
    assert fs == c_sizes(tmp_path)
    rc = int(re.search(r"dem_create (-?\d+)", out).group(1))
    # no device here: the error code and message come back through the bindings
    if rc != 0:
        assert rc == -2 and "device" in out.lower(), out


def test_integration_level1_excerpt_is_the_compiled_module():
    """The ComputeAccelerations shown in INTEGRATION.md is, line for line, the
    one in dem3d_amd/fortran/mforces_gpu.f90 (which oracle/build_ref.sh
    compiles with the reference's MDEM.f90 and DEM.f90)."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    src = open(os.path.join(ROOT, "dem3d_amd", "fortran", "mforces_gpu.f90")).read()
    m = re.search(r"```fortran\n(subroutine ComputeAccelerations.*?end subroutine)\n```", doc,
                  re.S)
    assert m, "level-1 excerpt not found"
    norm = lambda t: [re.sub(r"\s+", " ", ln.split("!")[0]).strip() for ln in t.splitlines()]
    excerpt = [ln for ln in norm(m.group(1)) if ln]
    module = [ln for ln in norm(src) if ln]
    # every line of the excerpt appears in the module, in the same order
    pos = 0
    for ln in excerpt:
        key = ln.replace("end subroutine", "end subroutine ComputeAccelerations")
        while pos < len(module) and module[pos] not in (ln, key):
            pos += 1
        assert pos < len(module), ln
        pos += 1


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"), reason="reference checkout absent")
def test_reference_programs_built():
    """oracle/build_ref.sh built the reference's program DEM and the same
    program over mforces_gpu.f90 (run by tests/test_gpu_reference_driver.py)."""
    for name in ("dem3d_ref", "dem3d_gpu"):
        p = os.path.join(ROOT, "oracle", "_ref", name)
        assert os.path.exists(p) and os.access(p, os.X_OK), p


@pytest.mark.skipif(not os.path.exists(LIB), reason="libdem3d_hip.so not built")
def test_footprint_needs_no_device():
    """dem_footprint answers on the host: positive, growing with the particle
    count by at least the two edge sets' history planes per particle, larger
    with more neighbour slots, and refusing bad arguments."""
    from dem3d_amd import native, workloads
    lib = native.Library(LIB, "dem_")
    red = workloads.ensemble_member_reduced(5000, 0.3, 0.8, 7, steps=100)
    b1 = native.footprint(lib, red, 5000)
    b2 = native.footprint(lib, red, 10000)
    assert b1 > 0 and b2 > b1
    # two edge sets of 48 slots per particle, 48 bytes of history each
    assert b2 - b1 >= 5000 * 2 * 48 * 48
    assert native.footprint(lib, red, 5000, max_neighbours=32) < b1
    with pytest.raises(native.DemError):
        native.footprint(lib, red, -1)
