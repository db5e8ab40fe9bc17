"""Configuration (.cfg) and state (.stt) files in the reference's binary
layout, through the library's dem_config_* / dem_state_file_* entry points
(host code; no device needed).

A configuration is a dict with the fields of SaveConfiguration
(src/MConfig.f90:109-236): form, version, length_scale, time_scale,
mass_scale, time, lbox, per component number/radius/density/stiffness/
shear_stiffness/mass/inertia, walls (dicts of the dem_wall_t fields) and the
particles comp[np], x/v/w[3, np].  A state file (SaveState,
src/DEM.f90:947-1014) adds the run records (a dict of the dem_run_t fields)
and the hot path's sections: ipc, ring[nprev, 9, np], the cell lists and the
contact lists (keys, ft, vt, wkeys, wft, wvt).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from .native import (ConfigT, DemError, Library, RunT, SectionsT, WallT, _dp, _ip, _up,
                     DEM_OK)

COMPLETE, COMPACT = 72, 28

_WALL_VEC = ("u", "w", "n", "position", "velocity", "omega")
_COMP = ("radius", "density", "stiffness", "shear_stiffness", "mass", "inertia")
_RUN_INT = ("finished", "n_interruptions", "hours", "minutes", "seconds", "milisecs", "dummy",
            "iix", "iiy", "iik", "conf", "next_step", "ncol")


def _check(lib: Library, rc, h=None):
    if rc != DEM_OK:
        raise DemError(rc, lib.last_error(h).decode(errors="replace"))


def _walls_to_c(walls):
    arr = (WallT * max(len(walls), 1))()
    for j, w in enumerate(walls):
        t = arr[j]
        t.shape = int(w["shape"])
        t.l1, t.l2, t.l3 = float(w["l1"]), float(w["l2"]), float(w["l3"])
        t.stiffness = float(w.get("stiffness", 0.0))
        t.shear_stiffness = float(w.get("shear_stiffness", 0.0))
        for k in _WALL_VEC:
            for d in range(3):
                getattr(t, k)[d] = float(w[k][d])
    return arr


def _walls_from_c(arr, nw):
    out = []
    for j in range(nw):
        t = arr[j]
        w = dict(shape=int(t.shape), l1=t.l1, l2=t.l2, l3=t.l3, stiffness=t.stiffness,
                 shear_stiffness=t.shear_stiffness)
        for k in _WALL_VEC:
            w[k] = np.array([getattr(t, k)[d] for d in range(3)])
        out.append(w)
    return out


class _Keep:
    """Holds the numpy buffers a ctypes struct points into."""

    def __init__(self):
        self.items = []

    def f64(self, a, shape=None):
        a = np.ascontiguousarray(a, np.float64)
        if shape is not None:
            a = a.reshape(shape)
        self.items.append(a)
        return a.ctypes.data_as(_dp)

    def i32(self, a):
        a = np.ascontiguousarray(a, np.int32)
        self.items.append(a)
        return a.ctypes.data_as(_ip)

    def u64(self, a):
        a = np.ascontiguousarray(a, np.uint64)
        self.items.append(a)
        return a.ctypes.data_as(_up)


def _config_to_c(cfg: dict, keep: _Keep) -> ConfigT:
    c = ConfigT()
    c.form = int(cfg.get("form", COMPLETE))
    c.version = int(cfg.get("version", 2))
    for k in ("length_scale", "time_scale", "mass_scale", "time"):
        setattr(c, k, float(cfg.get(k, 0.0)))
    for d in range(3):
        c.lbox[d] = float(cfg["lbox"][d])
    nc = len(cfg["radius"])
    comp = np.asarray(cfg.get("comp", np.zeros(0, np.int32)), np.int32)
    n = comp.shape[0]
    walls = cfg.get("walls", [])
    c.nc, c.nw, c.np = nc, len(walls), n
    number = cfg.get("number")
    if number is None:
        number = [int(np.count_nonzero(comp == k + 1)) for k in range(nc)]
    c.number = keep.i32(number)
    for k in _COMP:
        v = cfg.get(k)
        setattr(c, k, keep.f64(v) if v is not None else C.cast(None, _dp))
    warr = _walls_to_c(walls)
    keep.items.append(warr)
    c.walls = C.cast(warr, C.POINTER(WallT))
    c.comp = keep.i32(comp)
    for k in ("x", "v", "w"):
        v = cfg.get(k)
        setattr(c, k, keep.f64(v, (3, n)) if v is not None else C.cast(None, _dp))
    return c


def _config_arrays(c: ConfigT, keep: _Keep) -> dict:
    """Allocate the arrays of a configuration whose counts c holds."""
    nc, nw, n = c.nc, c.nw, c.np
    out = {"number": np.zeros(nc, np.int32), "comp": np.zeros(n, np.int32)}
    for k in _COMP:
        out[k] = np.zeros(nc)
    for k in ("x", "v", "w"):
        out[k] = np.zeros((3, n))
    c.number = out["number"].ctypes.data_as(_ip)
    c.comp = out["comp"].ctypes.data_as(_ip)
    for k in _COMP + ("x", "v", "w"):
        setattr(c, k, out[k].ctypes.data_as(_dp))
    warr = (WallT * max(nw, 1))()
    keep.items.append(warr)
    c.walls = C.cast(warr, C.POINTER(WallT))
    out["_walls"] = warr
    return out


def _config_dict(c: ConfigT, arrays: dict) -> dict:
    out = dict(form=c.form, version=c.version, length_scale=c.length_scale,
               time_scale=c.time_scale, mass_scale=c.mass_scale, time=c.time,
               lbox=np.array([c.lbox[d] for d in range(3)]))
    for k in ("number", "comp", "x", "v", "w") + _COMP:
        out[k] = arrays[k]
    out["walls"] = _walls_from_c(arrays["_walls"], c.nw)
    return out


def write_config(lib: Library, path: str, cfg: dict):
    keep = _Keep()
    c = _config_to_c(cfg, keep)
    _check(lib, lib.config_write(path.encode(), C.byref(c)))


def read_config(lib: Library, path: str) -> dict:
    c = ConfigT()
    _check(lib, lib.config_read(path.encode(), C.byref(c)))
    keep = _Keep()
    arrays = _config_arrays(c, keep)
    _check(lib, lib.config_read(path.encode(), C.byref(c)))
    return _config_dict(c, arrays)


def default_run() -> dict:
    run = {k: 0 for k in _RUN_INT}
    run.update(start_time="", finish_time="", interruption=[], retaking=[], aam=0.0)
    return run


def _run_to_c(run: dict, keep: _Keep) -> RunT:
    r = RunT()
    full = default_run()
    full.update(run or {})
    for k in _RUN_INT:
        setattr(r, k, int(full[k]))
    r.aam = float(full["aam"])
    r.start_time = full["start_time"].encode()[:28]
    r.finish_time = full["finish_time"].encode()[:28]
    n = len(full["interruption"])
    r.n_interruptions = n
    if n:
        pack = lambda lst: C.create_string_buffer(b"".join(s.encode()[:28].ljust(28) for s in lst),
                                                  28 * n)
        a, b = pack(full["interruption"]), pack(full["retaking"])
        keep.items += [a, b]
        r.interruption = C.cast(a, C.c_void_p)
        r.retaking = C.cast(b, C.c_void_p)
    return r


def _run_dict(r: RunT, inter=None, ret=None) -> dict:
    out = {k: int(getattr(r, k)) for k in _RUN_INT}
    out["aam"] = float(r.aam)
    out["start_time"] = bytes(r.start_time).decode(errors="replace")
    out["finish_time"] = bytes(r.finish_time).decode(errors="replace")
    n = r.n_interruptions
    cut = lambda buf: [bytes(buf.raw[28 * k: 28 * k + 28]).decode(errors="replace")
                       for k in range(n)] if buf is not None else []
    out["interruption"], out["retaking"] = cut(inter), cut(ret)
    return out


def _sections_to_c(sec: dict, n: int, keep: _Keep) -> SectionsT:
    s = SectionsT()
    nprev = int(sec["nprev"])
    s.nprev = nprev
    for k in range(nprev):
        s.ipc[k] = int(sec["ipc"][k])
    s.ring = keep.f64(sec["ring"], (max(nprev, 1) * 9 * n,)) if nprev else C.cast(None, _dp)
    cells = sec["cells"]
    s.ncell = len(cells)
    s.cell_len = keep.i32([len(c) for c in cells])
    s.cell_members = keep.i32([q for c in cells for q in c] or [0])
    s.npairs = len(sec["keys"])
    s.keys = keep.u64(sec["keys"])
    s.ft = keep.f64(np.asarray(sec["ft"]).reshape(-1))
    s.vt = keep.f64(np.asarray(sec["vt"]).reshape(-1))
    s.nwall = len(sec["wkeys"])
    s.wkeys = keep.u64(sec["wkeys"])
    s.wft = keep.f64(np.asarray(sec["wft"]).reshape(-1))
    s.wvt = keep.f64(np.asarray(sec["wvt"]).reshape(-1))
    return s


def write_state_file(lib: Library, path: str, run: Optional[dict], cfg: dict, sec: dict):
    keep = _Keep()
    r = _run_to_c(run, keep)
    c = _config_to_c(dict(cfg, form=COMPLETE), keep)
    s = _sections_to_c(sec, c.np, keep)
    _check(lib, lib.state_file_write(path.encode(), C.byref(r), C.byref(c), C.byref(s)))


def read_state_file(lib: Library, path: str, nprev: int):
    """(run, configuration, sections) of a state file; nprev = MethodOrder - 1."""
    r, c, s = RunT(), ConfigT(), SectionsT()
    s.nprev = nprev
    _check(lib, lib.state_file_read(path.encode(), C.byref(r), C.byref(c), C.byref(s)))
    keep = _Keep()
    arrays = _config_arrays(c, keep)
    n = c.np
    ninter = r.n_interruptions
    inter = C.create_string_buffer(28 * max(ninter, 1))
    ret = C.create_string_buffer(28 * max(ninter, 1))
    r.interruption = C.cast(inter, C.c_void_p)
    r.retaking = C.cast(ret, C.c_void_p)
    ring = np.zeros((max(nprev, 1), 9, n))
    cell_len = np.zeros(s.ncell, np.int32)
    members = np.zeros(max(n, 1), np.int32)
    keys, wkeys = np.zeros(s.npairs, np.uint64), np.zeros(s.nwall, np.uint64)
    ft, vt = np.zeros((s.npairs, 3)), np.zeros((s.npairs, 3))
    wft, wvt = np.zeros((s.nwall, 3)), np.zeros((s.nwall, 3))
    s.ring = ring.ctypes.data_as(_dp)
    s.cell_len = cell_len.ctypes.data_as(_ip)
    s.cell_members = members.ctypes.data_as(_ip)
    s.keys, s.wkeys = keys.ctypes.data_as(_up), wkeys.ctypes.data_as(_up)
    s.ft, s.vt = ft.ctypes.data_as(_dp), vt.ctypes.data_as(_dp)
    s.wft, s.wvt = wft.ctypes.data_as(_dp), wvt.ctypes.data_as(_dp)
    _check(lib, lib.state_file_read(path.encode(), C.byref(r), C.byref(c), C.byref(s)))
    cells, m = [], 0
    for L in cell_len:
        cells.append([int(q) for q in members[m:m + L]])
        m += L
    sec = dict(nprev=nprev, ipc=np.array([s.ipc[k] for k in range(nprev)], np.int32),
               ring=ring[:nprev], cells=cells, keys=keys, ft=ft, vt=vt, wkeys=wkeys, wft=wft,
               wvt=wvt)
    return _run_dict(r, inter, ret), _config_dict(c, arrays), sec


def cell_rank_from_lists(cells, n):
    """Rank of every particle in its cell list (the head of a list of m has
    rank m-1), the form dem_upload_state takes."""
    rank = np.zeros(n, np.int32)
    for c in cells:
        m = len(c)
        for k, q in enumerate(c):
            rank[q - 1] = m - 1 - k
    return rank


def save_state(system, path: str, meta: dict, run: Optional[dict] = None):
    """dem_save_state_file on a resident system; meta holds length_scale,
    time_scale, mass_scale and per component density, stiffness and
    shear_stiffness."""
    lib = system.lib
    keep = _Keep()
    m = ConfigT()
    for k in ("length_scale", "time_scale", "mass_scale"):
        setattr(m, k, float(meta[k]))
    for k in ("density", "stiffness", "shear_stiffness"):
        setattr(m, k, keep.f64(meta[k]))
    r = _run_to_c(run, keep)
    _check(lib, lib.save_state_file(system.h, path.encode(), C.byref(r), C.byref(m)), system.h)


def load_state(system, path: str):
    """dem_load_state_file into a system; returns the run records and the
    configuration's scalars and component table."""
    lib = system.lib
    r, m = RunT(), ConfigT()
    cap = 1024  # room for the interruption records
    inter, ret = C.create_string_buffer(28 * cap), C.create_string_buffer(28 * cap)
    r.n_interruptions = cap
    r.interruption = C.cast(inter, C.c_void_p)
    r.retaking = C.cast(ret, C.c_void_p)
    nc = system.red.nc
    comp = {k: np.zeros(nc) for k in _COMP}
    number = np.zeros(nc, np.int32)
    m.number = number.ctypes.data_as(_ip)
    for k in _COMP:
        setattr(m, k, comp[k].ctypes.data_as(_dp))
    _check(lib, lib.load_state_file(system.h, path.encode(), C.byref(r), C.byref(m)), system.h)
    system.np = m.np
    system.next_step = r.next_step
    meta = dict(length_scale=m.length_scale, time_scale=m.time_scale, mass_scale=m.mass_scale,
                time=m.time, number=number, **comp)
    got = r.n_interruptions <= cap
    return _run_dict(r, inter if got else None, ret if got else None), meta
