"""ctypes view of the C-ABI in include/dem3d.h and a thin host-side object.

`System` mirrors the module interface the Fortran driver uses
(src/MForces.f90:145 ComputeAccelerations, src/MDEM.f90:258-369
MoveParticlesEuler/MoveParticles) over any library that exports the ABI under
a given symbol prefix.  The product library is libdem3d_hip.so (prefix
"dem_"); `load_hip()` refuses to run without it -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .spec import Reduced

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.environ.get("DEM3D_HIP_LIB") or os.path.join(HERE, "libdem3d_hip.so")

DEM_OK = 0
DEM_ERR_ESCAPE = -6
DEM_FP64, DEM_MIXED = 0, 1
DEM_HIST_REFERENCE, DEM_HIST_CANONICAL = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)
_lp = C.POINTER(C.c_int64)


class Params(C.Structure):
    _fields_ = [
        ("nc", C.c_int32), ("nw", C.c_int32), ("pbc", C.c_int32 * 3),
        ("method_order", C.c_int32), ("n_euler", C.c_int32), ("max_ncell", C.c_int32),
        ("ref_grid", C.c_int32 * 3), ("precision", C.c_int32), ("history_mode", C.c_int32),
        ("max_neighbours", C.c_int32), ("max_contacts", C.c_int32),
        ("lbox", C.c_double * 3), ("gravity", C.c_double * 3), ("time_step", C.c_double),
        ("skin", C.c_double),
        ("radius", _dp), ("mass", _dp), ("inertia", _dp), ("inv_mass", _dp),
        ("inv_inertia", _dp), ("eff_mass", _dp), ("min_dij", _dp), ("min_dij_sq", _dp),
        ("kn", _dp), ("yn", _dp), ("kt", _dp), ("yt", _dp), ("musq", _dp),
        ("knw", _dp), ("ynw", _dp), ("ktw", _dp), ("ytw", _dp), ("muwsq", _dp),
    ]


class WallT(C.Structure):
    _fields_ = [
        ("shape", C.c_int32), ("order", C.c_int32), ("l1", C.c_double), ("l2", C.c_double),
        ("l3", C.c_double), ("stiffness", C.c_double), ("shear_stiffness", C.c_double),
        ("u", C.c_double * 3), ("w", C.c_double * 3), ("n", C.c_double * 3),
        ("position", C.c_double * 3), ("velocity", C.c_double * 3), ("omega", C.c_double * 3),
    ]


class WallMotionT(C.Structure):
    _fields_ = [("constant_factor", C.c_double * 3), ("amplitude", C.c_double * 3),
                ("velocity_amplitude", C.c_double * 3), ("two_pi_frequency", C.c_double * 3),
                ("phase_angle", C.c_double * 3)]


class ParticleT(C.Structure):
    _fields_ = [("index", C.c_int32), ("pad_", C.c_int32), ("position", C.c_double * 3),
                ("velocity", C.c_double * 3), ("acceleration", C.c_double * 3),
                ("omega", C.c_double * 3), ("alpha", C.c_double * 3)]


assert C.sizeof(ParticleT) == 128


NSHAPES = 8


def autotune_steps(steps_per_shape):
    """Steps that dem_autotune / dem_batch_autotune advance a system by."""
    return NSHAPES * (steps_per_shape + 1) + 2 * steps_per_shape + 1


class TuneT(C.Structure):
    """dem_tune_t: the autotuner's report."""
    _fields_ = [("nshapes", C.c_int32), ("chosen", C.c_int32), ("steps", C.c_int64),
                ("block", C.c_int32 * NSHAPES), ("waves", C.c_int32 * NSHAPES),
                ("front", C.c_int32 * NSHAPES), ("ms", C.c_double * NSHAPES),
                ("block_map", C.c_int32), ("pad_", C.c_int32), ("ms_map", C.c_double * 2)]

    def asdict(self):
        n = self.nshapes
        return dict(chosen=self.chosen, steps=self.steps, block=list(self.block[:n]),
                    waves=list(self.waves[:n]), front=list(self.front[:n]),
                    ms=list(self.ms[:n]), block_map=self.block_map,
                    ms_map=list(self.ms_map))


class Stats(C.Structure):
    _fields_ = [
        ("np", C.c_int64), ("pair_contacts", C.c_int64), ("wall_contacts", C.c_int64),
        ("stored_pairs", C.c_int64), ("ncol", C.c_int64), ("evaluations", C.c_int64),
        ("steps", C.c_int64), ("rebuilds", C.c_int64), ("neighbour_pairs", C.c_int64),
        ("stale_dropped", C.c_int64), ("ref_ncell", C.c_int64 * 3), ("nb_ncell", C.c_int64 * 3),
        ("kinetic_energy", C.c_double), ("max_overlap", C.c_double), ("ms_pass1", C.c_double),
        ("ms_pass2", C.c_double), ("ms_rebuild", C.c_double), ("ms_step", C.c_double),
        ("mean_pair_contacts", C.c_double), ("time", C.c_double), ("next_step", C.c_int64),
        ("max_halo", C.c_int64), ("halo_overflow", C.c_int64), ("device_bytes", C.c_int64),
    ]

    def asdict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class ObservablesT(C.Structure):
    """dem_observables_t."""
    _fields_ = [("kinetic_energy", C.c_double), ("potential_energy", C.c_double),
                ("pair_contacts", C.c_int64), ("wall_contacts", C.c_int64),
                ("coordination", C.c_double), ("packing_fraction", C.c_double),
                ("region", C.c_double * 2), ("solid_volume", C.c_double), ("top", C.c_double)]

    def asdict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class FluidT(C.Structure):
    _fields_ = [("fluid_density", C.c_double), ("viscosity", C.c_double),
                ("flow_rate", C.c_double), ("nl", C.c_int32), ("pad_", C.c_int32),
                ("density", _dp)]


class RngT(C.Structure):
    _fields_ = [("aam", C.c_double), ("dummy", C.c_int32), ("iix", C.c_int32),
                ("iiy", C.c_int32), ("iik", C.c_int32)]

    @classmethod
    def seeded(cls, seed: int):
        """The generator as a run starts it: Dummy = the specification's seed
        (negative), iix = iiy = -1 (src/MUtil.f90:10-11)."""
        return cls(0.0, int(seed), -1, -1, 0)


class PourT(C.Structure):
    _fields_ = [("target_number", C.POINTER(C.c_int32)), ("shape", C.c_int32),
                ("max_attempts", C.c_int32), ("size", C.c_double * 2),
                ("length_scale", C.c_double), ("time_scale", C.c_double)]


class ConfigT(C.Structure):
    _fields_ = [
        ("form", C.c_int32), ("version", C.c_int32), ("length_scale", C.c_double),
        ("time_scale", C.c_double), ("mass_scale", C.c_double), ("time", C.c_double),
        ("lbox", C.c_double * 3), ("nc", C.c_int32), ("nw", C.c_int32), ("np", C.c_int32),
        ("number", _ip), ("radius", _dp), ("density", _dp), ("stiffness", _dp),
        ("shear_stiffness", _dp), ("mass", _dp), ("inertia", _dp), ("walls", C.POINTER(WallT)),
        ("comp", _ip), ("x", _dp), ("v", _dp), ("w", _dp),
    ]


class RunT(C.Structure):
    _fields_ = [
        ("finished", C.c_int32), ("start_time", C.c_char * 28), ("n_interruptions", C.c_int32),
        ("interruption", C.c_void_p), ("retaking", C.c_void_p), ("finish_time", C.c_char * 28),
        ("hours", C.c_int32), ("minutes", C.c_int32), ("seconds", C.c_int32),
        ("milisecs", C.c_int32), ("aam", C.c_double), ("dummy", C.c_int32), ("iix", C.c_int32),
        ("iiy", C.c_int32), ("iik", C.c_int32), ("conf", C.c_int32), ("next_step", C.c_int32),
        ("ncol", C.c_int32),
    ]


class SectionsT(C.Structure):
    _fields_ = [
        ("nprev", C.c_int32), ("ipc", C.c_int32 * 3), ("ring", _dp), ("ncell", C.c_int32),
        ("cell_len", _ip), ("cell_members", _ip), ("npairs", C.c_int64), ("keys", _up),
        ("ft", _dp), ("vt", _dp), ("nwall", C.c_int64), ("wkeys", _up), ("wft", _dp),
        ("wvt", _dp),
    ]


def _ptr(a: Optional[np.ndarray], typ=_dp):
    if a is None:
        return C.cast(None, typ)
    return a.ctypes.data_as(typ)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class DemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Library:
    """Binds the ABI functions of one shared library under a symbol prefix."""

    FUNCS = {
        "create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
        "destroy": (None, [C.c_void_p]),
        "last_error": (C.c_char_p, [C.c_void_p]),
        "set_walls": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(WallT)]),
        "upload_state": (C.c_int, [C.c_void_p, C.c_int32, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip]),
        "download_state": (C.c_int, [C.c_void_p, C.c_int32, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip]),
        "download_predicted": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp, _dp]),
        "upload_contacts": (C.c_int, [C.c_void_p, C.c_int64, _up, _dp, _dp, C.c_int64, _up, _dp, _dp]),
        "download_contacts": (C.c_int, [C.c_void_p, C.c_int64, _lp, _up, _dp, _dp, C.c_int64, _lp,
                                        _up, _dp, _dp]),
        "download_cells": (C.c_int, [C.c_void_p, C.c_int32, _ip]),
        "compute_accelerations_aos": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ParticleT),
                                                C.c_double, C.c_int32]),
        "compute_accelerations": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
        "euler_startup": (C.c_int, [C.c_void_p]),
        "step": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, _lp, _ip]),
        "stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    }
    OPTIONAL = {
        "observables": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ObservablesT), _dp]),
        "potential_energy": (C.c_double, [C.c_void_p]),
        "observables_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ObservablesT), _dp]),
        "set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
        "set_wall_motion": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(WallMotionT)]),
        "download_walls": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(WallT)]),
        "synchronize": (C.c_int, [C.c_void_p]),
        "download_cell_rank": (C.c_int, [C.c_void_p, C.c_int32, _ip]),
        "ranf": (C.c_double, [C.POINTER(RngT)]),
        "set_fluid": (C.c_int, [C.c_void_p, C.POINTER(FluidT)]),
        "download_fluid_field": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp]),
        "pour": (C.c_int, [C.c_void_p, C.POINTER(PourT), C.POINTER(RngT), _ip]),
        "remove_escaped": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip]),
        "download_force_scale": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
        "config_write": (C.c_int, [C.c_char_p, C.POINTER(ConfigT)]),
        "config_read": (C.c_int, [C.c_char_p, C.POINTER(ConfigT)]),
        "state_file_write": (C.c_int, [C.c_char_p, C.POINTER(RunT), C.POINTER(ConfigT),
                                       C.POINTER(SectionsT)]),
        "state_file_read": (C.c_int, [C.c_char_p, C.POINTER(RunT), C.POINTER(ConfigT),
                                      C.POINTER(SectionsT)]),
        "save_state_file": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(RunT),
                                      C.POINTER(ConfigT)]),
        "load_state_file": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(RunT),
                                      C.POINTER(ConfigT)]),
        "batch_create": (C.c_int, [C.POINTER(C.c_void_p)]),
        "batch_destroy": (None, [C.c_void_p]),
        "batch_last_error": (C.c_char_p, [C.c_void_p]),
        "batch_add": (C.c_int, [C.c_void_p, C.c_void_p]),
        "batch_remove": (C.c_int, [C.c_void_p, C.c_void_p]),
        "batch_size": (C.c_int, [C.c_void_p]),
        "batch_step": (C.c_int, [C.c_void_p, C.c_int64, _lp, _ip]),
        "batch_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
        "batch_step_times": (C.c_int64, [C.c_void_p, C.c_int64, _dp]),
        "autotune": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(TuneT)]),
        "batch_autotune": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(TuneT)]),
        "footprint": (C.c_int, [C.POINTER(Params), C.c_int32, _lp]),
    }

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built")
        self.path = path
        self.lib = C.CDLL(path)
        self.prefix = prefix
        for name, (res, args) in list(self.FUNCS.items()) + list(self.OPTIONAL.items()):
            try:
                f = getattr(self.lib, prefix + name)
            except AttributeError:
                if name in self.OPTIONAL:
                    continue
                raise
            f.restype = res
            f.argtypes = args
            setattr(self, name, f)

    def symbols(self):
        return [self.prefix + n for n in self.FUNCS]


_hip: Optional[Library] = None


def load_hip() -> Library:
    """The product library.  Raises if it is missing: there is no fallback."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(
                f"{HIP_LIB} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        _hip = Library(HIP_LIB, "dem_")
    return _hip


def footprint(lib: "Library", red: Reduced, n: int, **kw) -> int:
    """dem_footprint: the device bytes a system of n particles with the
    reduced constants red (and System's keyword options) allocates."""
    out = C.c_int64(0)
    rc = lib.footprint(C.byref(make_params(red, **kw)), int(n), C.byref(out))
    if rc != DEM_OK:
        raise DemError(rc, lib.last_error(None).decode())
    return int(out.value)


def make_params(red: Reduced, precision=DEM_FP64, history_mode=DEM_HIST_REFERENCE,
                skin: float = 0.0, max_neighbours=0, max_contacts=0, max_ncell=None):
    """dem_params_t from reduced constants; the numpy buffers are kept alive on
    the returned object."""
    p = Params()
    keep = {}

    def arr(name, a, shape):
        a = np.asfortranarray(np.asarray(a, np.float64).reshape(shape)).ravel(order="F")
        a = np.ascontiguousarray(a)
        keep[name] = a
        setattr(p, name, a.ctypes.data_as(_dp))

    nc, nw = red.nc, red.nw
    p.nc, p.nw = nc, nw
    for d in range(3):
        p.pbc[d] = int(red.pbc[d])
        p.lbox[d] = float(red.lbox[d])
        p.gravity[d] = float(red.gravity[d])
    p.method_order = red.method_order
    p.n_euler = red.n_euler
    p.max_ncell = int(red.max_ncell if max_ncell is None else max_ncell)
    if red.ref_grid is not None:
        for d in range(3):
            p.ref_grid[d] = int(red.ref_grid[d])
    p.precision = precision
    p.history_mode = history_mode
    p.max_neighbours = max_neighbours
    p.max_contacts = max_contacts
    p.time_step = float(red.time_step)
    p.skin = float(skin)
    for name in ("radius", "mass", "inertia", "inv_mass", "inv_inertia", "eff_mass"):
        arr(name, getattr(red, name), (nc,))
    for name in ("min_dij", "min_dij_sq", "kn", "yn", "kt", "yt", "musq"):
        arr(name, getattr(red, name), (nc, nc))
    for name in ("knw", "ynw", "ktw", "ytw", "muwsq"):
        a = np.asarray(getattr(red, name), np.float64).reshape(nc, nw)
        if nw == 0:
            a = np.zeros((nc, 1))
        arr(name, a, a.shape)
    p._keep = keep
    return p


def make_walls(red: Reduced):
    arr = (WallT * max(red.nw, 1))()
    for j, w in enumerate(red.walls):
        t = arr[j]
        t.shape = int(w["shape"])
        t.order = 0
        t.l1, t.l2, t.l3 = float(w["l1"]), float(w["l2"]), float(w["l3"])
        t.stiffness = float(w["stiffness"])
        t.shear_stiffness = float(w["shear_stiffness"])
        for d in range(3):
            t.u[d] = float(w["u"][d])
            t.w[d] = float(w["w"][d])
            t.n[d] = float(w["n"][d])
            t.position[d] = float(w["position"][d])
            t.velocity[d] = float(w["velocity"][d])
            t.omega[d] = float(w["omega"][d])
    return arr


def make_wall_motion(motion):
    """dem_wall_motion_t[nw] from a list of dicts with the TWallMovement
    fields (constant_factor, amplitude, velocity_amplitude, two_pi_frequency,
    phase_angle), reduced units."""
    arr = (WallMotionT * max(len(motion), 1))()
    for j, m in enumerate(motion):
        for name, _ in WallMotionT._fields_:
            for d in range(3):
                getattr(arr[j], name)[d] = float(m[name][d])
    return arr


class System:
    """One granular system resident in a library (GPU product or a checker)."""

    def __init__(self, lib: Library, red: Reduced, **kw):
        self.lib = lib
        self.red = red
        self.params = make_params(red, **kw)
        h = C.c_void_p()
        rc = lib.create(C.byref(self.params), C.byref(h))
        if rc != DEM_OK:
            raise DemError(rc, lib.last_error(None).decode())
        self.h = h
        self.np = 0
        self.nprev = red.method_order - 1
        self.next_step = red.method_order  # InitialStep after the start-up (src/DEM.f90:87-92)
        if red.nw:
            self.walls = make_walls(red)
            self._check(lib.set_walls(self.h, red.nw, self.walls))
            if getattr(red, "wall_motion", None):
                self.set_wall_motion(red.wall_motion)
        if getattr(red, "fluid", None):
            self.set_fluid(red.fluid)

    def set_fluid(self, fluid):
        """FluidFlow with dict(density, viscosity, flow_rate, nl) (reduced);
        None turns it off."""
        if fluid is None:
            self._check(self.lib.set_fluid(self.h, None))
            return
        self._dens = np.ascontiguousarray(self.red.density, np.float64)
        f = FluidT(float(fluid["density"]), float(fluid["viscosity"]), float(fluid["flow_rate"]),
                   int(fluid["nl"]), 0, self._dens.ctypes.data_as(_dp))
        self._check(self.lib.set_fluid(self.h, C.byref(f)))

    def download_fluid_field(self):
        nl = int(self.red.fluid["nl"])
        vf, ef = np.zeros(nl + 1), np.zeros(nl + 1)
        self._check(self.lib.download_fluid_field(self.h, nl, _ptr(vf), _ptr(ef)))
        return vf, ef

    def set_wall_motion(self, motion):
        """WallsInMovement with WM = motion (list of dicts, one per wall);
        an empty list stops the walls."""
        self._wm = make_wall_motion(motion)
        self._check(self.lib.set_wall_motion(self.h, len(motion), self._wm))

    def download_walls(self):
        nw = self.red.nw
        arr = (WallT * max(nw, 1))()
        self._check(self.lib.download_walls(self.h, nw, arr))
        pos = np.array([[arr[j].position[d] for d in range(3)] for j in range(nw)]).reshape(nw, 3)
        vel = np.array([[arr[j].velocity[d] for d in range(3)] for j in range(nw)]).reshape(nw, 3)
        return dict(position=pos, velocity=vel)

    def set_time(self, t):
        self._check(self.lib.set_option(self.h, b"time", float(t)))

    def set_option(self, name: str, value: float):
        """dem_set_option ("timing", "resort_interval", "pass_shape", "block_map", ...)."""
        self._check(self.lib.set_option(self.h, name.encode(), float(value)))

    def autotune(self, steps_per_shape=20):
        """dem_autotune: times every launch shape of the force passes while
        stepping (the steps count, the trajectory is unchanged) and keeps the
        fastest.  Returns the report as a dict."""
        t = TuneT()
        rc = self.lib.autotune(self.h, int(steps_per_shape), int(self.next_step), C.byref(t))
        self.next_step += int(t.steps)
        self._check(rc)
        return t.asdict()

    def _check(self, rc):
        if rc != DEM_OK:
            raise DemError(rc, self.lib.last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state
    def upload(self, comp, x, v, w, a=None, alpha=None, ring=None, ipc=None, cell_rank=None):
        comp = np.ascontiguousarray(comp, np.int32)
        n = comp.shape[0]
        self.np = n
        conv = lambda q: None if q is None else _f64(np.asarray(q).reshape(3, n))
        x, v, w, a, alpha = map(conv, (x, v, w, a, alpha))
        ring = None if ring is None else _f64(ring)
        ipc = None if ipc is None else np.ascontiguousarray(ipc, np.int32)
        cr = None if cell_rank is None else np.ascontiguousarray(cell_rank, np.int32)
        self._keep = (comp, x, v, w, a, alpha, ring, ipc, cr)
        self._check(self.lib.upload_state(self.h, n, _ptr(comp, _ip), _ptr(x), _ptr(v), _ptr(w),
                                          _ptr(a), _ptr(alpha), _ptr(ring), _ptr(ipc, _ip),
                                          _ptr(cr, _ip)))

    def download(self):
        n = self.np
        comp = np.zeros(n, np.int32)
        x, v, w, a, al = (np.zeros((3, n)) for _ in range(5))
        ring = np.zeros((max(self.nprev, 1), 9, n))
        ipc = np.zeros(3, np.int32)
        self._check(self.lib.download_state(self.h, n, _ptr(comp, _ip), _ptr(x), _ptr(v), _ptr(w),
                                            _ptr(a), _ptr(al), _ptr(ring), _ptr(ipc, _ip)))
        return dict(comp=comp, x=x, v=v, w=w, a=a, alpha=al, ring=ring[: self.nprev],
                    ipc=ipc[: self.nprev])

    def download_predicted(self):
        n = self.np
        x, v, w, a, al = (np.zeros((3, n)) for _ in range(5))
        self._check(self.lib.download_predicted(self.h, n, _ptr(x), _ptr(v), _ptr(w), _ptr(a),
                                                _ptr(al)))
        return dict(x=x, v=v, w=w, a=a, alpha=al)

    def upload_contacts(self, keys, ft, vt, wkeys=None, wft=None, wvt=None):
        keys = np.ascontiguousarray(keys, np.uint64)
        ft, vt = _f64(np.asarray(ft).reshape(-1, 3)), _f64(np.asarray(vt).reshape(-1, 3))
        if wkeys is None:
            wkeys, wft, wvt = np.zeros(0, np.uint64), np.zeros((0, 3)), np.zeros((0, 3))
        wkeys = np.ascontiguousarray(wkeys, np.uint64)
        wft, wvt = _f64(np.asarray(wft).reshape(-1, 3)), _f64(np.asarray(wvt).reshape(-1, 3))
        self._check(self.lib.upload_contacts(self.h, keys.shape[0], _ptr(keys, _up), _ptr(ft),
                                             _ptr(vt), wkeys.shape[0], _ptr(wkeys, _up),
                                             _ptr(wft), _ptr(wvt)))

    def download_contacts(self):
        n = C.c_int64()
        m = C.c_int64()
        self._check(self.lib.download_contacts(self.h, 0, C.byref(n), None, None, None, 0,
                                               C.byref(m), None, None, None))
        keys = np.zeros(n.value, np.uint64)
        ft, vt = np.zeros((n.value, 3)), np.zeros((n.value, 3))
        wkeys = np.zeros(m.value, np.uint64)
        wft, wvt = np.zeros((m.value, 3)), np.zeros((m.value, 3))
        self._check(self.lib.download_contacts(self.h, n.value, C.byref(n), _ptr(keys, _up),
                                               _ptr(ft), _ptr(vt), m.value, C.byref(m),
                                               _ptr(wkeys, _up), _ptr(wft), _ptr(wvt)))
        return dict(keys=keys, ft=ft, vt=vt, wkeys=wkeys, wft=wft, wvt=wvt)

    def download_cells(self):
        c = np.zeros(self.np, np.int32)
        self._check(self.lib.download_cells(self.h, self.np, _ptr(c, _ip)))
        return c

    # ---- hot path
    def compute_accelerations_aos(self, comp, x, v, w, time_step=None, update=True):
        n = self.np
        arr = (ParticleT * n)()
        buf = np.frombuffer(arr, dtype=np.uint8).reshape(n, 128)
        rec = np.zeros((n, 16), np.float64)
        rec[:, 1:4] = np.asarray(x).reshape(3, n).T
        rec[:, 4:7] = np.asarray(v).reshape(3, n).T
        rec[:, 10:13] = np.asarray(w).reshape(3, n).T
        buf[:] = rec.view(np.uint8)
        np.frombuffer(arr, dtype=np.int32).reshape(n, 32)[:, 0] = np.asarray(comp, np.int32)
        dt = self.red.time_step if time_step is None else time_step
        self._check(self.lib.compute_accelerations_aos(self.h, n, arr, float(dt), int(update)))
        out = np.frombuffer(arr, dtype=np.float64).reshape(n, 16)
        return out[:, 7:10].T.copy(), out[:, 13:16].T.copy()

    def compute_accelerations(self, which=0, update=True):
        self._check(self.lib.compute_accelerations(self.h, int(which), int(update)))

    def euler_startup(self):
        self._check(self.lib.euler_startup(self.h))

    def startup(self):
        """The driver's start-up loop (src/DEM.f90:87-92)."""
        for _ in range(self.red.method_order - 1):
            self.euler_startup()

    def step(self, nsteps, first_step=None, allow_escape=False):
        done = C.c_int64()
        esc = C.c_int32()
        first = self.next_step if first_step is None else first_step
        rc = self.lib.step(self.h, int(nsteps), int(first), C.byref(done), C.byref(esc))
        if rc in (DEM_OK, DEM_ERR_ESCAPE):
            self.next_step = int(first) + done.value
        if rc == DEM_ERR_ESCAPE and allow_escape:
            return done.value, esc.value
        self._check(rc)
        return done.value, esc.value

    def pour(self, target, shape, size, max_attempts, rng: "RngT", length_scale, time_scale):
        """One PourParticle; returns 1 when a particle was inserted (rng advances)."""
        tn = np.ascontiguousarray(target, np.int32)
        p = PourT(tn.ctypes.data_as(_ip), int(shape), int(max_attempts),
                  (C.c_double * 2)(*(list(size) + [0.0] * (2 - len(size)))),
                  float(length_scale), float(time_scale))
        ins = C.c_int32()
        self._check(self.lib.pour(self.h, C.byref(p), C.byref(rng), C.byref(ins)))
        self.np += ins.value
        return ins.value

    def remove_escaped(self):
        """UpdateParticleList; returns the indices each removal freed."""
        cap = max(self.np, 1)
        out = np.zeros(cap, np.int32)
        n = C.c_int32()
        self._check(self.lib.remove_escaped(self.h, cap, out.ctypes.data_as(_ip), C.byref(n)))
        self.np -= n.value
        return out[: n.value].copy()

    def observables(self, nlayers: int = 0, noalloc: bool = False):
        """dem_observables (or dem_observables_noalloc): energies, contact
        counts, coordination number, bulk packing fraction and, with
        nlayers > 0, the vertical profile as an array [nlayers, 4] (solid
        fraction, mean vx, vy, vz)."""
        o = ObservablesT()
        prof = np.zeros((max(nlayers, 0), 4))
        f = self.lib.observables_noalloc if noalloc else self.lib.observables
        self._check(f(self.h, int(nlayers), C.byref(o), _ptr(prof) if nlayers > 0 else None))
        d = o.asdict()
        if nlayers > 0:
            d["profile"] = prof
        return d

    def synchronize(self):
        if hasattr(self.lib, "synchronize"):
            self._check(self.lib.synchronize(self.h))

    def stats(self):
        s = Stats()
        self._check(self.lib.stats(self.h, C.byref(s)))
        return s.asdict()


class Batch:
    """Independent systems stepped together (dem_batch_*): every phase of a
    step is one kernel launch for all of them."""

    def __init__(self, lib: Library):
        if not hasattr(lib, "batch_create"):
            raise RuntimeError(f"{lib.path} has no batch entry points")
        self.lib = lib
        b = C.c_void_p()
        rc = lib.batch_create(C.byref(b))
        if rc != DEM_OK:
            raise DemError(rc, lib.batch_last_error(None).decode())
        self.b = b
        self.systems = []

    def _check(self, rc):
        if rc != DEM_OK:
            raise DemError(rc, self.lib.batch_last_error(self.b).decode())
        return rc

    def add(self, sysm: "System"):
        self._check(self.lib.batch_add(self.b, sysm.h))
        self.systems.append(sysm)

    def remove(self, sysm: "System"):
        self._check(self.lib.batch_remove(self.b, sysm.h))
        self.systems.remove(sysm)

    def set_timing(self, on=True):
        self._check(self.lib.batch_set_option(self.b, b"timing", 1.0 if on else 0.0))

    def set_option(self, name: str, value: float):
        """dem_batch_set_option: "timing", "graph", "graph_steps", "pass_shape",
        "block_map", "persistent", "persist_steps", "spin_limit"."""
        self._check(self.lib.batch_set_option(self.b, name.encode(), float(value)))

    def autotune(self, steps_per_shape=20):
        """dem_batch_autotune (see System.autotune); every system advances by
        the report's "steps"."""
        t = TuneT()
        rc = self.lib.batch_autotune(self.b, int(steps_per_shape), C.byref(t))
        for sysm in self.systems:
            sysm.next_step += int(t.steps)
        self._check(rc)
        return t.asdict()

    def step(self, nsteps, allow_escape=False):
        n = len(self.systems)
        done = np.zeros(max(n, 1), np.int64)
        esc = np.zeros(max(n, 1), np.int32)
        rc = self.lib.batch_step(self.b, int(nsteps), done.ctypes.data_as(_lp),
                                 esc.ctypes.data_as(_ip))
        if rc in (DEM_OK, DEM_ERR_ESCAPE):
            for sysm, d in zip(self.systems, done[:n]):
                sysm.next_step += int(d)
        if not (rc == DEM_ERR_ESCAPE and allow_escape):
            self._check(rc)
        return done[:n], esc[:n]

    def step_times(self):
        k = self.lib.batch_step_times(self.b, 0, None)
        out = np.zeros(max(k, 1))
        self.lib.batch_step_times(self.b, k, out.ctypes.data_as(_dp))
        return out[:k]

    def close(self):
        if getattr(self, "b", None):
            self.lib.batch_destroy(self.b)
            self.b = None
            self.systems = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
