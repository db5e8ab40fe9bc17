"""Test-side access to the oracles: the C restatement (oracle/liboracle.so)
and the reference Fortran modules (oracle/_ref/libdem3d_ref*.so).

Only tests, smoke() and bench.py's cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import multiprocessing as mp
import os
import struct
import shutil
import tempfile
from typing import Optional

import numpy as np

from dem3d_amd.native import Library, System

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libdem3d_ref.so")
REF_LIB_BIG = os.path.join(ROOT, "oracle", "_ref", "libdem3d_ref_big.so")
# capacity-raised and built for speed (-O3, AVX2/FMA): timing only
REF_LIB_FAST = os.path.join(ROOT, "oracle", "_ref", "libdem3d_ref_fast.so")
# the same with -march=native for this host (built on request): timing only
REF_LIB_NATIVE = os.path.join(ROOT, "oracle", "_ref", "libdem3d_ref_native.so")

_oracle = None


def oracle_lib() -> Library:
    global _oracle
    if _oracle is None:
        _oracle = Library(ORACLE_LIB, "orc_")
    return _oracle


def oracle_system(red, **kw) -> System:
    return System(oracle_lib(), red, **kw)


def _ref_path(big=False, fast=False, native=False) -> str:
    if native:
        return REF_LIB_NATIVE
    return REF_LIB_FAST if fast else (REF_LIB_BIG if big else REF_LIB)


def ref_available(big=False, fast=False, native=False) -> bool:
    return os.path.exists(_ref_path(big, fast, native))


# ---------------------------------------------------------------------------
# The reference's binary state layout (BinFiles byte stream, src/BinFiles.f90):
# SaveCellStructure (src/MCells.f90:958-983), SaveIntegrationVariables
# (src/MDEM.f90:52-68), SaveContactList (src/MContacts.f90:75-118).

def parse_state_file(path, np_, nprev):
    data = open(path, "rb").read()
    off = 0

    def i32():
        nonlocal off
        v = struct.unpack_from("<i", data, off)[0]
        off += 4
        return v

    def f64(n):
        nonlocal off
        v = np.frombuffer(data, "<f8", n, off).copy()
        off += 8 * n
        return v

    nt = i32()
    cells = []
    for _ in range(nt):
        k = i32()
        cells.append([i32() for _ in range(k)])
    ipc = [i32() for _ in range(nprev)]
    ring = np.zeros((nprev, 9, np_))
    for j in range(nprev):
        for i in range(np_):
            ring[j, :, i] = f64(9)
    npl = i32()
    pairs, walls = [], []
    for i in range(npl):
        for m in range(2):
            k = i32()
            for _ in range(k):
                idx = i32()
                v1 = f64(3)
                v2 = f64(3)
                (pairs if m == 0 else walls).append((i + 1, idx, v1, v2))
    return dict(cells=cells, ipc=np.array(ipc, np.int32), ring=ring, pairs=pairs, walls=walls)


def write_state_file(path, cells, ipc, ring, pairs, walls, np_):
    """Inverse of parse_state_file; pairs/walls are (i, j, v1, v2) 1-based."""
    out = bytearray()
    out += struct.pack("<i", len(cells))
    for c in cells:
        out += struct.pack("<i", len(c))
        for q in c:
            out += struct.pack("<i", q)
    for q in ipc:
        out += struct.pack("<i", int(q))
    for j in range(len(ipc)):
        for i in range(np_):
            out += np.asarray(ring[j, :, i], "<f8").tobytes()
    out += struct.pack("<i", np_)
    by_i = {}
    for (i, j, v1, v2) in pairs:
        by_i.setdefault((i, 0), []).append((j, v1, v2))
    for (i, j, v1, v2) in walls:
        by_i.setdefault((i, 1), []).append((j, v1, v2))
    for i in range(1, np_ + 1):
        for m in range(2):
            lst = sorted(by_i.get((i, m), []), key=lambda t: t[0])
            out += struct.pack("<i", len(lst))
            for (j, v1, v2) in lst:
                out += struct.pack("<i", int(j))
                out += np.asarray(v1, "<f8").tobytes() + np.asarray(v2, "<f8").tobytes()
    pad = (-len(out)) % 1024
    out += b"\0" * pad
    with open(path, "wb") as f:
        f.write(bytes(out))


def contacts_to_keys(pairs):
    keys = np.array([(i << 32) | j for (i, j, _, _) in pairs], np.uint64)
    ft = np.array([v1 for (_, _, v1, _) in pairs]).reshape(-1, 3)
    vt = np.array([v2 for (_, _, _, v2) in pairs]).reshape(-1, 3)
    order = np.argsort(keys, kind="stable")
    return keys[order], ft[order], vt[order]


def cells_to_rank(cells, np_):
    rank = np.zeros(np_, np.int32)
    for c in cells:
        n = len(c)
        for k, q in enumerate(c):
            rank[q - 1] = n - 1 - k
    return rank


# ---------------------------------------------------------------------------
# Running the reference modules.  They keep global module state, so every job
# runs in a fresh process.

def _ref_job(job, q):
    # The Fortran modules keep array temporaries on the stack; beyond ~200k
    # particles they overflow a default 8 MiB main-thread stack, so the job
    # runs on a thread with a 1 GiB stack.
    import threading

    res = {}

    def body():
        try:
            res["v"] = ("ok", _ref_job_body(job))
        except BaseException:  # pragma: no cover - reported to the parent
            import traceback
            res["v"] = ("err", traceback.format_exc())

    threading.stack_size(1 << 30)
    th = threading.Thread(target=body)
    th.start()
    th.join()
    q.put(res.get("v", ("err", "reference job ended without a result")))


def _d(a):
    return np.asfortranarray(np.asarray(a, np.float64))


def _ref_job_body(job):
    red = job["red"]
    lib = C.CDLL(_ref_path(job.get("big", False), job.get("fast", False), job.get("native", False)))
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    P = lambda a: a.ctypes.data_as(dp)
    nc, nw = red.nc, red.nw
    keep = []

    def tab(a, shape):
        a = _d(np.asarray(a, np.float64).reshape(shape))
        keep.append(a)
        return P(a)

    fl = job.get("fluid_si")
    if fl is not None:
        # the specification's fluid block as ReadFluidSpecifications reads it
        import tempfile as _tf
        fd, fpath = _tf.mkstemp(suffix=".fluid", dir=os.path.join(ROOT, ".scratch")
                                if os.path.isdir(os.path.join(ROOT, ".scratch")) else None)
        with os.fdopen(fd, "w") as f:
            f.write("[Simulate fluid flow?]\nyes\n[Fluid viscosity]\n%.17g\n"
                    "[Flow rate]\n%.17g\n[Layers]\n%d\n" % (fl["viscosity"], fl["flow_rate"],
                                                            int(fl["layers"])))
        b = fpath.encode()
        lib.ref_prepare_fluid(C.c_double(red.fluid["density"]), C.c_char_p(b), C.c_int32(len(b)),
                              C.c_double(red.length_scale), C.c_double(red.time_scale),
                              C.c_double(red.mass_scale))
        os.remove(fpath)
    pbc = np.array(red.pbc, np.int32)
    lbox = _d(red.lbox)
    grav = _d(red.gravity)
    wshape = (nc, max(nw, 1))
    zw = np.zeros(wshape)
    lib.ref_init(C.c_int32(nc), C.c_int32(nw), pbc.ctypes.data_as(ip), P(lbox), P(grav),
                 C.c_double(red.time_step), C.c_int32(red.method_order), C.c_int32(red.n_euler),
                 tab(red.radius, nc), tab(red.mass, nc), tab(red.inertia, nc), tab(red.density, nc),
                 tab(red.kn, (nc, nc)), tab(red.yn, (nc, nc)), tab(red.kt, (nc, nc)),
                 tab(red.yt, (nc, nc)), tab(red.mu, (nc, nc)),
                 tab(red.knw if nw else zw, wshape), tab(red.ynw if nw else zw, wshape),
                 tab(red.ktw if nw else zw, wshape), tab(red.ytw if nw else zw, wshape),
                 tab(red.muw if nw else zw, wshape))
    for j, w in enumerate(red.walls):
        vec = lambda k: tab(w[k], 3)
        lib.ref_set_wall(C.c_int32(j + 1), C.c_int32(int(w["shape"])), C.c_double(w["l1"]),
                         C.c_double(w["l2"]), C.c_double(w["l3"]), vec("u"), vec("w"), vec("n"),
                         vec("position"), vec("velocity"), vec("omega"))
    lib.ref_configure_walls()
    wm = getattr(red, "wall_motion", None)
    if wm:
        fields = ("constant_factor", "amplitude", "velocity_amplitude", "two_pi_frequency",
                  "phase_angle")
        cols_wm = [tab(np.array([m[f] for m in wm], np.float64).T, (3, nw)) for f in fields]
        lib.ref_set_wall_motion(C.c_int32(nw), *cols_wm)
    lib.ref_get_time.restype = C.c_double
    out = {}
    md, mdsq, musq = (np.zeros((nc, nc), order="F") for _ in range(3))
    invm, invi, effm = (np.zeros(nc) for _ in range(3))
    muwsq = np.zeros(wshape, order="F")
    lib.ref_get_tables(P(md), P(mdsq), P(invm), P(invi), P(effm), P(musq), P(muwsq))
    out["tables"] = dict(min_dij=md, min_dij_sq=mdsq, inv_mass=invm, inv_inertia=invi,
                         eff_mass=effm, musq=musq, muwsq=muwsq[:, :nw])
    grid = np.zeros(3, np.int32)
    lib.ref_get_grid(grid.ctypes.data_as(ip))
    out["grid"] = grid
    scratch = os.path.join(ROOT, ".scratch")
    os.makedirs(scratch, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="ref_", dir=scratch)
    nprev = red.method_order - 1
    if job.get("state_file") is not None:
        # ReadState: the configuration, cell lists, ring and contacts from a file
        path = job["state_file"].encode()
        ints = np.zeros(12, np.int32)
        aam = C.c_double()
        st_, ft_ = C.create_string_buffer(28), C.create_string_buffer(28)
        lib.ref_read_state_file(C.c_char_p(path), C.c_int32(len(path)), ints.ctypes.data_as(ip),
                                C.byref(aam), st_, ft_)
        lib.ref_get_np.restype = C.c_int32
        n = int(lib.ref_get_np())
        out["read_run"] = dict(ints=ints.copy(), aam=aam.value, start_time=st_.raw,
                               finish_time=ft_.raw)
        comp = np.zeros(n, np.int32)
        cols = lambda a: _d(np.asarray(a).reshape(3, n).T)
    else:
        n = int(job["comp"].shape[0])
        comp = np.ascontiguousarray(job["comp"], np.int32)
        cols = lambda a: _d(np.asarray(a).reshape(3, n).T)
        x, v, w = cols(job["x"]), cols(job["v"]), cols(job["w"])
        a = cols(job["a"]) if job.get("a") is not None else np.zeros((n, 3), order="F")
        al = cols(job["alpha"]) if job.get("alpha") is not None else np.zeros((n, 3), order="F")
        lib.ref_set_particles(C.c_int32(n), comp.ctypes.data_as(ip), P(x), P(v), P(w), P(a),
                              P(al))
    if job.get("meta") is not None:
        m = job["meta"]
        lib.ref_set_meta(C.c_double(m["length_scale"]), C.c_double(m["time_scale"]),
                         C.c_double(m["mass_scale"]), C.c_double(m.get("time", 0.0)),
                         tab(m["stiffness"], nc), tab(m["shear_stiffness"], nc),
                         tab(m["wall_stiffness"] if nw else [0.0], max(nw, 1)),
                         tab(m["wall_shear_stiffness"] if nw else [0.0], max(nw, 1)))
    if job.get("state_file") is not None:
        pass
    elif job.get("state") is not None:
        st = job["state"]
        path = os.path.join(tmp, "in.stt").encode()
        write_state_file(path, st["cells"], st["ipc"], st["ring"], st["pairs"], st["walls"], n)
        lib.ref_load_state(C.c_char_p(path), C.c_int32(len(path)))
    else:
        lib.ref_fresh_state()
    lib.ref_kinetic_energy.restype = C.c_double
    lib.ref_potential_energy.restype = C.c_double
    lib.ref_ncol.restype = C.c_int32

    lib.ref_get_np.restype = C.c_int32

    def snapshot(tag):
        n = int(lib.ref_get_np())
        c2 = np.zeros(n, np.int32)
        xs, vs, ws, as_, als = (np.zeros((n, 3), order="F") for _ in range(5))
        lib.ref_get_particles(C.c_int32(n), c2.ctypes.data_as(ip), P(xs), P(vs), P(ws), P(as_),
                              P(als))
        ci = np.zeros(n, np.int32)
        lib.ref_get_cell_index(C.c_int32(n), ci.ctypes.data_as(ip))
        path = os.path.join(tmp, f"{tag}.stt").encode()
        lib.ref_save_state(C.c_char_p(path), C.c_int32(len(path)))
        st = parse_state_file(path, n, nprev)
        os.remove(path)
        wpos, wvel = np.zeros((3, max(nw, 1)), order="F"), np.zeros((3, max(nw, 1)), order="F")
        if nw:
            lib.ref_get_walls(C.c_int32(nw), P(wpos), P(wvel))
        return dict(x=xs.T.copy(), v=vs.T.copy(), w=ws.T.copy(), a=as_.T.copy(),
                    alpha=als.T.copy(), cell_index=ci, state=st, ncol=int(lib.ref_ncol()),
                    ke=float(lib.ref_kinetic_energy()), pe=float(lib.ref_potential_energy()),
                    wall_pos=wpos[:, :nw].T.copy(),
                    wall_vel=wvel[:, :nw].T.copy(), time=float(lib.ref_get_time()))

    out["snapshots"] = []
    for act in job["actions"]:
        kind = act[0]
        if kind == "compute":
            lib.ref_compute_accelerations(C.c_int32(int(act[1])))
        elif kind == "compute_soa":
            _, upd, px, pv, pw, dt = act
            xs, vs, ws = cols(px), cols(pv), cols(pw)
            as_, als = np.zeros((n, 3), order="F"), np.zeros((n, 3), order="F")
            lib.ref_compute_accelerations_soa(C.c_int32(n), comp.ctypes.data_as(ip), P(xs),
                                              P(vs), P(ws), P(as_), P(als), C.c_double(dt),
                                              C.c_int32(int(upd)))
            out.setdefault("soa", []).append(dict(a=as_.T.copy(), alpha=als.T.copy()))
        elif kind == "euler":
            lib.ref_move_particles_euler()
        elif kind == "step":
            lib.ref_move_particles(C.c_int32(int(act[1])), C.c_int32(int(act[2])))
        elif kind == "snapshot":
            out["snapshots"].append(snapshot(act[1]))
        elif kind == "save_config":
            path = act[1].encode()
            lib.ref_save_configuration(C.c_char_p(path), C.c_int32(len(path)), C.c_int32(act[2]))
        elif kind == "read_config":
            path = act[1].encode()
            lib.ref_read_configuration(C.c_char_p(path), C.c_int32(len(path)))
            out.setdefault("configs", []).append(_ref_config(lib))
        elif kind == "save_state_file":
            _, fpath, run = act
            path = fpath.encode()
            pad = lambda s_: C.create_string_buffer(s_.encode()[:28].ljust(28), 28)
            ninter = len(run.get("interruption", []))
            inter = C.create_string_buffer(b"".join(q.encode()[:28].ljust(28)
                                                    for q in run.get("interruption", [])) or b" " * 28)
            ret = C.create_string_buffer(b"".join(q.encode()[:28].ljust(28)
                                                  for q in run.get("retaking", [])) or b" " * 28)
            i32 = lambda k: C.c_int32(int(run.get(k, 0)))
            lib.ref_save_state_file(C.c_char_p(path), C.c_int32(len(path)), i32("finished"),
                                    pad(run.get("start_time", "")), C.c_int32(ninter), inter, ret,
                                    pad(run.get("finish_time", "")), i32("hours"), i32("minutes"),
                                    i32("seconds"), i32("milisecs"),
                                    C.c_double(run.get("aam", 0.0)), i32("dummy"), i32("iix"),
                                    i32("iiy"), i32("iik"), i32("conf"), i32("next_step"))
        elif kind == "pour":
            _, pr = act
            rng = out.setdefault("rng", dict(aam=0.0, st=np.array([pr["seed"], -1, -1, 0],
                                                                  np.int32)))
            aam_ = C.c_double(rng["aam"])
            st = np.ascontiguousarray(rng["st"], np.int32)
            tn = np.ascontiguousarray(pr["target"], np.int32)
            size = np.ascontiguousarray(list(pr["size"]) + [0.0] * (2 - len(pr["size"])),
                                        np.float64)
            lib.ref_pour(tn.ctypes.data_as(ip), C.c_int32(pr["shape"]), P(size),
                         C.c_int32(pr["max_attempts"]), C.byref(aam_), st.ctypes.data_as(ip))
            rng["aam"], rng["st"] = aam_.value, st.copy()
            out.setdefault("np_after", []).append(int(lib.ref_get_np()))
        elif kind == "remove":
            lib.ref_update_particle_list()
            out.setdefault("np_after", []).append(int(lib.ref_get_np()))
        elif kind == "time":
            import time as _t
            t0 = _t.perf_counter()
            lib.ref_move_particles(C.c_int32(int(act[1])), C.c_int32(int(act[2])))
            out["seconds"] = _t.perf_counter() - t0
        else:
            raise ValueError(kind)
    shutil.rmtree(tmp, ignore_errors=True)
    return out


def _ref_config(lib):
    """The MConfig variables ReadConfiguration set, as a configuration dict."""
    s = np.zeros(7)
    counts = np.zeros(3, np.int32)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    lib.ref_get_config_scalars(s.ctypes.data_as(dp), counts.ctypes.data_as(ip))
    nc, nw, n = (int(q) for q in counts)
    number = np.zeros(nc, np.int32)
    props = np.zeros((nc, 6))
    lib.ref_get_components(C.c_int32(nc), number.ctypes.data_as(ip), props.ctypes.data_as(dp))
    walls = []
    for j in range(nw):
        shape = C.c_int32()
        v = np.zeros(23)
        lib.ref_get_wall(C.c_int32(j + 1), C.byref(shape), v.ctypes.data_as(dp))
        walls.append(dict(shape=shape.value, l1=v[0], l2=v[1], l3=v[2], stiffness=v[3],
                          shear_stiffness=v[4], u=v[5:8], w=v[8:11], n=v[11:14],
                          position=v[14:17], velocity=v[17:20], omega=v[20:23]))
    comp = np.zeros(n, np.int32)
    xs, vs, ws, as_, als = (np.zeros((n, 3), order="F") for _ in range(5))
    lib.ref_get_particles(C.c_int32(n), comp.ctypes.data_as(ip), xs.ctypes.data_as(dp),
                          vs.ctypes.data_as(dp), ws.ctypes.data_as(dp), as_.ctypes.data_as(dp),
                          als.ctypes.data_as(dp))
    out = dict(length_scale=s[0], time_scale=s[1], mass_scale=s[2], time=s[3], lbox=s[4:7],
               number=number, comp=comp, x=xs.T.copy(), v=vs.T.copy(), w=ws.T.copy(),
               walls=walls)
    for k, name in enumerate(("radius", "density", "stiffness", "shear_stiffness", "mass",
                              "inertia")):
        out[name] = props[:, k].copy()
    return out


def run_reference(job: dict, timeout: Optional[float] = 600, method: str = "fork"):
    """Run one job of the reference modules in a fresh process (fork by
    default: callers that have initialised the GPU pass method="spawn" from
    a __main__-guarded script)."""
    ctx = mp.get_context(method)
    q = ctx.Queue()
    p = ctx.Process(target=_ref_job, args=(job, q))
    p.start()
    import queue as _queue
    import time as _time
    t_end = None if timeout is None else _time.monotonic() + timeout
    try:
        while True:
            try:
                status, payload = q.get(timeout=1.0)
                break
            except _queue.Empty:
                if not p.is_alive():
                    raise RuntimeError(f"reference job died (exit code {p.exitcode})")
                if t_end is not None and _time.monotonic() > t_end:
                    raise TimeoutError(f"reference job exceeded {timeout} s")
    finally:
        p.join(timeout=30)
        if p.is_alive():
            p.kill()
    if status != "ok":
        raise RuntimeError(payload)
    return payload
