"""Continuous batching of simulation requests on one GPU.

Requests (workloads.Request: arrival time, size, steps, contact parameters)
arrive on their own clock.  The server admits a waiting request as soon as
its device footprint -- the library's own count of what the system will
allocate (dem_footprint) -- fits the budget next to the systems already
resident.  A worker thread builds and starts each admitted request -- the
packing on the host, then upload and Euler start-up on the system's own
stream -- while the main thread keeps stepping the batch, so the device
runs both at once; a built system joins the batch between two chunks.  The
batch advances all resident systems together in chunks of at most `chunk`
steps (bounded by the fewest remaining steps, so a finished system leaves at
once and its memory is freed for the next admission).  A system in which
particles escape stops there and leaves with the steps it did.  Every
system's state evolves exactly as it would alone (the batched kernels run
the same bodies with the same arguments), which tests/test_gpu_serve.py and
tests/test_gpu_ensemble_serving.py check.

Reported: per-step device time of the batch (p50/p99), per-request waiting
time (arrival to admission) and completion latency (arrival to the end of
its last step), and the aggregate particle-steps/s over the run.
"""
from __future__ import annotations

import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

from . import workloads
from .native import DEM_MIXED, Batch, Library, System, footprint

def request_bytes(lib: Library, req: workloads.Request, precision: int = DEM_MIXED) -> int:
    """The device bytes the request's system will hold: the library's own
    count of what it allocates for those constants and that many particles
    (dem_footprint), computed without building the particles."""
    red = workloads.ensemble_member_reduced(req.n, req.mu, req.e, req.seed, steps=req.steps)
    return footprint(lib, red, req.n, precision=precision)


@dataclass
class Record:
    request: workloads.Request
    index: int
    admitted: float = float("nan")
    started: float = float("nan")
    finished: float = float("nan")
    remaining: int = 0
    steps_done: int = 0
    bytes: int = -1  # device footprint (request_bytes), once computed
    escaped: int = 0
    system: Optional[System] = None
    result: Optional[dict] = None
    build: Optional[Future] = None


@dataclass
class Report:
    records: List[Record]
    step_ms: np.ndarray
    wall_s: float
    particle_steps: float
    chunks: int
    max_resident: int
    peak_bytes: int

    def summary(self) -> dict:
        lat = np.array([r.finished - r.request.arrival for r in self.records])
        wait = np.array([r.admitted - r.request.arrival for r in self.records])
        st = self.step_ms if self.step_ms.size else np.zeros(1)
        return {
            "requests": len(self.records),
            "particle_steps": self.particle_steps,
            "wall_s": self.wall_s,
            "particle_steps_per_s": self.particle_steps / self.wall_s if self.wall_s else 0.0,
            "step_ms_p50": float(np.percentile(st, 50)),
            "step_ms_p99": float(np.percentile(st, 99)),
            "completion_s_p50": float(np.percentile(lat, 50)),
            "completion_s_p99": float(np.percentile(lat, 99)),
            "wait_s_p50": float(np.percentile(wait, 50)),
            "wait_s_p99": float(np.percentile(wait, 99)),
            "batch_steps": int(st.size),
            "chunks": self.chunks,
            "max_resident": self.max_resident,
            "peak_device_bytes": self.peak_bytes,
        }


def build_system(lib: Library, req: workloads.Request, precision: int) -> System:
    wl = workloads.ensemble_member(req.n, req.mu, req.e, req.seed, steps=req.steps)
    s = System(lib, wl.red, precision=precision)
    s.upload(wl.comp, wl.x, wl.v, wl.w)
    s.startup()
    return s


def serve(lib: Library, trace: List[workloads.Request], precision: int = DEM_MIXED,
          budget_bytes: int = 200 << 30, chunk: int = 100, realtime: bool = True,
          keep: Optional[Callable[[Record], dict]] = None, log=None) -> Report:
    """Runs the trace to completion.  realtime=False admits requests as if
    each had already arrived when the previous chunk ended (a closed loop,
    for tests).  keep(record) may extract results before a finished system
    is destroyed."""
    recs = [Record(q, k) for k, q in enumerate(sorted(trace, key=lambda q: q.arrival))]
    pending = list(recs)
    building: List[Record] = []
    active: List[Record] = []
    batch = Batch(lib)
    batch.set_timing(True)
    builder = ThreadPoolExecutor(max_workers=1)
    step_ms: List[np.ndarray] = []
    used = 0
    peak = 0
    max_res = 0
    chunks = 0
    psteps = 0.0
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0

    def retire(r: Record, t: float):
        nonlocal used
        r.finished = t
        if keep is not None:
            r.result = keep(r)
        batch.remove(r.system)
        r.system.close()
        r.system = None
        used -= r.bytes

    try:
        while pending or building or active:
            # admit everything that has arrived and fits: its build starts on
            # the worker thread
            while pending and (not realtime or pending[0].request.arrival <= now()):
                if pending[0].bytes < 0:
                    pending[0].bytes = request_bytes(lib, pending[0].request, precision)
                need = pending[0].bytes
                if (active or building) and used + need > budget_bytes:
                    break
                r = pending.pop(0)
                r.admitted = now()
                r.remaining = r.request.steps
                r.build = builder.submit(build_system, lib, r.request, precision)
                building.append(r)
                used += need
                peak = max(peak, used)
            # built systems join the batch between chunks; with nothing to
            # step, wait for the next build to finish
            if not active and building:
                building[0].build.result()
            for r in [q for q in building if q.build.done()]:
                r.system = r.build.result()
                r.build = None
                batch.add(r.system)
                r.started = now()
                building.remove(r)
                active.append(r)
            if not active:
                if pending and not building:  # idle until the next arrival
                    time.sleep(max(0.0, min(0.05, pending[0].request.arrival - now())))
                continue
            max_res = max(max_res, len(active))
            k = min(chunk, min(r.remaining for r in active))
            members = list(batch.systems)
            done, esc = batch.step(k, allow_escape=True)
            chunks += 1
            step_ms.append(batch.step_times())
            by_sys = {id(sy): (int(d), int(e)) for sy, d, e in zip(members, done, esc)}
            t = now()
            still = []
            for r in active:
                d, e = by_sys[id(r.system)]
                psteps += float(r.request.n) * d
                r.remaining -= d
                r.steps_done += d
                if e > 0:
                    r.escaped = e
                if r.remaining <= 0 or e > 0:
                    retire(r, t)
                else:
                    still.append(r)
            active = still
            if log is not None:
                log(f"t={t:.2f}s chunk={k} resident={len(active)} building={len(building)} "
                    f"pending={len(pending)}")
    finally:
        builder.shutdown(wait=True)
        for r in building:
            if r.build is not None and r.build.done() and r.build.exception() is None:
                r.build.result().close()
        batch.close()
    wall = now()
    sm = np.concatenate(step_ms) if step_ms else np.zeros(0)
    return Report(recs, sm, wall, psteps, chunks, max_res, peak)
