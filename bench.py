"""Benchmark of the DEM per-timestep path on one MI355X.

Workload: the 4 000 000-sphere dense sheared bed -- bidisperse spheres
(d = 1.0 / 1.4 mm, half of each, species by a seeded shuffle) packed
disordered at volume fraction 0.60 by radius growth on the GPU
(dem3d_amd/packing.py, seed 20240003), periodic x,y, floor and a rough top
wall moving at 10 s^-1 * H; mu = 0.3, e = 0.8; FP64.  The bed is packed and
then sheared under gravity for 1 000 steps before the timing.  One step =
MoveParticles (src/MDEM.f90:285-369: two force evaluations, AB predictor, AM
corrector) + the escape scan, with the state resident in HBM.

Prints one JSON line:
  value        particle-steps/s over exactly --steps timed steps
  roofline     the dominant force pass: algorithmic bytes per launch / mean
               launch time from HIP events on the library's stream, against
               8 TB/s
  cpu_baseline the reference Fortran modules (oracle/_ref, capacity constants
               raised, built -O3 with AVX2/FMA) timed on one host core on a
               bounded sample: --cpu-steps steps of the same packed bed
  mixed        the same run in the mixed-precision mode (FP32 force
               arithmetic, FP64 state and spring history): rate, speed-up over
               FP64, and its kinetic energy and packing fraction against the
               FP64 run's after the same steps
  crystal      the ordered bed of the same spheres (rock-salt packing,
               phi = 0.567) as a second, labelled leg
  ensemble     256 independent systems (500..20 000 spheres each) stepped as
               one batch (one launch per phase, and persistent: one
               cooperative launch per 250 steps, a group of resident blocks
               per system) in mixed and in FP64, against the plain FP64 path
               that steps each system alone and against the reference
               stepping every system on the host's cores, one process per
               system, --cpu-workers at a time (--cpu-ens-steps steps each)
  serving      the Poisson request trace served with continuous batching:
               p50/p99 step time, waiting and completion latency, aggregate
               particle-steps/s
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of the fused two-pass design (FP64): per particle 436 B in
# each pass; per contact 112 B in pass 1 (history read + write) and 56 B in
# pass 2 (history read).  Neighbour-list and rebuild traffic excluded.
PASS1_B_PER_PARTICLE, PASS1_B_PER_CONTACT = 436.0, 112.0
PASS2_B_PER_PARTICLE, PASS2_B_PER_CONTACT = 436.0, 56.0


def cpu_baseline(wl, steps):
    """The reference on `steps` MoveParticles steps of the benchmark's own bed
    (zero ABM ring), in a spawned process (the GPU is initialised here)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from checkers import ref_available, run_reference
    if not ref_available(fast=True):
        return None
    res = run_reference(dict(red=wl.red, comp=wl.comp, x=wl.x, v=wl.v, w=wl.w, fast=True,
                             actions=[("time", 1, steps)]), timeout=900, method="spawn")
    sec = res["seconds"]
    return {
        "value": wl.np * steps / sec,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "reference",
        "host_cpus": os.cpu_count(),
        "cpu_model": cpu_model(),
        "sample": (f"reference Fortran MoveParticles (MaxNeighbours/MaxNCell raised; amdflang "
                   f"-O3 -march=x86-64-v3), {steps} steps of the benchmark's own packed bed "
                   f"({wl.np} spheres) from a zero ABM ring, {sec:.2f} s, 1 core, {cpu_model()}"),
    }


def cpu_ensemble_baseline(members, steps, workers):
    """The reference on every member of the ensemble, `steps` MoveParticles
    steps each from a zero ABM ring, `workers` members at a time, each in a
    spawned process of its own (one core each): the host's ensemble
    throughput mode.  The value is workers x particle-steps / the members'
    summed stepping time (set-up and process start-up excluded; the wall time
    including them is reported beside it)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from concurrent.futures import ThreadPoolExecutor
    from checkers import ref_available, run_reference
    if not ref_available(fast=True) or steps <= 0:
        return None

    def one(w):
        res = run_reference(dict(red=w.red, comp=w.comp, x=w.x, v=w.v, w=w.w, fast=True,
                                 actions=[("time", 1, steps)]), timeout=900, method="spawn")
        return res["seconds"]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        secs = list(ex.map(one, members))
    wall = time.perf_counter() - t0
    ps = sum(w.np for w in members) * steps
    return {
        "value": workers * ps / sum(secs),
        "unit": "particle-steps/s",
        "cores": workers,
        "kind": "reference",
        "host_cpus": os.cpu_count(),
        "cpu_model": cpu_model(),
        "wall_particle_steps_per_s": ps / wall,
        "sample": (f"reference Fortran MoveParticles (amdflang -O3 -march=x86-64-v3), all "
                   f"{len(members)} systems x {steps} steps from a zero ABM ring, {workers} "
                   f"systems at a time, one process and core each; stepping time summed over "
                   f"the systems {sum(secs):.1f} s, wall time with process start-up "
                   f"{wall:.1f} s, {cpu_model()}"),
    }


def cpu_configs(args):
    """The reference timed on this host's cores alone (no GPU needed when the
    200 k packing comes from --poly-cache): the 2 000-sphere settling case over
    its full 20 000 steps and the 200 000-sphere loose packing over
    --cpu-poly-steps steps, both from a zero ABM ring, one core each, with
    the -march=native build when it exists (oracle/build_ref.sh with
    DEM3D_REF_NATIVE=1), else the -O3 AVX2/FMA one.  One JSON line each."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from checkers import ref_available, run_reference
    from dem3d_amd import workloads
    native = ref_available(native=True)
    if not native and not ref_available(fast=True):
        print("bench.py: the reference timing build (oracle/_ref) is absent", file=sys.stderr)
        return 2
    build = ("amdflang -O3 -march=native" if native else "amdflang -O3 -march=x86-64-v3")

    def time_case(wl, steps):
        res = run_reference(dict(red=wl.red, comp=wl.comp, x=wl.x, v=wl.v, w=wl.w,
                                 fast=not native, native=native,
                                 actions=[("time", 1, steps)]), timeout=3600, method="spawn")
        sec = res["seconds"]
        print(json.dumps({"workload": wl.name, "spheres": wl.np, "steps": steps,
                          "seconds": sec, "value": wl.np * steps / sec,
                          "unit": "particle-steps/s", "cores": 1, "kind": "reference",
                          "build": build + " (MaxNeighbours/MaxNCell raised)",
                          "cpu": cpu_model()}), flush=True)

    time_case(workloads.mono_settling_2k(), 20000)
    state = None
    if args.poly_cache and os.path.exists(args.poly_cache):
        with np.load(args.poly_cache) as z:
            state = (z["x"], z["v"], z["w"])
    lib = None
    if state is None:
        from dem3d_amd.native import load_hip
        lib = load_hip()
    poly = workloads.poly_loose_packed(lib, state=state)
    if args.poly_cache and state is None:
        np.savez(args.poly_cache, x=poly.x, v=poly.v, w=poly.w)
    time_case(poly, args.cpu_poly_steps)
    return 0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def traffic_from_profiles(workload, kernel, shape, block_map):
    """HBM bytes per launch of kernel ("pass1" or "pass2") in the given launch
    shape and block order, measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE) for this workload over the benchmark's timed steps and
    committed in profiles/; None when not measured."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    for e in d.get("launches", []):
        if (e["shape"], e["block_map"]) == (shape, block_map):
            return e.get(f"{kernel}_bytes_per_launch")
    return None


def timed_run(wl, args, precision, hmode):
    """Upload, Euler start-up, warm-up, then exactly args.steps timed steps
    bracketed by device synchronisations.  Returns the rate, the stats of the
    library (pass timings from HIP events on its stream) and the final
    positions."""
    from dem3d_amd.native import System, autotune_steps, load_hip
    g = System(load_hip(), wl.red, skin=args.skin, history_mode=hmode, precision=precision)
    g.upload(wl.comp, wl.x, wl.v, wl.w)
    g.startup()
    tune = None
    shape = (0, 0)  # the library's default launch shape and block order
    if args.pass_shape >= 0:
        shape = (args.pass_shape, max(args.block_map, 0))
        g.set_option("pass_shape", args.pass_shape)
        if args.block_map >= 0:
            g.set_option("block_map", args.block_map)
        if args.tune_steps > 0:
            # the steps the autotuner would take first: the timed steps are
            # then those of an autotuned run (every shape gives the same
            # trajectory), as profiling runs of one shape need
            g.step(autotune_steps(args.tune_steps))
    elif args.tune_steps > 0:
        # dem_autotune: every launch shape of the passes timed on the live bed
        # (these are ordinary steps: the trajectory does not change)
        tune = g.autotune(steps_per_shape=args.tune_steps)
        shape = (tune["chosen"], tune["block_map"])
        print(f"autotune: shape {tune['chosen']} of {tune['block']}/{tune['waves']}/"
              f"{tune['front']}, block order {tune['block_map']} {tune['ms_map']}, "
              f"ms {[round(m, 4) for m in tune['ms']]}", file=sys.stderr, flush=True)
    if args.warmup:
        g.step(args.warmup)
    g.lib.set_option(g.h, b"timing", 1.0)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    t1 = time.perf_counter()
    st = g.stats()
    sec = t1 - t0
    out = {"ms": 1e3 * sec / args.steps, "value": wl.np * args.steps / sec, "stats": st,
           "obs": g.observables(), "tune": tune, "shape": shape,
           "steps_before": g.next_step - args.steps - wl.red.method_order}
    g.close()
    return out


def ensemble_run(args):
    """The 256-system ensemble (sizes log-uniform in [500, 20 000]): batched
    mixed precision for --ens-steps steps, batched FP64, and the plain FP64
    path (each system stepped alone, on a --alone-steps sample)."""
    from dem3d_amd import workloads
    from dem3d_amd.native import DEM_FP64, DEM_MIXED, Batch, System, autotune_steps, load_hip
    lib = load_hip()
    members = workloads.ensemble_256()
    ntot = sum(w.np for w in members)

    def build(precision):
        out = []
        for w in members:
            s = System(lib, w.red, precision=precision)
            s.upload(w.comp, w.x, w.v, w.w)
            s.startup()
            out.append(s)
        for s in out:
            s.synchronize()
        return out

    def batched(systems, steps, graph=False, shape=None, tune_steps=0):
        """Launches one by one (the default) or replayed from captured HIP
        graphs; autotuned launch shape (tune_steps > 0) or a given one.
        Throughput over `steps` steps, then the per-step device times of
        args.latency_steps more steps."""
        b = Batch(lib)
        b.set_option("graph", 1 if graph else 0)
        if shape is not None:
            b.set_option("pass_shape", shape)
        for s in systems:
            b.add(s)
        b.step(20)  # warm-up
        tune = None
        if tune_steps:
            tune = b.autotune(steps_per_shape=tune_steps)
        else:
            b.step(autotune_steps(args.tune_steps))  # the same trajectory as a tuned run
        b.step(8)
        t0 = time.perf_counter()
        b.step(steps)
        sec = time.perf_counter() - t0
        b.set_timing(True)
        b.step(args.latency_steps)
        st = b.step_times()
        b.close()
        return ntot * steps / sec, st, tune

    res = {"systems": len(members), "spheres": ntot,
           "sizes": [int(min(w.np for w in members)), int(max(w.np for w in members))]}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_ensemble_baseline(members, args.cpu_ens_steps,
                                    min(args.cpu_workers, os.cpu_count() or 1))
        print(f"ensemble cpu baseline: {cpu}", file=sys.stderr, flush=True)
    sysm = build(DEM_MIXED)
    v_mx, st_mx, tune_mx = batched(sysm, args.ens_steps, tune_steps=args.tune_steps)
    for s in sysm:
        s.close()
    print(f"ensemble mixed batched: {v_mx:.4g} particle-steps/s, tune {tune_mx}",
          file=sys.stderr, flush=True)
    sys64 = build(DEM_FP64)
    v_64b, st_64, tune_64 = batched(sys64, args.ens_steps, tune_steps=args.tune_steps)
    for s in sys64:
        s.close()
    # the same steps from the same start, replayed from captured graphs
    sys64 = build(DEM_FP64)
    v_64g, st_64g = None, None
    if not args.no_ens_graph:
        v_64g, st_64g, _ = batched(sys64, args.ens_steps, graph=True,
                                   shape=tune_64["chosen"] if tune_64 else 0)
        print(f"ensemble fp64 batched: {v_64b:.4g} launched one by one, {v_64g:.4g} from graphs",
              file=sys.stderr, flush=True)
    # the same in persistent launches (one cooperative launch per 250 steps,
    # a group of resident blocks per system)
    def persistent(precision, shape):
        systems = build(precision)
        b = Batch(lib)
        b.set_option("persistent", 1)
        if shape is not None:
            b.set_option("pass_shape", shape)
        for s_ in systems:
            b.add(s_)
        b.step(20)  # warm-up
        t0 = time.perf_counter()
        b.step(args.ens_steps)
        sec = time.perf_counter() - t0
        b.close()
        for s_ in systems:
            s_.close()
        return ntot * args.ens_steps / sec

    v_mxp = v_64p = None
    if not args.no_ens_persistent:
        v_mxp = persistent(DEM_MIXED, tune_mx["chosen"] if tune_mx else None)
        v_64p = persistent(DEM_FP64, tune_64["chosen"] if tune_64 else None)
        print(f"ensemble persistent: mixed {v_mxp:.4g}, fp64 {v_64p:.4g} particle-steps/s",
              file=sys.stderr, flush=True)
    t_alone = 0.0
    ps_alone = 0.0
    for s, w in zip(sys64, members):
        s.synchronize()
        t0 = time.perf_counter()
        s.step(args.alone_steps)
        s.synchronize()
        t_alone += time.perf_counter() - t0
        ps_alone += w.np * args.alone_steps
    for s in sys64:
        s.close()
    v_alone = ps_alone / t_alone
    print(f"ensemble fp64 batched: {v_64b:.4g}, alone: {v_alone:.4g}", file=sys.stderr,
          flush=True)
    res.update({
        "mixed_batched_particle_steps_per_s": v_mx,
        "mixed_batched_steps": args.ens_steps,
        "step_ms_p50": float(np.percentile(st_mx, 50)),
        "step_ms_p99": float(np.percentile(st_mx, 99)),
        "mixed_pass_shape": tune_mx,
        "fp64_batched_particle_steps_per_s": v_64b,
        "fp64_batched_step_ms_p50": float(np.percentile(st_64, 50)),
        "fp64_pass_shape": tune_64,
        "fp64_batched_graph_particle_steps_per_s": v_64g,
        "fp64_batched_graph_step_ms_p50": (float(np.percentile(st_64g, 50))
                                           if st_64g is not None else None),
        "graph_speedup_fp64": v_64g / v_64b if v_64g else None,
        "mixed_persistent_particle_steps_per_s": v_mxp,
        "fp64_persistent_particle_steps_per_s": v_64p,
        "persistent_speedup_mixed": v_mxp / v_mx if v_mxp else None,
        "persistent_speedup_fp64": v_64p / v_64b if v_64p else None,
        "fp64_alone_particle_steps_per_s": v_alone,
        "fp64_alone_sample_steps": args.alone_steps,
        "speedup_mixed_batched_vs_fp64_alone": v_mx / v_alone,
        "speedup_mixed_persistent_vs_fp64_alone": v_mxp / v_alone if v_mxp else None,
        "speedup_fp64_batched_vs_fp64_alone": v_64b / v_alone,
        "cpu_reference_throughput": cpu,
        "speedup_mixed_batched_vs_cpu_reference": v_mx / cpu["value"] if cpu else None,
    })
    return res


def serving_run(args):
    """The Poisson serving trace (200 requests at 50 req/s, sizes log-uniform
    in [1 000, 500 000], steps U[1 000, 20 000]; --serve-requests cuts it),
    served in real time with continuous batching in mixed precision."""
    from dem3d_amd import serve, workloads
    from dem3d_amd.native import DEM_MIXED, load_hip
    trace = workloads.serving_trace()[: args.serve_requests]
    last = [0.0]

    def log(m):  # a progress line every few seconds
        t = time.perf_counter()
        if t - last[0] >= 5.0:
            last[0] = t
            print(m, file=sys.stderr, flush=True)

    rep = serve.serve(load_hip(), trace, precision=DEM_MIXED, chunk=100, log=log)
    out = rep.summary()
    out["trace"] = (f"{'all' if len(trace) == 200 else f'first {len(trace)} of'} 200 requests, "
                    f"Poisson 50/s, n log-uniform [1000, 500000], steps U[1000, 20000], "
                    f"mixed precision; built on a worker thread while the batch steps")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=4_000_000, help="spheres of the packed bed")
    ap.add_argument("--settle-steps", type=int, default=1000,
                    help="steps of the packed bed sheared under gravity before the timing")
    ap.add_argument("--cells", default="100,100,50",
                    help="rock-salt cells (8 spheres each) of the crystal leg")
    ap.add_argument("--bed-cache", default="",
                    help="npz file holding the packed bed: loaded when it exists, else written "
                         "after packing (profiler runs pack once, outside the profiler)")
    ap.add_argument("--pack-only", action="store_true",
                    help="pack the bed (and write --bed-cache), then exit")
    ap.add_argument("--crystal-only", action="store_true",
                    help="time only the ordered (rock-salt) bed of --cells for --steps steps "
                         "and print its pass timings (profiling runs)")
    ap.add_argument("--crystal-steps", type=int, default=500,
                    help="timed steps of the crystal leg (0 = skip it)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-configs", action="store_true",
                    help="only time the reference on the 2 000-sphere settling case (20 000 "
                         "steps) and the 200 000-sphere loose packing (--cpu-poly-steps) on "
                         "this host, one JSON line each")
    ap.add_argument("--cpu-poly-steps", type=int, default=100)
    ap.add_argument("--poly-cache", default="",
                    help="npz file holding the 200 000-sphere loose packing: loaded when it "
                         "exists, else grown on the GPU and written")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skin", type=float, default=0.0)
    ap.add_argument("--precision", choices=("fp64", "both"), default="both",
                    help="fp64: the FP64 path only; both: also time the mixed-precision mode and "
                         "compare its energy and packing fraction with the FP64 run")
    ap.add_argument("--ens-steps", type=int, default=10000,
                    help="steps of the batched 256-system ensemble (0 = skip the ensemble)")
    ap.add_argument("--latency-steps", type=int, default=2000,
                    help="batched steps timed one by one for the p50/p99 step latency")
    ap.add_argument("--tune-steps", type=int, default=20,
                    help="timed steps per launch shape of the autotuner (0 = default shape)")
    ap.add_argument("--block-map", type=int, default=-1,
                    help="with --pass-shape: block order of the force passes (0 = particle "
                         "ranges per XCD, 1 = dispatch order)")
    ap.add_argument("--no-ens-graph", action="store_true",
                    help="skip the graph-replay run of the ensemble (profiler runs)")
    ap.add_argument("--no-ens-persistent", action="store_true",
                    help="skip the persistent (cooperative) runs of the ensemble (profiler runs: "
                         "rocprofv3 can fail at exit after cooperative launches)")
    ap.add_argument("--pass-shape", type=int, default=-1,
                    help="launch shape of the 4M runs' force passes instead of autotuning "
                         "(profiling runs; with --tune-steps > 0 the run first takes the "
                         "steps the autotuner would, so its timed steps are the same)")
    ap.add_argument("--cpu-ens-steps", type=int, default=20,
                    help="steps per system of the reference's ensemble throughput sample "
                         "(0 = skip it)")
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="host cores (one process each) of the ensemble throughput sample")
    ap.add_argument("--alone-steps", type=int, default=100,
                    help="steps per system of the one-at-a-time FP64 sample")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the serving trace to run (0 = skip serving)")
    ap.add_argument("--history", choices=("reference", "canonical"), default="reference",
                    help="orientation of stored tangential history (see include/dem3d.h)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device-to-device copy kernel (reads 1 GiB, writes "
                         "1 GiB) to calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        import torch
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return 0
    if args.cpu_configs:
        return cpu_configs(args)
    if args.gpus != 1:
        print(f"bench.py: the DEM path runs on one GPU; --gpus {args.gpus} is not supported "
              f"(a multi-GPU run would decompose the box into slabs, one per GPU)",
              file=sys.stderr)
        return 2
    from dem3d_amd import workloads
    from dem3d_amd.native import DEM_FP64, DEM_HIST_CANONICAL, DEM_HIST_REFERENCE, DEM_MIXED, load_hip
    lib = load_hip()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.crystal_only:
        # the ordered bed alone, --steps timed steps (profiling runs)
        hmode = DEM_HIST_REFERENCE if args.history == "reference" else DEM_HIST_CANONICAL
        cwl = workloads.bidisperse_sheared_4m(cells=tuple(int(x) for x in args.cells.split(",")))
        rc4 = timed_run(cwl, args, DEM_FP64, hmode)
        cst = rc4["stats"]
        print(json.dumps({"workload": cwl.name, "spheres": cwl.np, "value": rc4["value"],
                          "unit": "particle-steps/s", "ms_per_step": rc4["ms"],
                          "steps": args.steps, "ms_pass1": cst["ms_pass1"],
                          "ms_pass2": cst["ms_pass2"], "ms_rebuild_per_step": cst["ms_rebuild"],
                          "pair_contacts_mean": cst["mean_pair_contacts"],
                          "wall_contacts": cst["wall_contacts"],
                          "neighbour_entries_per_sphere": cst["neighbour_pairs"] / cwl.np,
                          "autotune": rc4["tune"]}), flush=True)
        return 0
    t0 = time.perf_counter()
    state = None
    if args.bed_cache and os.path.exists(args.bed_cache):
        with np.load(args.bed_cache) as z:
            state = (z["x"], z["v"], z["w"])
    wl = workloads.packed_bed(lib, n=args.n, settle_steps=args.settle_steps, log=log,
                              state=state)
    if args.bed_cache and state is None:
        np.savez(args.bed_cache, x=wl.x, v=wl.v, w=wl.w)
    if args.pack_only:
        return 0
    print(f"workload: {wl.name} np={wl.np} {wl.note} ({time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, args.cpu_steps)
        print(f"cpu baseline: {cpu}", file=sys.stderr, flush=True)
    hmode = DEM_HIST_REFERENCE if args.history == "reference" else DEM_HIST_CANONICAL
    r64 = timed_run(wl, args, DEM_FP64, hmode)
    print(f"fp64: {r64['ms']:.4f} ms/step", file=sys.stderr, flush=True)
    st, ms, value = r64["stats"], r64["ms"], r64["value"]
    obs = r64["obs"]
    mixed = None
    if args.precision == "both":
        rmx = timed_run(wl, args, DEM_MIXED, hmode)
        print(f"mixed: {rmx['ms']:.4f} ms/step", file=sys.stderr, flush=True)
        phi64 = r64["obs"]["packing_fraction"]
        phimx = rmx["obs"]["packing_fraction"]
        ke64, kemx = r64["stats"]["kinetic_energy"], rmx["stats"]["kinetic_energy"]
        mixed = {
            "value": rmx["value"],
            "ms_per_step": rmx["ms"],
            "speedup_vs_fp64": rmx["value"] / value,
            "ms_pass1": rmx["stats"]["ms_pass1"],
            "ms_pass2": rmx["stats"]["ms_pass2"],
            "pair_contacts": rmx["stats"]["pair_contacts"],
            "autotune": rmx["tune"],
            "after_steps": rmx["steps_before"] + args.steps,
            "kinetic_energy_rel_diff_vs_fp64": (kemx - ke64) / ke64 if ke64 else None,
            "packing_fraction_fp64": phi64,
            "packing_fraction_mixed": phimx,
            "packing_fraction_diff": phimx - phi64,
        }
    crystal = None
    if args.crystal_steps > 0:
        cwl = workloads.bidisperse_sheared_4m(cells=tuple(int(x) for x in args.cells.split(",")))
        cargs = argparse.Namespace(**vars(args))
        cargs.steps = args.crystal_steps
        cargs.precision = "fp64"
        rc4 = timed_run(cwl, cargs, DEM_FP64, hmode)
        cst = rc4["stats"]
        crystal = {"workload": cwl.name, "packing": cwl.note, "spheres": cwl.np,
                   "value": rc4["value"], "ms_per_step": rc4["ms"], "steps": args.crystal_steps,
                   "ms_pass1": cst["ms_pass1"], "ms_pass2": cst["ms_pass2"],
                   "ms_rebuild_per_step": cst["ms_rebuild"],
                   "pair_contacts_mean": cst["mean_pair_contacts"],
                   "neighbour_entries_per_sphere": cst["neighbour_pairs"] / cwl.np,
                   "autotune": rc4["tune"]}
        print(f"crystal: {rc4['ms']:.4f} ms/step", file=sys.stderr, flush=True)
    ensemble = ensemble_run(args) if args.ens_steps > 0 else None
    serving = serving_run(args) if args.serve_requests > 0 else None
    nbr = st["neighbour_pairs"] / wl.np
    # contacts averaged over the timed steps (the bed relaxes during the run)
    P = st["mean_pair_contacts"] or st["pair_contacts"]
    W = st["wall_contacts"]
    b1 = PASS1_B_PER_PARTICLE * wl.np + PASS1_B_PER_CONTACT * (P + W)
    b2 = PASS2_B_PER_PARTICLE * wl.np + PASS2_B_PER_CONTACT * (P + W)
    dom_ms, dom_b, dom = (st["ms_pass1"], b1, "pass1")
    if st["ms_pass2"] > st["ms_pass1"]:
        dom_ms, dom_b, dom = (st["ms_pass2"], b2, "pass2")
    achieved = dom_b / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    traffic = traffic_from_profiles(wl.name, dom, *r64["shape"])
    out = {
        "metric": "particle-steps/sec at 4M spheres (FP64 vs mixed) + batched p50/p99 step latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "spheres": wl.np,
            "packing": wl.note,
            "packing_fraction_bulk": obs["packing_fraction"],
            "coordination_number": obs["coordination"],
            "wall_contacts_final": obs["wall_contacts"],
            "neighbour_entries_per_sphere": nbr,
            "precision": "fp64",
            "history_mode": args.history,
            "pair_contacts_mean": P,
            "pair_contacts_final": st["pair_contacts"],
            "wall_contacts": W,
            "rebuilds": st["rebuilds"],
            "ms_pass1": st["ms_pass1"],
            "ms_pass2": st["ms_pass2"],
            "ms_rebuild_per_step": st["ms_rebuild"],
            "max_halo": st["max_halo"],
            "halo_overflow_tiles": st["halo_overflow"],
            "steps_before_timing": r64["steps_before"],
            "autotune": r64["tune"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_launch_shape": list(r64["shape"]),
            "bytes_per_launch": dom_b,
            "step_frac": (b1 + b2) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "cpu_baseline": cpu,
        "mixed": mixed,
        "crystal": crystal,
        "ensemble": ensemble,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
