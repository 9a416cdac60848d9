#!/usr/bin/env python3
"""FLIP particle-steps/s on one MI355X.

Workload (N=1): the large single scene -- 256^3 dam break [0,0.4]x[0,0.6]x[0,1] plus a falling
sphere (centre (0.7,0.75,0.5), radius 0.12, v0 = (0,-2,0)), 8 particles per cell, solid walls on
all faces, dt = 1/600, FP32 particle state, pressure by FP32 PCG under an FP64 defect-correction
loop to 1e-10 relative residual. A step = one full FLIP substep (move, bin, P2G, extrapolate,
constrain, gravity, fluid SDF, assembly, PCG, velocity update, extrapolate, constrain, G2P) with the
scene resident in HBM.

Prints one JSON line. `value` and `ms_per_step` come from the timed steps, run without any event
instrumentation. A second, instrumented pass (--profile-steps) records a HIP event pair around every
launch of the main kernels: `kernels` lists each one's launches, average duration, algorithmic bytes
(KERNEL_BYTES below) and achieved GB/s; `roofline` is the entry with the largest total time.
`cpu_baseline` times the oracle (FP32, the reference's IC(0)-PCG) on the host cores on bounded
samples: the 128^3 version of this scene, and the dam-break configurations at 64^3 (1 and 10 steps,
both boundary modes, all threads and one thread) and 128^3 (1 step).

Two more legs ride on the same line:
  batched  -- 512 independent scenes, edge length log-uniform 24..96 cells, fluid-box extents
              uniform in [0.2, 0.8] per axis, N(0, 0.5^2) initial velocities, refined pressure to
              1e-9: scenes*steps/s of fused batch calls (every stage one launch over all scenes'
              blocks, one host wait per sync point for all scenes, repeated PCG launch sequences
              replayed from hipGraphs, launch shape chosen by the autotuner during warm-up) against
              the plain path (each scene stepped through its own handle, one after the other).
  served   -- Poisson arrivals (200/s for 120 s) of step or pressure-solve requests (50/50) on a
              resident pool of scenes with edge length log-uniform 32..192, HBM admission cap
              200 GB: completed requests/s, p50/p99 latency and the measured mean number of requests
              in flight, served with continuous batching (every request in flight joins the fused
              launches) and one request at a time; then doubled rates until the server falls behind
              (its saturation throughput).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FLIP particle-steps/s at 256³; served scenes·steps/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic bytes per launch of each timed kernel: every input element it must read once and every
# output element it writes once, at storage precision (neighbour re-reads, atomics and temporaries not
# counted). Np particles, F faces (u+v+w), C cells, V nodes, A fluid cells (rows of the system).
KERNEL_BYTES = {
    "p2g": lambda Np, F, C, V, A: 24 * Np + 5 * F,          # grid coords + velocities; velocity + valid
    "g2p": lambda Np, F, C, V, A: 36 * Np + 8 * F,          # coords + velocities in/out; grid and backup
    "move": lambda Np, F, C, V, A: 24 * Np + 4 * F,         # positions in/out; grid velocities
    "fluid_sdf": lambda Np, F, C, V, A: 12 * Np + 8 * C + 4 * V,  # coords, cell ranges, phi; boundary SDF
    "permute": lambda Np, F, C, V, A: 72 * Np,              # order + 7 words in, 7 words + coords out
    "pcg_spmv": lambda Np, F, C, V, A: 32 * A,              # z, s, 4 coefficients in; s, q out
    "pcg_update": lambda Np, F, C, V, A: 44 * A,            # multigrid: fused with the finest pre-smoother -- x, r, s, q,
                                                            # 3 couplings, 1/diag in; x, r, pre-smoothed z out
    "mg_pre": lambda Np, F, C, V, A: 24 * A,                # 3 couplings, b, 1/diag in; x out
    "mg_restrict": lambda Np, F, C, V, A: 24.5 * A,         # b, diag, 3 couplings, x in; coarse b out
    "mg_prolong": lambda Np, F, C, V, A: 12.5 * A,          # x, 1/diag, coarse x in; x out
    "mg_post": lambda Np, F, C, V, A: 28 * A,               # 3 couplings, b, 1/diag, x in; x out
    "extrap": lambda Np, F, C, V, A: 11 * F / 3,            # one component, 4 layers: value + valid in; value, valid, extrapolated out
}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_step_timed(o, dt, stage_s):
    """One substep of the oracle, stage by stage as in its step() (PIC_Solver.cpp:292-294,
    FLIP_Solver.cpp:42-53, PIC_Solver.cpp:533-539), with the wall time of each stage added to
    stage_s (stage names after the reference's printRunTime calls); returns the solve's seconds."""
    seq = [("move", lambda: o.move_particles(dt)), ("p2g", o.map_particles_to_grid),
           ("extrapolate", o.extrapolate_velocity), ("constrain", o.constrain_grid_velocity),
           ("backup+gravity", lambda: (o.backup_grid_velocity(), o.add_gravity(dt))),
           ("weights", o.compute_cell_weights), ("phi", o.compute_fluid_sdf), ("assembly", lambda: o.compute_system(dt)),
           ("pcg", o.solve_system), ("update", lambda: o.update_projected_velocity(dt)),
           ("extrapolate", o.extrapolate_velocity), ("constrain", o.constrain_grid_velocity),
           ("g2p", lambda: (o.compute_changes_grid_velocity(), o.map_grid_to_particles()))]
    pcg_s = 0.0
    for name, fn in seq:
        t = time.perf_counter()
        fn()
        e = time.perf_counter() - t
        stage_s[name] = stage_s.get(name, 0.0) + e
        if name == "pcg":
            pcg_s += e
    return pcg_s


def cpu_run(n, sphere, walls, steps, threads, seed=2024):
    """The oracle in FP32 (the reference's algorithm: CAS P2G, IC(0)-PCG with the infinity-norm stop at
    rel-tol 1e-6) on an n^3 dam break; per-run totals."""
    import numpy as np

    import banana_amd
    from banana_amd import scenes
    import oracle_lib

    params, spec = scenes.dam_break(n, sphere=sphere, boundary_sdf_at_init=walls, cg_tol=1e-6)
    pos, vel = banana_amd.generate_scene_host(params, spec, seed)
    o = oracle_lib.Oracle(params, precision=32, threads=threads)
    o.set_particles(pos, vel)
    dt = np.float32(1.0 / 600.0)
    stage_s, iters, pcg_s, res = {}, 0, 0.0, 0.0
    t0 = time.perf_counter()
    first = None
    for k in range(steps):
        pcg_s += oracle_step_timed(o, dt, stage_s)
        it, res, _ = o.pcg_stats()
        iters += it
        if k == 0:
            first = time.perf_counter() - t0
    secs = time.perf_counter() - t0
    return {"workload": f"dam break {n}^3{' + sphere' if sphere else ''}, {'walls (boundary SDF at init)' if walls else 'reference default boundary'}",
            "particles": int(pos.shape[0]), "steps": steps, "threads": threads, "seconds": round(secs, 3),
            "first_step_s": round(first, 3), "particle_steps_per_s": round(pos.shape[0] * steps / secs, 1),
            "pcg_iterations": iters, "pcg_iterations_per_s": round(iters / pcg_s, 1) if pcg_s > 0 else None,
            "final_residual_inf": res, "stage_s": {k: round(v, 3) for k, v in stage_s.items()}}


def cpu_baseline(args, log):
    """Bounded CPU samples on the host: the 128^3 version of the headline scene (the `value`), and the
    dam-break configurations of the reference's CPU plan (64^3: 1 and 10 steps, both boundary modes,
    all threads and one thread; 128^3: 1 step)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    if not os.path.exists(oracle_lib.ORACLE_SO):
        oracle_lib.build()
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = nproc
    # the GPU box grants 16 host threads to one GPU's job, whatever nproc reports
    threads = max(1, min(aff, 16))
    model = cpu_model()
    log(f"cpu baseline: {model}, nproc {nproc}, affinity {aff}, using {threads} threads")
    head = cpu_run(args.cpu_n, True, True, args.cpu_steps, threads)
    log(f"cpu baseline {head['workload']}: {head}")
    runs = []
    if args.cpu_configs:
        for walls in (False, True):
            runs.append(cpu_run(64, False, walls, 10, threads))
            runs.append(cpu_run(64, False, walls, 1, 1))
        runs.append(cpu_run(128, False, True, 1, threads))
        for r in runs:
            log(f"cpu baseline run: {r}")
    return {"value": head["particle_steps_per_s"], "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{head['steps']} substeps of the {args.cpu_n}^3 version of this scene ({head['particles']} particles), "
                      f"FP32 oracle with the reference's IC(0)-PCG at rel-tol 1e-6 on {threads} threads, {head['seconds']} s",
            "cpu_model": model, "nproc": nproc, "affinity": aff, "headline_run": head, "runs": runs}


def batched_leg(args, log):
    """scenes*steps/s of the uneven batch: fused launches (bnn_flip_batch_step) vs scene-by-scene (same scenes, same kernels)."""
    import numpy as np

    import banana_amd
    from banana_amd import scenes
    from banana_amd.serving import Batch

    rng = np.random.default_rng(7)
    dt = np.float32(1.0 / 600.0)
    t0 = time.perf_counter()
    sims, np_total, cells = [], 0, 0
    for i in range(args.batch_scenes):
        p, s = scenes.random_scene(rng, 24, 96)
        p.update(pressure_mode=1, refine_tolerance=1e-9)
        g = banana_amd.FLIPSolver(p)
        g.generate_scene(s, 1000 + i)
        np_total += g.num_particles()
        cells += g.n[0] * g.n[1] * g.n[2]
        sims.append(g)
    log(f"batched: {len(sims)} scenes, {np_total} particles, {cells} cells, set up in {time.perf_counter() - t0:.1f}s")
    b = Batch()
    t_tune = time.perf_counter()
    for _ in range(5):  # warm the batched path; the autotuner tries its launch shapes on the first four calls
        b.step(sims, dt)
    tune_s = time.perf_counter() - t_tune
    for g in sims:
        g.step(dt)  # warm the plain path
    t1 = time.perf_counter()
    for _ in range(args.batch_steps):
        for g in sims:
            g.step(dt)
    plain_s = time.perf_counter() - t1
    t2 = time.perf_counter()
    worst = 0
    for _ in range(args.batch_steps):
        st = b.step(sims, dt)
        worst = max(worst, int(np.max(st)))
    batched_s = time.perf_counter() - t2
    its = [g.stats()["pcg_iterations"] for g in sims]
    ok = all(g.stats()["pcg_success"] for g in sims)
    rel = max(g.stats()["pcg_residual"] / max(g.stats()["rhs_max"], 1e-300) for g in sims)
    info = b.info()
    n_ss = len(sims) * args.batch_steps
    out = {"scenes": len(sims), "steps": args.batch_steps, "particles": np_total, "cells": cells,
           "scene_steps_per_s": round(n_ss / batched_s, 2), "plain_scene_steps_per_s": round(n_ss / plain_s, 2),
           "speedup_vs_plain": round(plain_s / batched_s, 3),
           "particle_steps_per_s": round(np_total * args.batch_steps / batched_s, 1),
           "pcg_iterations_mean": round(float(np.mean(its)), 1), "all_converged": bool(ok and worst == 0),
           "fp64_relative_residual_max": rel, "pressure": "FP32 PCG under FP64 defect correction to 1e-9",
           "fused": {"shape": info["shape"], "warmup_and_tuning_s": round(tune_s, 2), "flushes": info["flushes"],
                     "launches": info["launches"], "scene_launches": info["scene_launches"],
                     "graph_replays": info["graph_replays"], "graph_captures": info["graph_captures"],
                     "mean_scenes_per_flush": round(info["mean_scenes_per_flush"], 1)}}
    for g in sims:
        g.close()
    log(f"batched: {out}")
    return out


def served_leg(args, log, batching):
    """Poisson request streams against a resident pool, with the server's own accounting: the
    configured rate for serve_seconds, then higher rates for sat_seconds each until the server no
    longer keeps up (completions below 90 % of arrivals) -- its saturation throughput."""
    import numpy as np

    from banana_amd import scenes
    from banana_amd.serving import Server

    rng = np.random.default_rng(11)
    dt = np.float32(1.0 / 600.0)
    srv = Server(hbm_budget_bytes=200e9, workers=4, batching=batching, barrier_rounds=args.barrier_rounds,
                 max_batch_mcells=args.batch_mcells)
    t0 = time.perf_counter()
    ids = []
    for i in range(args.serve_pool):
        p, s = scenes.random_scene(rng, 32, 192)
        p.update(pressure_mode=1, refine_tolerance=1e-9)
        ids.append(srv.add_scene(p, s, 5000 + i))
    warm = [srv.submit(i, 0, dt) for i in ids]  # every scene resident and stepped once
    for r in warm:
        srv.wait(r)
    mode = "batched" if batching else "one at a time"
    log(f"served ({mode}): pool of {len(ids)} resident in {time.perf_counter() - t0:.1f}s, "
        f"{srv.stats()['resident_bytes'] / 1e9:.1f} GB")
    arr = np.random.default_rng(13)

    def stream(rate, seconds):
        srv.reset_stats()
        t_next = srv.now()
        end = t_next + seconds
        n = 0
        while True:
            t_next += arr.exponential(1.0 / rate)
            if t_next > end:
                break
            wait = t_next - srv.now()
            if wait > 0:
                time.sleep(wait)
            srv.submit(ids[int(arr.integers(len(ids)))], int(arr.integers(2)), dt)
            n += 1
        srv.drain(600.0)
        st = srv.stats()
        out = {"arrival_rate_per_s": rate, "seconds": seconds, "requests": n, "completed": st["completed"], "failed": st["failed"],
               "completed_per_s": round(st["throughput_rps"], 2), "latency_p50_ms": round(st["latency_p50_ms"], 2),
               "latency_p99_ms": round(st["latency_p99_ms"], 2), "latency_mean_ms": round(st["latency_mean_ms"], 2),
               "mean_batch": round(st["mean_batch"], 2), "peak_resident_gb": round(st["peak_resident_bytes"] / 1e9, 2)}
        log(f"served ({mode}) at {rate}/s: {out}")
        return out

    main = stream(args.serve_rate, args.serve_seconds if batching else min(args.serve_seconds, args.plain_serve_seconds))
    sweep = []
    rate = args.serve_rate
    for _ in range(args.sat_steps):
        rate *= 2
        r = stream(rate, args.sat_seconds)
        sweep.append(r)
        if r["completed_per_s"] < 0.9 * rate:
            break
    srv.close()
    main["saturation"] = sweep
    main["saturation_completed_per_s"] = max([main["completed_per_s"]] + [r["completed_per_s"] for r in sweep])
    return main


def measured_traffic(kernel, grid):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE counter runs, scaled by a copy of known
    size) as recorded in profiles/ for this workload's grid, else None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "flip_step_dambreak256_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("grid") != grid:
        return None
    return rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"]


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (its copy kernel reads and writes
    exactly nbytes), to scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:  # 3: device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="grid edge in cells (256 = the headline workload)")
    ap.add_argument("--pressure", choices=["refined", "fp32"], default="refined")
    ap.add_argument("--precond", choices=["mg", "jacobi"], default="mg")
    ap.add_argument("--warm-start", type=int, default=1, help="1: refined solves start from the previous pressure (0: from 0)")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="relative tolerance of each inner FP32 solve under the FP64 correction (library default when omitted)")
    ap.add_argument("--mg-sweeps", type=int, default=None, help="smoothing sweeps per multigrid level (library default when omitted)")
    ap.add_argument("--mg-coarsest-sweeps", type=int, default=None,
                    help="symmetric sweeps of the coarsest multigrid level (library default when omitted)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="after the steps, copy this many bytes device-to-device once (a known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    ap.add_argument("--cpu-n", type=int, default=128, help="grid edge of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=3, help="substeps of the CPU baseline sample")
    ap.add_argument("--profile-steps", type=int, default=2,
                    help="extra steps after the timed ones with HIP events around every stage and main kernel launch")
    ap.add_argument("--cpu-configs", type=int, default=1, help="also time the 64^3 and 128^3 dam-break CPU samples")
    ap.add_argument("--batch-scenes", type=int, default=512, help="scenes of the batched leg (0: skip the leg)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=120.0, help="length of the Poisson request stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="request arrivals per second")
    ap.add_argument("--plain-serve-seconds", type=float, default=30.0, help="length of the stream served one request at a time")
    ap.add_argument("--sat-seconds", type=float, default=10.0, help="length of each higher-rate stream of the saturation sweep")
    ap.add_argument("--sat-steps", type=int, default=4, help="rate doublings of the saturation sweep at most")
    ap.add_argument("--barrier-rounds", type=int, default=0, help="server rounds a request may wait at a phase boundary")
    ap.add_argument("--batch-mcells", type=int, default=0, help="server admission cap on grid cells in flight, millions (0: none)")
    ap.add_argument("--serve-pool", type=int, default=48, help="resident scenes the requests address")
    args = ap.parse_args()

    if args.gpus != 1:
        log(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
            "(scenes are independent and would shard one per GPU; see the design notes)")
        sys.exit(2)

    import numpy as np

    import banana_amd
    from banana_amd import scenes

    params, spec = scenes.dam_break(args.n, sphere=True, boundary_sdf_at_init=True, cg_tol=1e-6)
    params.update(pressure_mode=1 if args.pressure == "refined" else 0, refine_tolerance=1e-10, timing=0,
                  preconditioner=1 if args.precond == "mg" else 0)
    if args.inner_tol is not None:
        params["refine_inner_tolerance"] = args.inner_tol
    params["warm_start"] = args.warm_start
    if args.mg_sweeps is not None:
        params["mg_sweeps"] = args.mg_sweeps
    if args.mg_coarsest_sweeps is not None:
        params["mg_coarsest_sweeps"] = args.mg_coarsest_sweeps
    # the z-sort only runs when called (set_particles does not sort; steps never do); it is timed
    # once after the timed steps
    params.update(enable_sort_particle=1, sort_frequency=10)
    dt = np.float32(1.0 / 600.0)
    t0 = time.perf_counter()
    pos, vel = banana_amd.generate_scene_host(params, spec, 2024)
    log(f"generated {pos.shape[0]} particles in {time.perf_counter() - t0:.1f}s")
    sim = banana_amd.FLIPSolver(params)
    sim.set_particles(pos, vel)
    del pos, vel
    log(f"grid {sim.n}, particles {sim.num_particles()}")

    for i in range(args.warmup):
        t1 = time.perf_counter()
        sim.step(dt)
        st = sim.stats()
        log(f"warmup {i + 1}/{args.warmup}: {1e3 * (time.perf_counter() - t1):.1f} ms, pcg {st['pcg_iterations']} "
            f"(outer {st['outer_iterations']}), active cells {st['num_active_cells']}")
    sim.synchronize()

    # timed steps: no events in the stream
    pcg_iters = 0
    per_step = []
    rel_res = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        sim.step(dt)
        st = sim.stats()  # the step has synchronized
        per_step.append(time.perf_counter() - ts)
        pcg_iters += st["pcg_iterations"]
        rel_res.append(st["pcg_residual"] / st["rhs_max"] if st["rhs_max"] > 0 else 0.0)
        log(f"step {i + 1}/{args.steps}: {1e3 * per_step[-1]:.1f} ms, pcg {st['pcg_iterations']} (outer {st['outer_iterations']}), "
            f"residual {st['pcg_residual']:.2e}/{st['rhs_max']:.2e}, active cells {st['num_active_cells']}")
    sim.synchronize()
    elapsed = time.perf_counter() - t_start
    np_ = sim.num_particles()
    value = np_ * args.steps / elapsed

    # instrumented steps: HIP events around every stage and every launch of the main kernels
    kernels, stage_ms, stage_bytes, prof_iters = {}, {}, {}, 0
    if args.profile_steps > 0:
        sim.set_timing(2)
        acc_ms, acc_n = {}, {}
        work = {}
        for i in range(args.profile_steps):
            sim.step(dt)
            st = sim.stats()
            it, outer = st["pcg_iterations"], st["outer_iterations"]
            prof_iters += it
            Np, F = st["num_particles"], sum(int(np.prod(sim.field_shape(c))) for c in ("u", "v", "w"))
            C = st["num_cells"]
            V = (sim.n[0] + 1) * (sim.n[1] + 1) * (sim.n[2] + 1)
            A = st["num_active_cells"]
            for k, v in st["kernel_ms"].items():
                n = st["kernel_launches"][k]
                if not n:
                    continue
                acc_ms[k] = acc_ms.get(k, 0.0) + v
                acc_n[k] = acc_n.get(k, 0) + n
                # launches that did the work: PCG kernels past the stop flag return at once
                wl = it if k in ("pcg_spmv", "pcg_update") else (it + outer if k.startswith("mg_") and k != "mg_coarse" else n)
                wl = min(wl, n)
                b = KERNEL_BYTES[k](Np, F, C, V, A) if k in KERNEL_BYTES else None
                w = work.setdefault(k, [0, 0.0, b is not None])
                w[0] += wl
                if b is not None:
                    w[1] += b * wl
            for k, v in st["stage_ms"].items():
                stage_ms[k] = stage_ms.get(k, 0.0) + v
            for k, v in st["stage_bytes"].items():
                stage_bytes[k] = stage_bytes.get(k, 0.0) + v
            log(f"profiled step {i + 1}/{args.profile_steps}: pcg {it} (outer {outer})")
        sim.set_timing(0)
        for k in acc_ms:
            n, ms = acc_n[k], acc_ms[k]
            wl, bytes_, has_b = work[k]
            e = {"launches": n, "working_launches": wl, "avg_us": round(1e3 * ms / n, 2),
                 "ms_per_step": round(ms / args.profile_steps, 3)}
            if has_b and ms > 0:
                gbs = bytes_ / (ms * 1e-3) / 1e9
                e.update({"bytes_per_working_launch": round(bytes_ / wl), "achieved_gbs": round(gbs, 1),
                          "frac": round(gbs / HBM_PEAK_GBS, 4)})
            kernels[k] = e
    kernels = dict(sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_step"]))

    zsort_ms = []
    for _ in range(2):  # the first call also allocates its histogram buffer
        tz = time.perf_counter()
        sim.sort_particles()  # synchronous
        zsort_ms.append(1e3 * (time.perf_counter() - tz))
    log(f"z-sort of {np_} particles: {zsort_ms[-1]:.2f} ms")
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        log(f"copy calibration: {args.copy_calibration} bytes device-to-device")

    roofline = None
    dom = next((k for k, e in kernels.items() if "frac" in e), None)
    if dom is not None:
        e = kernels[dom]
        roofline = {"kernel": dom, "bound": "hbm", "achieved": e["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": e["frac"], "traffic": measured_traffic(dom, list(sim.n)), "avg_launch_us": e["avg_us"],
                    "launches": e["launches"], "working_launches": e["working_launches"],
                    "bytes_per_launch": e["bytes_per_working_launch"]}

    cpu = None if args.no_cpu_baseline else cpu_baseline(args, log)

    batched = batched_leg(args, log) if args.batch_scenes > 0 else None
    served = None
    if args.serve_seconds > 0:
        served = {"arrival_rate_per_s": args.serve_rate, "pool": args.serve_pool,
                  "batched": served_leg(args, log, True), "plain": served_leg(args, log, False)}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"FLIP dam break {args.n}^3 + falling sphere, 8 particles/cell, solid walls, dt=1/600, "
                               f"pressure {'FP32 PCG under FP64 defect correction to 1e-10' if args.pressure == 'refined' else 'FP32 PCG rel-tol 1e-6'}",
                   "grid": list(sim.n), "particles": np_, "pcg_iterations_per_step": pcg_iters / args.steps,
                   "pcg_iterations_per_s": round(pcg_iters / elapsed, 1),
                   "fp64_relative_residual_max": max(rel_res) if rel_res else None,
                   "preconditioner": args.precond, "warm_start": bool(args.warm_start),
                   "stage_ms_per_step": {k: round(v / args.profile_steps, 3) for k, v in stage_ms.items()} if args.profile_steps else None},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "zsort": {"ms_per_call": round(zsort_ms[-1], 3), "particles": np_},
        # per stage (instrumented steps): the library's algorithmic bytes (stats.stage_bytes) over the
        # stage's event time
        "stage_roofline": {k: {"gbs": round(stage_bytes[k] / (stage_ms[k] * 1e6), 1),
                               "frac": round(stage_bytes[k] / (stage_ms[k] * 1e6) / HBM_PEAK_GBS, 3)}
                           for k in stage_ms if stage_ms[k] > 0 and stage_bytes.get(k, 0) > 0},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
