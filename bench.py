#!/usr/bin/env python3
"""FLIP particle-steps/s on one MI355X.

Workload (N=1): the large single scene -- 256^3 dam break [0,0.4]x[0,0.6]x[0,1] plus a falling
sphere (centre (0.7,0.75,0.5), radius 0.12, v0 = (0,-2,0)), 8 particles per cell, solid walls on
all faces, dt = 1/600, FP32 particle state, pressure by FP32 PCG under an FP64 defect-correction
loop to 1e-10 relative residual. A step = one full FLIP substep (move, bin, P2G, extrapolate,
constrain, gravity, fluid SDF, assembly, PCG, velocity update, extrapolate, constrain, G2P) with the
scene resident in HBM.

Prints one JSON line. `roofline` is for the dominant kernel (the PCG SpMV), from HIP events around
every launch over the timed steps; `cpu_baseline` times the oracle (FP32, the reference's IC(0)
PCG) on the host cores on a bounded sample of the same workload.

Two more legs ride on the same line:
  batched  -- 512 independent scenes, edge length log-uniform 24..96 cells, fluid-box extents
              uniform in [0.2, 0.8] per axis, N(0, 0.5^2) initial velocities, refined pressure to
              1e-9: scenes*steps/s with every stage launched once for all scenes, against the plain
              path (each scene stepped through its own handle, one after the other).
  served   -- Poisson arrivals (200/s) of step or pressure-solve requests (50/50) on a resident pool
              of scenes with edge length log-uniform 32..192, HBM admission cap 200 GB: completed
              requests/s and p50/p99 latency, with continuous batching and with one request at a time.
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


def spmv_bytes_per_cell():
    # k_pcg_spmv per active cell: read z, s (8 B) + diag, cx, cy, cz (16 B); write s, q (8 B)
    return 32.0


def cpu_baseline(n, threads, steps=3):
    """The oracle in FP32 (reference algorithm incl. serial IC(0)-PCG, threaded particle/grid loops)
    on the n^3 version of the workload, one substep; returns (particle-steps/s, seconds, particles)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np

    import banana_amd
    from banana_amd import scenes
    import oracle_lib

    if not os.path.exists(oracle_lib.ORACLE_SO):
        oracle_lib.build()
    params, spec = scenes.dam_break(n, sphere=True, boundary_sdf_at_init=True, cg_tol=1e-6)
    pos, vel = banana_amd.generate_scene_host(params, spec, 2024)
    o = oracle_lib.Oracle(params, precision=32, threads=threads)
    o.set_particles(pos, vel)
    t0 = time.perf_counter()
    iters = 0
    for _ in range(steps):
        o.step(np.float32(1.0 / 600.0))
        iters += o.pcg_stats()[0]
    dt = time.perf_counter() - t0
    return pos.shape[0] * steps / dt, dt, pos.shape[0], iters / steps


def batched_leg(args, log):
    """scenes*steps/s of the uneven batch: fused launches vs scene-by-scene (same scenes, same kernels)."""
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
    b.step(sims, dt)  # warm the batched path (work lists, arenas)
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
    n_ss = len(sims) * args.batch_steps
    out = {"scenes": len(sims), "steps": args.batch_steps, "particles": np_total, "cells": cells,
           "scene_steps_per_s": round(n_ss / batched_s, 2), "plain_scene_steps_per_s": round(n_ss / plain_s, 2),
           "speedup_vs_plain": round(plain_s / batched_s, 3),
           "particle_steps_per_s": round(np_total * args.batch_steps / batched_s, 1),
           "pcg_iterations_mean": round(float(np.mean(its)), 1), "all_converged": bool(ok and worst == 0),
           "pressure": "FP32 PCG under FP64 defect correction to 1e-9"}
    for g in sims:
        g.close()
    log(f"batched: {out}")
    return out


def served_leg(args, log, batching):
    """Poisson request stream against a resident pool; returns the server's own accounting."""
    import numpy as np

    from banana_amd import scenes
    from banana_amd.serving import Server

    rng = np.random.default_rng(11)
    dt = np.float32(1.0 / 600.0)
    srv = Server(hbm_budget_bytes=200e9, workers=4 if batching else 1, batching=batching)
    t0 = time.perf_counter()
    ids = []
    for i in range(args.serve_pool):
        p, s = scenes.random_scene(rng, 32, 192)
        p.update(pressure_mode=1, refine_tolerance=1e-9)
        ids.append(srv.add_scene(p, s, 5000 + i))
    warm = [srv.submit(i, 0, dt) for i in ids]  # every scene resident and stepped once
    for r in warm:
        srv.wait(r)
    srv.reset_stats()
    log(f"served ({'batched' if batching else 'one at a time'}): pool of {len(ids)} resident in "
        f"{time.perf_counter() - t0:.1f}s, {srv.stats()['resident_bytes'] / 1e9:.1f} GB")
    arr = np.random.default_rng(13)
    t_next = srv.now()
    end = t_next + args.serve_seconds
    n = 0
    while True:
        t_next += arr.exponential(1.0 / args.serve_rate)
        if t_next > end:
            break
        wait = t_next - srv.now()
        if wait > 0:
            time.sleep(wait)
        srv.submit(ids[int(arr.integers(len(ids)))], int(arr.integers(2)), dt)
        n += 1
    srv.drain(600.0)
    st = srv.stats()
    srv.close()
    out = {"requests": n, "completed": st["completed"], "failed": st["failed"],
           "completed_per_s": round(st["throughput_rps"], 2), "latency_p50_ms": round(st["latency_p50_ms"], 2),
           "latency_p99_ms": round(st["latency_p99_ms"], 2), "latency_mean_ms": round(st["latency_mean_ms"], 2),
           "mean_batch": round(st["mean_batch"], 2), "peak_resident_gb": round(st["peak_resident_bytes"] / 1e9, 2)}
    log(f"served: {out}")
    return out


def measured_traffic(grid):
    """HBM bytes per k_pcg_spmv launch (FETCH_SIZE + WRITE_SIZE counter runs, scaled by a 1 GiB copy)
    as recorded in profiles/ for this workload's grid, else None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_pcg_spmv_dambreak256_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("grid") != grid:
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
    ap.add_argument("--mg-sweeps", type=int, default=None, help="smoothing sweeps per multigrid level (library default when omitted)")
    ap.add_argument("--mg-coarsest-sweeps", type=int, default=None,
                    help="symmetric sweeps of the coarsest multigrid level (library default when omitted)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="after the steps, copy this many bytes device-to-device once (a known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    ap.add_argument("--cpu-n", type=int, default=128, help="grid edge of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=3, help="substeps of the CPU baseline sample")
    ap.add_argument("--timing", type=int, default=2, help="0: none, 1: stage events, 2: + per-kernel events")
    ap.add_argument("--batch-scenes", type=int, default=512, help="scenes of the batched leg (0: skip the leg)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="length of the Poisson request stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="request arrivals per second")
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
    params.update(pressure_mode=1 if args.pressure == "refined" else 0, refine_tolerance=1e-10, timing=args.timing,
                  preconditioner=1 if args.precond == "mg" else 0)
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

    spmv_ms = 0.0
    spmv_launches = 0
    spmv_bytes = 0.0
    pcg_iters = 0
    per_step = []
    stage_ms = {}
    stage_bytes = {}
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        sim.step(dt)
        st = sim.stats()  # the step has synchronized (end of step collects its events)
        per_step.append(time.perf_counter() - ts)
        it = st["pcg_iterations"]
        pcg_iters += it
        spmv_ms += st["kernel_ms"]["pcg_spmv"]
        spmv_launches += st["kernel_launches"]["pcg_spmv"]
        spmv_bytes += spmv_bytes_per_cell() * st["num_active_cells"] * it
        for k, v in st["stage_ms"].items():
            stage_ms[k] = stage_ms.get(k, 0.0) + v
        for k, v in st["stage_bytes"].items():
            stage_bytes[k] = stage_bytes.get(k, 0.0) + v
        log(f"step {i + 1}/{args.steps}: {1e3 * per_step[-1]:.1f} ms, pcg {it} (outer {st['outer_iterations']}), "
            f"residual {st['pcg_residual']:.2e}/{st['rhs_max']:.2e}, active cells {st['num_active_cells']}")
    sim.synchronize()
    elapsed = time.perf_counter() - t_start
    np_ = sim.num_particles()
    value = np_ * args.steps / elapsed

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
    if spmv_launches and spmv_ms > 0:
        avg_ms = spmv_ms / spmv_launches
        achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
        roofline = {"kernel": "k_pcg_spmv", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(list(sim.n)),
                    "avg_launch_us": round(avg_ms * 1e3, 2), "launches": spmv_launches,
                    "bytes_per_launch": spmv_bytes / max(pcg_iters, 1)}

    cpu = None
    if not args.no_cpu_baseline:
        threads = os.cpu_count() or 1
        try:
            aff = len(os.sched_getaffinity(0))
            threads = min(threads, aff)
        except Exception:
            pass
        threads = min(threads, 16)
        log(f"cpu baseline: oracle FP32 on {threads} threads, {args.cpu_n}^3 ...")
        v, secs, npc, iters = cpu_baseline(args.cpu_n, threads, args.cpu_steps)
        cpu = {"value": round(v, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_steps} substeps of the {args.cpu_n}^3 version of this scene ({npc} particles), "
                         f"FP32 oracle with the reference's IC(0)-PCG at rel-tol 1e-6 ({iters:.0f} iterations/step), {secs:.1f} s"}

    batched = batched_leg(args, log) if args.batch_scenes > 0 else None
    served = None
    if args.serve_seconds > 0:
        served = {"arrival_rate_per_s": args.serve_rate, "seconds": args.serve_seconds, "pool": args.serve_pool,
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
                   "preconditioner": args.precond, "stage_ms_per_step": {k: round(v / args.steps, 3) for k, v in stage_ms.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "zsort": {"ms_per_call": round(zsort_ms[-1], 3), "particles": np_},
        # per stage: algorithmic bytes (the library's byte model; PCG counts 64 B per fluid cell and
        # inner iteration, the Jacobi form, so the multigrid's extra traffic is not in it) over the
        # stage's event time
        "stage_roofline": {k: {"gbs": round(stage_bytes[k] / (stage_ms[k] * 1e6), 1),
                               "frac": round(stage_bytes[k] / (stage_ms[k] * 1e6) / HBM_PEAK_GBS, 3)}
                           for k in stage_ms if stage_ms[k] > 0 and stage_bytes.get(k, 0) > 0},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
