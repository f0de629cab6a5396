#!/usr/bin/env python3
"""bench.py -- particle-substeps/s of the FLIP step on one MI355X.

Workload: the dam break at 1024x1024 cells, 8 jittered particles per cell
(2,342,752 FP32 particles), FP32 grids/SpMV/IPP with FP64 PCG reductions and
FP64 iterative refinement to relative residual 1e-6.  A "step" is one
FlipSolver::stepFrame() (a 1/30 s frame, CFL-split into substeps).  The scene
state stays resident in HBM; the timed region covers K frames after W warmup
frames, bracketed by device synchronisation.

Prints one JSON line (rank 0).  `roofline` is for the dominant kernel,
measured with HIP events around each of its launches on the stream it runs
on; `cpu_baseline` times the CPU restatement of the reference path (its own
PCG, IPP and density solve, the reference's thread-pool decomposition) on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s + p99 frame latency, batched vs plain GPU @1e-6 resid"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per work unit (per cell).  One PCG iteration is 54 B/cell:
# SpMV 9 + direction update 12 (k_pcg_a: 21) and axpy pair 24 + IPP 9
# (k_pcg_b: 33).
KERNEL_BYTES = {
    "pcg_b": ("K_PCG_B", 33.0, "cells"),
    "pcg_a": ("K_PCG_A", 21.0, "cells"),
}
# HBM bytes per cell actually moved, at the 1024^2 dam break: rocprofv3
# FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000) per dispatch, the factors being
# what the same counters read for a 1 GiB copy kernel
# (profiles/pcg_dam_break_1024_hbm_traffic.md).
MEASURED_TRAFFIC_PER_CELL = {1024: {"pcg_b": 34.18, "pcg_a": 21.83}}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3, help="untimed frames")
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--roofline-kernel", default="pcg_b", choices=sorted(KERNEL_BYTES))
    ap.add_argument("--cpu-substeps", type=int, default=4, help="substeps in the CPU-baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cpu count)")
    ap.add_argument("--stage-timing", action="store_true", help="extra run with per-stage event timing")
    return ap.parse_args()


def cpu_baseline(res, ppc, substeps, threads):
    """The oracle's reference-replica mode (reference PCG + IPP + max-abs test,
    cap 200, density solve gated as in the reference) on the same dam break."""
    import oracle as orc

    p, st = orc.dam_break(res, ppc=ppc, solver=orc.SOLVER_REF_REPLICA, density_mode=orc.DENSITY_GATED,
                          threads=threads, tol=1e-2)
    L = orc.lib()
    stats = orc.OrcStats()
    frame_dt = 1.0 / p.fps
    t_frame, count, done = 0.0, 0, C.c_int32(0)
    psub = 0
    t0 = time.perf_counter()
    for _ in range(substeps):
        vel = L.orc_max_particle_velocity(st.ref())
        dt = L.orc_frame_dt(vel, p.cfl_number, frame_dt, t_frame, count, p.max_substeps, C.byref(done))
        psub += st.n
        L.orc_substep(C.byref(p), st.ref(), dt, C.byref(stats))
        t_frame += dt
        count += 1
        if done.value:
            t_frame, count = 0.0, 0
    el = time.perf_counter() - t0
    return {
        "value": psub / el,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {substeps} substeps of the {res}x{res} dam break ({st.n} particles), "
                  f"reference PCG/IPP replica incl. the gated density solve, {threads} threads; {el:.1f} s",
        "seconds": el,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: scenes are independent and "
              "would shard whole across GPUs, which this library does not do", file=sys.stderr)
        sys.exit(2)
    import flipsolver2d_amd as fg

    ctx = fg.Context(0)
    scene = fg.dam_break_scene(ctx, a.res, ppc=a.ppc, tol=a.tol, pcg_iter_limit=4000)
    n0 = scene.particle_count()
    print(f"# dam break {a.res}^2 ppc {a.ppc}: {n0} particles", file=sys.stderr, flush=True)
    for w in range(a.warmup):
        s = scene.step_frame()
        print(f"# warmup frame {w}: {s.substeps} substeps, {s.pressure_iters} PCG iters, "
              f"{s.frame_ms:.1f} ms", file=sys.stderr, flush=True)

    kname, per_unit, unit_name = KERNEL_BYTES[a.roofline_kernel]
    ctx.profile_kernel(getattr(fg, kname))
    fg.lib().fg_ctx_profile_read(ctx.h, None, None, None)  # sync + reset
    frame_ms, psub, subs, iters, resid, conv = [], 0, 0, [], 0.0, True
    t0 = time.perf_counter()
    for k in range(a.steps):
        s = scene.step_frame()
        frame_ms.append(s.frame_ms)
        psub += s.particle_substeps
        subs += s.substeps
        iters.append(s.pressure_iters)
        resid = max(resid, s.pressure_rel_resid)
        conv = conv and bool(s.converged)
    ktime, klaunch, kunits = ctx.profile_read()   # synchronises the stream
    el = time.perf_counter() - t0
    ctx.profile_kernel(-1)
    print(f"# timed {a.steps} frames: {subs} substeps, {el:.3f} s", file=sys.stderr, flush=True)

    avg_ms = ktime / max(klaunch, 1)
    achieved = (kunits * per_unit) / (ktime * 1e-3) / 1e9 if ktime > 0 else 0.0
    bytes_per_launch = kunits * per_unit / max(klaunch, 1)
    copy_gbs, _ = ctx.bandwidth_probe(1 << 30, 5)
    tpc = MEASURED_TRAFFIC_PER_CELL.get(a.res, {}).get(a.roofline_kernel)
    traffic = tpc * kunits / max(klaunch, 1) if tpc else None

    stage = None
    if a.stage_timing:
        ctx.set_timing(True)
        s = scene.step_frame()
        ctx.set_timing(False)
        stage = {k: round(v, 3) for k, v in s.as_dict()["stage_ms"].items()}

    fs = sorted(frame_ms)
    p50 = statistics.median(fs)
    p99 = fs[min(len(fs) - 1, int(round(0.99 * (len(fs) - 1))))]
    scene.close()
    ctx.close()

    cpu = None
    if a.cpu_substeps > 0:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(a.res, a.ppc, a.cpu_substeps, threads)
        print(f"# cpu baseline: {cpu['value']:.3e} particle-substeps/s ({cpu['seconds']:.1f} s)",
              file=sys.stderr, flush=True)
        cpu.pop("seconds")

    value = psub / el
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain single-scene dam break {a.res}x{a.res}, {a.ppc} ppc ({n0} particles), "
                        f"FP32 state, FP64 PCG reductions + refinement to rel resid {a.tol:g}",
            "resolution": a.res, "particles_per_cell": a.ppc, "particles": n0, "step": "one stepFrame (1/30 s)",
            "substeps": subs, "solver_tol": a.tol,
        },
        "p50_frame_ms": p50,
        "p99_frame_ms": p99,
        "pcg_iters_per_substep_max": max(iters) if iters else 0,
        "max_rel_residual": resid,
        "all_converged": conv,
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{a.roofline_kernel}",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_us": avg_ms * 1e3,
            "launches": klaunch,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "bytes_per_unit": per_unit,
            "unit_of_work": unit_name,
            "measured_copy_gbs": copy_gbs,
        },
        "cpu_baseline": cpu,
    }
    if stage:
        out["stage_ms_one_frame"] = stage
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
