#!/usr/bin/env python3
"""bench.py -- particle-substeps/s of the FLIP step on one MI355X, batched vs plain.

Headline workload (`value`): 256 independent dam-break scenes of uneven size
(resolution per scene from {64, 96, 128, 192, 256}, fluid-rectangle corners
uniform in [0.1, 0.6] of the box, gravity 9.81 m/s^2 at +-30 degrees, 4-8
particles per cell), FP32 state, FP64 PCG reductions + FP64 refinement to
relative residual 1e-6.  A "step" is one frame (1/30 s, CFL-split) of every
scene, all advanced together by fg_batch_step_frame (one launch per phase).
The same scenes, stepped one at a time by fg_scene_step_frame on the same
frames, give the plain-GPU number; both sets do identical work (their
results are bit-identical).  State stays resident in HBM.

Also reported on the same line: the plain single-scene 1024^2 dam break
(8 particles per cell), the same scene under NBFlipSolver (narrow-band
FLIP, with the oracle's NB substep timed on the GPU's state beside it), the
isolated 4096^2 pressure projection, a served
stream (Poisson arrivals, HBM-budget
admission, continuous batching), `roofline` for the dominant kernel measured
with HIP events attached to its launches on the stream they run on (over as
many frames of the same workload right after the timed ones: events on every
launch slow the launch pipeline by ~7 %, so the timed frames run without
them), and `cpu_baseline`: the CPU restatement
of the reference path (its own PCG, IPP and gated density solve, the
reference's thread-pool decomposition) on a bounded sample of the workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s + p99 frame latency, batched vs plain GPU @1e-6 resid"
HBM_PEAK_GBS = 8000.0

# Kernel families bench.py reports a roofline for: (library family id,
# algorithmic bytes per work unit, unit, kernels of the family).  The library
# counts the bytes itself where the unit is 1 byte (P particles, C cells,
# F = F_U + F_V faces, C1 level-1 cells):
#   p2g        16 P + 5 F per substep (positions, velocities; faces + validity)
#   particles  36 P + 8 F + 4 C per substep: G2P + advection (k_g2p_count and
#              k_advect: positions/velocities read and written, keys; U, V,
#              saved U, saved V; solid SDF)
#   sdf        8 P + 6 C per substep (positions; fluid SDF + material)
#   pcg_a      17 B per cell and iteration (stencil code 1, z 4, old
#              direction 4; p 4, q = A' p 4)
#   pcg_b      32 B per cell and iteration (r -= alpha q in FP32, x += alpha p
#              in FP64)
#   mg_down0   5 C + 5 C1 per V-cycle (code, r; level-1 activity, r1)
#   mg_up0     9 C + 4 C1 per V-cycle (code, r, z; level-1 y)
#   mg_resid   k_mg_rr of every level l between level 1 and the coarse
#              block: 9 B per level-l cell (r, dinv; r_{l+1}) and the 100 B of
#              each stored row, per V-cycle
#   mg_pp      k_mg_pp of the same levels: 14 B per level-l cell (act, r,
#              dinv, y_{l+1}; y) and the 100 B of each stored row
#   nb_euler   NB-FLIP's Euler pass: U, V and the level set read once and
#              their advected copies written, 8 (F + C) per substep
#   visc_cg    the viscosity CG (U and V together): 216 B a cell and iteration,
#              timed over the whole iteration stretch with its polls
#   mg_coarse  per scene and V-cycle: 4 B per cell of the levels' dinv, of its
#              input and of its output, once each the 100 B of every row above
#              the bottom that is not its level's canonical interior stencil,
#              and the 100 B of every bottom row
KERNEL_BYTES = {
    "p2g": ("K_P2G", 1.0, "bytes", ["k_p2g"]),
    "particles": ("K_PARTICLES", 1.0, "bytes", ["k_advect", "k_g2p_count"]),
    "sdf": ("K_SDF", 1.0, "bytes", ["k_sdf_materials"]),
    "pcg_a": ("K_PCG_A", 17.0, "cells", ["k_pcg_a"]),
    "pcg_b": ("K_PCG_B", 32.0, "cells", ["k_pcg_b"]),
    "mg_down0": ("K_MG_DOWN0", 1.0, "bytes", ["k_mg_down0"]),
    "mg_up0": ("K_MG_UP0", 1.0, "bytes", ["k_mg_up0", "k_mg_up0_wide"]),
    "mg_resid": ("K_MG_RESID", 1.0, "bytes", ["k_mg_rr"]),
    "mg_pp": ("K_MG_PP", 1.0, "bytes", ["k_mg_pp"]),
    "mg_coarse": ("K_MG_COARSE", 1.0, "bytes", ["k_mg_coarse"]),
    "nb_euler": ("K_NB_EULER", 1.0, "bytes", ["k_nb_euler"]),
    "visc_cg": ("K_VISC", 1.0, "bytes", ["k_visc_spmv", "k_visc_update", "k_visc_dir", "k_visc_red"]),
}
STEP_FAMILIES = ["p2g", "particles", "sdf", "pcg_a", "pcg_b", "mg_down0", "mg_up0", "mg_resid", "mg_pp",
                 "mg_coarse"]
SOLVE_FAMILIES = ["pcg_a", "pcg_b", "mg_down0", "mg_up0", "mg_resid", "mg_pp", "mg_coarse"]
# Bytes moved past the L2 per algorithmic byte, per workload and family:
# rocprofv3's read requests by size (TCC_EA0_RDREQ_32B/64B/128B, times 32,
# 64 and 128 bytes) plus WRITE_SIZE, summed over every dispatch of the
# family's kernels in a run of that workload alone, divided by the library's
# algorithmic bytes over the same launches (profiles/hbm_traffic.md; on a
# 1 GiB copy kernel the same counters read 1.000 GiB each way).
MEASURED_TRAFFIC_PER_UNIT = {
    "mixed256": {"p2g": 1.367, "particles": 0.655, "sdf": 2.114, "pcg_a": 1.089, "pcg_b": 1.012,
                 "mg_down0": 1.063, "mg_up0": 0.948, "mg_resid": 1.306, "mg_pp": 1.143, "mg_coarse": 1.243},
    "dam1024": {"p2g": 1.107, "particles": 2.041, "sdf": 0.63, "pcg_a": 1.022, "pcg_b": 1.003,
                "mg_down0": 1.029, "mg_up0": 0.939, "mg_resid": 1.21, "mg_pp": 1.113, "mg_coarse": 1.513},
    "nb1024": {"p2g": 1.3, "particles": 0.261, "sdf": 2.155, "pcg_a": 1.022, "pcg_b": 1.003, "mg_down0": 1.018,
               "mg_up0": 0.939, "mg_resid": 1.211, "mg_pp": 1.114, "mg_coarse": 1.516, "nb_euler": 1.646},
    "visc256": {"p2g": 1.224, "particles": 1.704, "sdf": 1.285, "pcg_a": 1.054, "pcg_b": 1.005, "mg_down0": 1.024,
                "mg_up0": 0.937, "mg_resid": 1.203, "mg_pp": 1.093, "mg_coarse": 1.311, "visc_cg": 1.193},
    "proj4096": {"pcg_a": 1.863, "pcg_b": 1.0, "mg_down0": 1.006, "mg_up0": 1.983, "mg_resid": 1.259,
                 "mg_pp": 1.222, "mg_coarse": 1.386},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed frames of the batched workload")
    ap.add_argument("--warmup", type=int, default=3, help="untimed frames")
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--master-seed", type=int, default=2024)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--legs", default="batched,plain,dam,nb,visc,proj,served,cpu",
                    help="comma list of batched, plain, dam, nb, visc, proj, served, cpu (profiling runs pick one)")
    ap.add_argument("--kernel-frames", type=int, default=2,
                    help="frames (solves) each kernel family is timed over, after the timed ones (0: skip)")
    ap.add_argument("--skip-plain", action="store_true", help="batched leg only (profiling runs)")
    ap.add_argument("--lanes", type=int, default=0, help="lanes of the batched leg (0: the autotuner's choice)")
    ap.add_argument("--tune-rounds", type=int, default=1,
                    help="autotuner rounds (each steps 4 frames; 0: no tuning, --lanes or 1 lane)")
    ap.add_argument("--dam-frames", type=int, default=10, help="timed frames of the 1024^2 leg (0: skip)")
    ap.add_argument("--nb-frames", type=int, default=10, help="timed frames of the NB-FLIP 1024^2 leg (0: skip)")
    ap.add_argument("--visc-frames", type=int, default=5, help="timed frames of the viscous batch leg (0: skip)")
    ap.add_argument("--visc-scenes", type=int, default=32, help="256^2 scenes of the viscous batch leg")
    ap.add_argument("--visc", type=float, default=10.0, help="the viscous leg's fluid viscosity")
    ap.add_argument("--proj-res", type=int, default=4096, help="isolated projection grid (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="served-stream window (0: skip)")
    ap.add_argument("--serve-max-resident", type=int, default=4096,
                    help="resident-scene cap of the served leg (high, so that the HBM budget is the gate)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0, help="HBM budget of the served leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU-baseline budget per solver mode, seconds of timed frames (0: skip)")
    ap.add_argument("--cpu-frames", type=int, default=3, help="frames of each scene in the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cpu count)")
    return ap.parse_args()


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def pctl(v, q):
    if not v:
        return 0.0
    s = sorted(v)
    pos = q * (len(s) - 1)
    lo, hi = int(pos), min(len(s) - 1, int(pos) + 1)
    return s[lo] + (s[hi] - s[lo]) * (pos - lo)


class KernelTimer:
    """Per-launch HIP-event timing of one kernel family (events attached to
    each launch on the stream it runs on), accumulated over chosen regions."""

    def __init__(self, ctx, fg, name):
        self.ctx, self.fg, self.name = ctx, fg, name
        self.kid = getattr(fg, KERNEL_BYTES[name][0])
        self.per_unit = KERNEL_BYTES[name][1]
        self.unit_of_work = KERNEL_BYTES[name][2]
        self.ms = 0.0
        self.busy = 0.0
        self.launches = 0
        self.units = 0

    def __enter__(self):
        self.ctx.profile_kernel(self.kid)
        return self

    def __exit__(self, *a):
        t, b, n, u = self.ctx.profile_read()
        self.ctx.profile_kernel(-1)
        self.ms += t
        self.busy += b
        self.launches += n
        self.units += u

    def roofline(self, traffic_per_unit=None, copy_gbs=None, per=1):
        """per: the frames (solves) the timer covered, for the per-frame time."""
        achieved = self.units * self.per_unit / (self.ms * 1e-3) / 1e9 if self.ms > 0 else 0.0
        per_launch = self.units / max(self.launches, 1)
        return {
            "kernel": "+".join(KERNEL_BYTES[self.name][3]),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_per_unit * per_launch * self.per_unit if traffic_per_unit else None,
            "traffic_per_algorithmic_byte": traffic_per_unit,
            "avg_launch_us": 1e3 * self.ms / max(self.launches, 1),
            "launches": self.launches,
            "ms_per_frame": self.ms / max(per, 1),
            "algorithmic_bytes_per_launch": per_launch * self.per_unit,
            "bytes_per_unit": self.per_unit,
            "unit_of_work": self.unit_of_work,
            "measured_copy_gbs": copy_gbs,
            # launches of several lanes overlap: algorithmic bytes over the
            # time at least one launch of the kernel was running
            "busy_ms": self.busy,
            "busy_achieved": self.units * self.per_unit / (self.busy * 1e-3) / 1e9 if self.busy > 0 else 0.0,
        }


def kernel_table(fg, ctx, families, step, reps, workload, copy_gbs, where):
    """Each family timed with HIP events on its launches over `reps` further
    frames (solves) of the workload, one family at a time; returns the
    per-family rooflines and the dominant family (most GPU time per frame)."""
    table = {}
    if reps <= 0:
        return table, None
    for name in families:
        kt = KernelTimer(ctx, fg, name)
        with kt:
            for _ in range(reps):
                step()
        if kt.launches == 0 or kt.units == 0:   # the workload has no such launch
            continue
        table[name] = dict(kt.roofline(MEASURED_TRAFFIC_PER_UNIT[workload].get(name), copy_gbs, per=reps),
                           measured_on=f"{reps} {where} after the timed ones")
    dom = max(table, key=lambda k: table[k]["ms_per_frame"]) if table else None
    return table, dom


def work_totals(fg, ctx):
    """Algorithmic bytes of every reported family over the leg (all launches)."""
    w = ctx.work_read(reset=True)
    return {name: int(w[getattr(fg, kid)] * per) for name, (kid, per, _, _) in KERNEL_BYTES.items()}


def batched_leg(fg, ctx, a, copy_gbs):
    specs = fg.mixed_scene_specs(a.scenes, master_seed=a.master_seed)
    t0 = time.perf_counter()
    ctx.work_read(reset=True)
    batch = [fg.make_scene(ctx, sp, tol=a.tol, pcg_iter_limit=4000) for sp in specs]
    plain = [] if "plain" not in a.legs else [fg.make_scene(ctx, sp, tol=a.tol, pcg_iter_limit=4000)
                                              for sp in specs]
    n0 = sum(s.particle_count() for s in batch)
    cells = sum(sp["n"] ** 2 for sp in specs)
    log(f"{a.scenes} mixed scenes: {cells} cells, {n0} particles, set up in {time.perf_counter() - t0:.1f} s")
    for w in range(a.warmup):
        st = fg.batch_step_frame(ctx, batch)
        for sc in plain:
            sc.step_frame()
        log(f"warmup frame {w}: max substeps {max(s.substeps for s in st)}")
    # the lane schedule (size classes on separate streams) is picked by
    # measurement on live frames; the plain set advances the same frames
    if a.tune_rounds > 0:
        lanes, rates = fg.batch_autotune(ctx, batch, rounds=a.tune_rounds)
        for _ in range(a.tune_rounds * fg.MAX_LANES):
            for sc in plain:
                sc.step_frame()
        log(f"autotuned lanes: {lanes} (rates {[round(r / 1e6, 1) for r in rates]} M)")
    else:   # no tuning: the lane count given (default 1)
        lanes, rates = max(a.lanes, 1), []
        ctx.set_batch_lanes(lanes)
    # the solver tiling, tuned on a copy of the batch: a tiling sets the FP64
    # sums' order, and the timed batch and plain sets must see one order
    tiling, tiling_rates = None, []
    if a.tune_rounds > 0:
        copy = [fg.make_scene(ctx, sp, tol=a.tol, pcg_iter_limit=4000) for sp in specs]
        fg.batch_step_frame(ctx, copy)
        k, rates_t = fg.batch_autotune_tiling(ctx, copy, rounds=2 * a.tune_rounds)   # both orders
        for sc in copy:
            sc.close()
        tiling, tiling_rates = list(fg.TILING_CANDIDATES[k]), [float(r) for r in rates_t]
        log(f"autotuned solver tiling: {tiling} (rates {[round(r / 1e6, 1) for r in rates_t]} M)")
    if a.lanes > 0:   # a fixed lane count (profiling runs compare like with like)
        lanes = a.lanes
        ctx.set_batch_lanes(lanes)
        log(f"lanes fixed at {lanes}")
    el_b = el_p = 0.0
    ps_b = ps_p = 0
    lat_b, lat_p, frame_b, frame_p = [], [], [], []
    conv, resid, iters, subs = True, 0.0, 0, 0
    for k in range(a.steps):
        t0 = time.perf_counter()
        st = fg.batch_step_frame(ctx, batch)
        dt = time.perf_counter() - t0
        el_b += dt
        frame_b.append(dt * 1e3)
        for s in st:
            ps_b += s.particle_substeps
            lat_b.append(s.frame_ms)
            conv = conv and bool(s.converged)
            resid = max(resid, s.pressure_rel_resid)
            iters = max(iters, s.pressure_iters)
            subs += s.substeps
        if plain:
            t0 = time.perf_counter()
            for sc in plain:
                s = sc.step_frame()
                ps_p += s.particle_substeps
                lat_p.append(s.frame_ms)
            dt = time.perf_counter() - t0
            el_p += dt
            frame_p.append(dt * 1e3)
        log(f"frame {k}: batched {frame_b[-1]:.1f} ms" + (f", plain {frame_p[-1]:.1f} ms" if plain else ""))
    # the named kernels: HIP events attached to each of their launches on the
    # stream they run on, over further frames of the same batch (events on
    # every launch slow the launch pipeline by ~7 %, so the timed frames above
    # run without them), with the batch on one lane so that other lanes'
    # kernels do not stretch the intervals
    ctx.set_batch_lanes(1)
    table, dom = kernel_table(fg, ctx, STEP_FAMILIES, lambda: fg.batch_step_frame(ctx, batch), a.kernel_frames,
                              "mixed256", copy_gbs, "frames of the same batch on one lane")
    ctx.set_batch_lanes(lanes)
    for sc in batch + plain:
        sc.close()
    out = {
        "value": ps_b / el_b,
        "ms_per_step": 1e3 * el_b / a.steps,
        "particles": n0,
        "cells": cells,
        "substeps": subs,
        "p50_frame_ms": pctl(lat_b, 0.5),
        "p99_frame_ms": pctl(lat_b, 0.99),
        "all_converged": conv,
        "max_rel_residual": resid,
        "pcg_iters_per_substep_max": iters,
        "kernels": table,
        "dominant": dom,
        "algorithmic_bytes_total": work_totals(fg, ctx),
        "lanes": lanes,
        "autotune_rates": [float(r) for r in rates],
        "solver_tiling": tiling,
        "tiling_rates": tiling_rates,
    }
    if plain:
        assert ps_p == ps_b, "plain and batched sets diverged"
        out["plain"] = {
            "value": ps_p / el_p,
            "ms_per_step": 1e3 * el_p / a.steps,
            "p50_frame_ms": pctl(lat_p, 0.5),
            "p99_frame_ms": pctl(lat_p, 0.99),
        }
        out["batched_over_plain"] = out["value"] / out["plain"]["value"]
    return out


def dam_leg(fg, ctx, a, copy_gbs):
    res, ppc = 1024, 8
    ctx.work_read(reset=True)
    scene = fg.dam_break_scene(ctx, res, ppc=ppc, tol=a.tol, pcg_iter_limit=4000)
    n0 = scene.particle_count()
    for _ in range(a.warmup):
        scene.step_frame()
    lat, psub, subs, resid, conv = [], 0, 0, 0.0, True
    t0 = time.perf_counter()
    for _ in range(a.dam_frames):
        s = scene.step_frame()
        lat.append(s.frame_ms)
        psub += s.particle_substeps
        subs += s.substeps
        resid = max(resid, s.pressure_rel_resid)
        conv = conv and bool(s.converged)
    el = time.perf_counter() - t0
    table, dom = kernel_table(fg, ctx, STEP_FAMILIES, scene.step_frame, a.kernel_frames, "dam1024", copy_gbs,
                              "frames")
    scene.close()
    log(f"dam break {res}^2: {a.dam_frames} frames, {subs} substeps, {el:.2f} s")
    return {
        "workload": f"plain single-scene dam break {res}x{res}, {ppc} ppc ({n0} particles)",
        "value": psub / el,
        "ms_per_step": 1e3 * el / a.dam_frames,
        "frames": a.dam_frames,
        "substeps": subs,
        "p50_frame_ms": pctl(lat, 0.5),
        "p99_frame_ms": pctl(lat, 0.99),
        "max_rel_residual": resid,
        "all_converged": conv,
        "roofline": table.get(dom),
        "kernels": table,
        "algorithmic_bytes_total": work_totals(fg, ctx),
    }


def nb_cpu_sample(scene, res, ppc, dt, a):
    """One NB-FLIP substep of the oracle from the GPU's state after the timed
    frames (same scene, same dt), its converged FP64 solve to the same
    tolerance on `threads` threads."""
    import oracle as orc

    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
    g = scene.get_state()
    p = orc.make_params(res, ppc=ppc, solver=orc.SOLVER_CONVERGED, threads=threads, tol=a.tol, sim=orc.SIM_NBFLIP)
    n = len(g["px"])
    st = orc.State(res, res, 2 * res * res * ppc)
    for k in ("px", "py", "vx", "vy", "pid"):
        getattr(st, k)[:n] = g[k]
    for k in ("u", "v", "u_valid", "v_valid", "material", "solid_sdf", "fluid_sdf", "solid_id", "emitter_id",
              "particle_counts"):
        getattr(st, k)[...] = g[k]
    st.n = n
    st.c.rng_epoch, st.c.next_pid = int(g["rng_epoch"]), int(g["next_pid"])
    t0 = time.perf_counter()
    orc.lib().orc_nb_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "particle-substeps/s", "ms_per_substep": 1e3 * el, "cores": threads,
            "kind": "port", "sample": f"one NB-FLIP substep of the {res}x{res} scene ({n} particles) from the "
                                      "GPU's state after the timed frames, dt of the last GPU substep"}


def nb_leg(fg, ctx, a, copy_gbs):
    """NBFlipSolver (nbflipsolver.cpp) on the 1024^2 dam break, 8 ppc: the
    first frame seeds no particle, later frames keep particles in the band
    below the surface (the timed frames follow --warmup frames)."""
    res, ppc = 1024, 8
    ctx.work_read(reset=True)
    scene = fg.nb_dam_break_scene(ctx, res, ppc=ppc, tol=a.tol, pcg_iter_limit=4000)
    for _ in range(a.warmup):
        scene.step_frame()
    lat, psub, subs, resid, conv, last_dt = [], 0, 0, 0.0, True, 0.0
    t0 = time.perf_counter()
    for _ in range(a.nb_frames):
        s = scene.step_frame()
        lat.append(s.frame_ms)
        psub += s.particle_substeps
        subs += s.substeps
        resid = max(resid, s.pressure_rel_resid)
        conv = conv and bool(s.converged)
        last_dt = s.last_dt
    el = time.perf_counter() - t0
    n_end = scene.particle_count()
    cpu = nb_cpu_sample(scene, res, ppc, last_dt, a) if "cpu" in a.legs and a.cpu_seconds > 0 else None
    table, dom = kernel_table(fg, ctx, STEP_FAMILIES + ["nb_euler"], scene.step_frame, a.kernel_frames, "nb1024",
                              copy_gbs, "frames")
    scene.close()
    log(f"NB-FLIP dam break {res}^2: {a.nb_frames} frames, {subs} substeps, {el:.2f} s")
    return {
        "workload": f"plain single-scene NB-FLIP dam break {res}x{res}, {ppc} ppc ({n_end} particles after the "
                    f"timed frames: the band below the surface)",
        "value": psub / el,
        "cell_substeps_per_s": res * res * subs / el,
        "ms_per_step": 1e3 * el / a.nb_frames,
        "ms_per_substep": 1e3 * el / max(subs, 1),
        "frames": a.nb_frames,
        "substeps": subs,
        "p50_frame_ms": pctl(lat, 0.5),
        "p99_frame_ms": pctl(lat, 0.99),
        "max_rel_residual": resid,
        "all_converged": conv,
        "roofline": table.get(dom),
        "kernels": table,
        "algorithmic_bytes_total": work_totals(fg, ctx),
        "cpu_sample": cpu,
    }


def visc_cpu_sample(scene, res, ppc, dt, a):
    """One viscous substep of the oracle from the GPU's state after the
    timed frames (same scene, same dt; converged FP64 solves)."""
    import oracle as orc

    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
    g = scene.get_state()
    p = orc.make_params(res, ppc=ppc, solver=orc.SOLVER_CONVERGED, threads=threads, tol=a.tol, viscosity=True)
    n = len(g["px"])
    st = orc.State(res, res, res * res * ppc)
    for k in ("px", "py", "vx", "vy", "pid"):
        getattr(st, k)[:n] = g[k]
    st.pvisc[:n] = g["particle_viscosity"]
    for k in ("u", "v", "u_valid", "v_valid", "material", "solid_sdf", "fluid_sdf", "solid_id", "emitter_id",
              "particle_counts"):
        getattr(st, k)[...] = g[k]
    st.n = n
    st.c.rng_epoch, st.c.next_pid = int(g["rng_epoch"]), int(g["next_pid"])
    t0 = time.perf_counter()
    orc.lib().orc_substep(C.byref(p), st.ref(), C.c_float(dt), C.byref(orc.OrcStats()))
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "particle-substeps/s", "ms_per_substep": 1e3 * el, "cores": threads,
            "kind": "port", "sample": f"one viscous substep of the {res}x{res} scene ({n} particles) from the GPU's "
                                      "state after the timed frames, dt of the last GPU substep"}


def visc_leg(fg, ctx, a, copy_gbs):
    """Viscous dam breaks (viscosityEnabled: the light model and a second
    projection every substep), a batch of a.visc_scenes 256^2 scenes with 8
    ppc and fluid rectangles of varied size, stepped together by
    fg_batch_step_frame.  (A single 1024^2 viscous dam break runs away from
    its first substep from rest, on the oracle exactly as here: the light
    model's rows of air cells beside the liquid are not diagonally dominant
    once 4 nu dt > 1 and the second projection reuses the substep's stale
    matrix; at 256^2 the flow stays bounded.)"""
    import numpy as np

    res, ppc = 256, 8
    rng = np.random.default_rng(a.master_seed)
    ctx.work_read(reset=True)
    scenes = []
    walls, _ = fg.dam_break_geometry(res)
    for k in range(a.visc_scenes):
        x1, y1 = rng.uniform(0.4, 0.9), rng.uniform(0.25, 0.6)
        _, fluid = fg.dam_break_geometry(res, (x1 - rng.uniform(0.2, 0.4), x1, 0.0, y1))
        sc = fg.Scene(ctx, fg.make_params(res, ppc=ppc, tol=a.tol, pcg_iter_limit=4000, viscosity=True))
        sc.init_objects(solids=walls, fluids=fluid, friction=np.zeros(len(walls), np.float32),
                        fluid_viscosity=[a.visc], seed=k)
        scenes.append(sc)
    n0 = sum(sc.particle_count() for sc in scenes)
    for _ in range(a.warmup):
        fg.batch_step_frame(ctx, scenes)
    lat, psub, subs, resid, conv, vit, vmax = [], 0, 0, 0.0, True, 0, 0.0
    t0 = time.perf_counter()
    for _ in range(a.visc_frames):
        tf = time.perf_counter()
        st = fg.batch_step_frame(ctx, scenes)
        lat.append(1e3 * (time.perf_counter() - tf))
        for s in st:
            psub += s.particle_substeps
            subs = max(subs, s.substeps)
            resid = max(resid, s.pressure_rel_resid)
            conv = conv and bool(s.converged)
            vit = max(vit, s.viscosity_iters)
            vmax = max(vmax, s.max_velocity)
    el = time.perf_counter() - t0
    last_dt = st[0].last_dt
    cpu = visc_cpu_sample(scenes[0], res, ppc, last_dt, a) if "cpu" in a.legs and a.cpu_seconds > 0 else None
    table, dom = kernel_table(fg, ctx, STEP_FAMILIES + ["visc_cg"], lambda: fg.batch_step_frame(ctx, scenes),
                              a.kernel_frames, "visc256", copy_gbs, "frames")
    for sc in scenes:
        sc.close()
    log(f"viscous batch {a.visc_scenes} x {res}^2: {a.visc_frames} frames, {el:.2f} s")
    return {
        "workload": f"batched: {a.visc_scenes} viscous dam-break scenes {res}x{res}, {ppc} ppc ({n0} particles), "
                    f"fluid viscosity {a.visc:g}: the light model + a second projection every substep",
        "value": psub / el,
        "ms_per_step": 1e3 * el / a.visc_frames,
        "frames": a.visc_frames,
        "max_substeps_per_frame": subs,
        "viscosity_iters_max": vit,
        "max_particle_speed": vmax,
        "p50_frame_ms": pctl(lat, 0.5),
        "p99_frame_ms": pctl(lat, 0.99),
        "max_rel_residual": resid,
        "all_converged": conv,
        "roofline": table.get(dom),
        "kernels": table,
        "algorithmic_bytes_total": work_totals(fg, ctx),
        "cpu_sample": cpu,
    }


def projection_leg(fg, ctx, a, copy_gbs):
    n = a.proj_res
    mat, _, _, b, scale = fg.isolated_projection_system(n, 1)
    ctx.work_read(reset=True)
    ctx.poisson_solve(mat, scale, b, tol=a.tol)          # warm-up (allocation, code load)
    x, st = ctx.poisson_solve(mat, scale, b, tol=a.tol)
    table, dom = kernel_table(fg, ctx, SOLVE_FAMILIES, lambda: ctx.poisson_solve(mat, scale, b, tol=a.tol),
                              max(1, a.kernel_frames), "proj4096", copy_gbs, "solves of the same system")
    liquid = int(((mat & fg.FLUID) != 0).sum())
    log(f"isolated projection {n}^2: {st.iterations} iterations, {st.solve_ms:.1f} ms, resid {st.rel_residual:.2e}")
    return {
        "workload": f"isolated pressure projection {n}x{n}: smoothed-noise liquid mask ({liquid} unknowns), "
                    "one-cell solid border + 2% random solids, rhs = divergence of U[-1,1] faces",
        "iterations": st.iterations,
        "sweeps": st.sweeps,
        "rel_residual": st.rel_residual,
        "converged": bool(st.converged),
        "solve_ms_incl_rhs_upload": st.solve_ms,
        "roofline": table.get(dom),
        "kernels": table,
        "algorithmic_bytes_total": work_totals(fg, ctx),
    }


def served_leg(fg, ctx, a):
    rate, budget = 200.0, int(a.serve_budget_gb * (1 << 30))
    reqs = fg.served_request_specs(rate=rate, duration=a.serve_seconds)
    srv = fg.Server(ctx, hbm_budget=budget, max_resident=a.serve_max_resident)
    t0 = time.perf_counter()
    nxt, psub, ticks = 0, 0, 0
    while True:
        now = time.perf_counter() - t0
        if now >= a.serve_seconds:
            break
        while nxt < len(reqs) and reqs[nxt]["t"] <= now:
            srv.submit_spec(reqs[nxt], reqs[nxt]["frames"], tol=a.tol, pcg_iter_limit=4000)
            nxt += 1
        t = srv.tick()
        ticks += 1
        psub += t.particle_substeps
        if t.resident == 0 and t.queued == 0 and nxt < len(reqs):
            time.sleep(max(0.0, min(reqs[nxt]["t"] - (time.perf_counter() - t0), 0.01)))
    el = time.perf_counter() - t0
    s = srv.summary()
    srv.close()
    log(f"served {a.serve_seconds:.0f} s: {s.frames} scene frames, {psub / el / 1e6:.2f} M particle-substeps/s")
    return {
        "workload": f"Poisson arrivals {rate:g}/s, a {a.serve_seconds:g} s window of the stream (the full stream "
                    "runs 120 s), resolution log-uniform [64, 1024], 4-8 ppc, 30-300 frames per request, "
                    f"{a.serve_budget_gb:g} GiB HBM budget, at most {a.serve_max_resident} resident scenes",
        "window_s": a.serve_seconds,
        "budget_gb": budget / 1e9,
        "budget_waits": s.budget_waits,
        "cap_waits": s.cap_waits,
        "value": psub / el,
        "unit": "particle-substeps/s",
        "p50_frame_ms": s.p50_frame_ms,
        "p99_frame_ms": s.p99_frame_ms,
        "p50_admit_wait_ms": s.p50_admit_wait_ms,
        "p99_admit_wait_ms": s.p99_admit_wait_ms,
        "submitted": s.submitted,
        "completed": s.completed,
        "scene_frames": s.frames,
        "peak_resident": s.peak_resident,
        "peak_resident_gb": s.peak_resident_bytes / 1e9,
        "lanes": ctx.batch_lanes(),
    }


def cpu_baseline(fg, a):
    """The CPU restatement of the reference path (the oracle) on the same
    mixed scenes, in both solver modes:
      ref_replica     the reference's solver (its PCG with its IPP, max-abs
                      1e-2, cap 200, gated density solve);
      converged_fp64  PCG in FP64 to relative residual 1e-6, the GPU path's
                      accuracy.
    Each scene runs single-threaded, the configuration that matches the
    reference's own build: 128^2, 4 ppc, 10 frames ran at 1.44e5
    particle-substeps/s in replica mode where the reference's single-thread
    build ran at about 1.4e5 (same 8-core host; the reference's 8-thread
    pool was slower there, 4.0e4, and this restatement's pool faster, 7.0e4,
    so neither pool stands in for the other).  Independent scenes step side
    by side, `threads` at a time (one per core), frames 0 .. cpu_frames-1 of
    each; set-up is not timed.  Batches of scenes run until the mode's time
    budget is spent."""
    import concurrent.futures as cf

    import oracle as orc

    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
    specs = fg.mixed_scene_specs(a.scenes, master_seed=a.master_seed)
    L = orc.lib()

    def setup(sp, mode):
        solver, tol = (orc.SOLVER_REF_REPLICA, 1e-2) if mode == "ref_replica" else (orc.SOLVER_CONVERGED, a.tol)
        p = orc.make_params(sp["n"], ppc=sp["ppc"], gravity=sp["gravity"], solver=solver,
                            density_mode=orc.DENSITY_GATED, threads=1, tol=tol)
        walls, fluids = orc.dam_break_geometry(sp["n"], sp["fluid_box"])
        return p, orc.init_scene(p, walls, fluids, seed=sp["seed"])

    def frames(job):
        p, st = job
        frame_dt = 1.0 / p.fps
        psub = 0
        for _ in range(a.cpu_frames):
            t_frame, count, fin = 0.0, 0, C.c_int32(0)
            while not fin.value:
                vel = L.orc_max_particle_velocity(st.ref())
                dt = L.orc_frame_dt(vel, p.cfl_number, frame_dt, t_frame, count, p.max_substeps, C.byref(fin))
                psub += st.n
                L.orc_substep(C.byref(p), st.ref(), dt, C.byref(orc.OrcStats()))
                t_frame += dt
                count += 1
        return psub

    out = {}
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        for mode in ("ref_replica", "converged_fp64"):
            psub, el, done = 0, 0.0, 0
            while done < len(specs) and el < a.cpu_seconds:
                batch = specs[done:done + threads]
                jobs = list(ex.map(lambda sp: setup(sp, mode), batch))
                t0 = time.perf_counter()
                psub += sum(ex.map(frames, jobs))
                el += time.perf_counter() - t0
                done += len(batch)
            out[mode] = {"value": psub / el, "scenes": done, "seconds": el}
            log(f"cpu baseline {mode}: {psub / el:.3e} particle-substeps/s ({done} scenes, {el:.1f} s)")
    rep, conv = out["ref_replica"], out["converged_fp64"]
    return {
        "value": rep["value"],
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"frames 0-{a.cpu_frames - 1} of the first {rep['scenes']} of the {a.scenes} mixed scenes, CPU "
                  f"restatement in reference-replica mode (reference PCG/IPP, cap 200, max-abs 1e-2, gated density "
                  f"solve), one thread per scene, {threads} scenes at a time; {rep['seconds']:.1f} s of frames",
        "converged_fp64": {
            "value": conv["value"],
            "unit": "particle-substeps/s",
            "sample": f"the same with FP64 PCG to relative residual {a.tol:g} (the GPU path's accuracy): frames "
                      f"0-{a.cpu_frames - 1} of {conv['scenes']} scenes; {conv['seconds']:.1f} s of frames",
        },
        "calibration": "replica mode, one thread, 128^2 dam break, 4 ppc, 10 frames: 1.44e5 particle-substeps/s "
                       "vs the reference's single-thread build at about 1.4e5 on the same 8-core host",
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: scenes are independent and "
              "would shard whole across GPUs, which this library does not do", file=sys.stderr)
        sys.exit(2)
    import flipsolver2d_amd as fg

    a.legs = set(a.legs.split(","))
    if a.skip_plain:
        a.legs.discard("plain")
    b = dam = nb = visc = proj = served = None
    if a.legs & {"batched", "plain", "dam", "nb", "visc", "proj", "served"}:
        ctx = fg.Context(0)
        copy_gbs, _ = ctx.bandwidth_probe(1 << 30, 5)
        b = batched_leg(fg, ctx, a, copy_gbs) if "batched" in a.legs else None
        dam = dam_leg(fg, ctx, a, copy_gbs) if a.dam_frames > 0 and "dam" in a.legs else None
        nb = nb_leg(fg, ctx, a, copy_gbs) if a.nb_frames > 0 and "nb" in a.legs else None
        visc = visc_leg(fg, ctx, a, copy_gbs) if a.visc_frames > 0 and "visc" in a.legs else None
        proj = projection_leg(fg, ctx, a, copy_gbs) if a.proj_res > 0 and "proj" in a.legs else None
        served = served_leg(fg, ctx, a) if a.serve_seconds > 0 and "served" in a.legs else None
        ctx.close()
    cpu = cpu_baseline(fg, a) if a.cpu_seconds > 0 and "cpu" in a.legs else None
    if cpu:
        log(f"cpu baseline: {cpu['value']:.3e} particle-substeps/s")
    if b is None:   # a profiling run of other legs: their results alone
        print(json.dumps({"dam_break_1024": dam, "nbflip_dam_break_1024": nb, "viscous_batch_256": visc,
                          "isolated_projection": proj, "served": served, "cpu_baseline": cpu}), flush=True)
        return

    dom = b.pop("dominant")
    out = {
        "metric": METRIC,
        "value": b["value"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched: {a.scenes} independent dam-break scenes, resolution from {{64,96,128,192,256}}, "
                        f"4-8 ppc ({b['particles']} particles, {b['cells']} cells), FP32 state, FP64 PCG "
                        f"reductions + refinement to rel resid {a.tol:g}; one step = one frame of every scene",
            "scenes": a.scenes, "master_seed": a.master_seed, "solver_tol": a.tol,
            "substeps": b["substeps"], "lanes": b["lanes"],
            "autotune_rates": b["autotune_rates"],
            "solver_tiling": b["solver_tiling"], "tiling_candidates": [list(c) for c in fg.TILING_CANDIDATES],
            "tiling_rates": b["tiling_rates"],
        },
        "p50_frame_ms": b["p50_frame_ms"],
        "p99_frame_ms": b["p99_frame_ms"],
        "all_converged": b["all_converged"],
        "max_rel_residual": b["max_rel_residual"],
        "pcg_iters_per_substep_max": b["pcg_iters_per_substep_max"],
        "plain": b.get("plain"),
        "batched_over_plain": b.get("batched_over_plain"),
        "roofline": b["kernels"].get(dom),
        "kernels": b["kernels"],
        "algorithmic_bytes_total": b["algorithmic_bytes_total"],
        "cpu_baseline": cpu,
        "dam_break_1024": dam,
        "nbflip_dam_break_1024": nb,
        "viscous_batch_256": visc,
        "isolated_projection": proj,
        "served": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
