"""flipsolver2d_amd -- MI355X-native FLIP substep of FlipSolver2d behind a C-ABI.

This module is a thin ctypes binding of libflipgpu.so (include/flipgpu.h) used
by the tests and bench.py.  The C++ facade with the reference's class names
lives in flipsolver2d_amd/cpp/.  There is no CPU fallback: if the HIP library
is missing or no GPU is visible, every compute entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libflipgpu.so")

FG_OK = 0
FLUID, SOURCE, SOLID, SINK, EMPTY = 0x40, 0x41, 0x20, 0x12, 0x10
DENSITY_OFF, DENSITY_GATED, DENSITY_ON = 0, 1, 2
SIM_FLIP, SIM_NBFLIP = 0, 1
(STAGE_ADVECT, STAGE_REBIN, STAGE_DENSITY, STAGE_P2G, STAGE_SDF, STAGE_LEVELSET, STAGE_EXTRAPOLATE,
 STAGE_FORCES, STAGE_PROJECT, STAGE_SOLIDS, STAGE_G2P, STAGE_COUNT) = range(12)
STAGE_NAMES = ["advect", "rebin", "density", "p2g", "sdf", "levelset", "extrapolate", "forces", "project",
               "solids", "g2p", "count"]

EXPORTS = [
    "fg_last_error", "fg_abi_version", "fg_ctx_create", "fg_ctx_destroy", "fg_ctx_bytes_in_use",
    "fg_ctx_set_timing", "fg_scene_bytes", "fg_scene_create", "fg_scene_destroy", "fg_scene_init_geometry",
    "fg_scene_set_friction", "fg_scene_set_state", "fg_scene_get_state", "fg_scene_particle_count",
    "fg_scene_substep", "fg_scene_step_frame", "fg_scene_stage", "fg_batch_step_frame", "fg_poisson_solve",
    "fg_ctx_profile_kernel", "fg_ctx_profile_read", "fg_ctx_bandwidth_probe",
    "fg_server_create", "fg_server_destroy", "fg_server_submit", "fg_server_tick", "fg_server_status",
    "fg_server_scene", "fg_server_release", "fg_server_summary_get", "fg_ctx_set_batch_lanes",
    "fg_ctx_batch_lanes", "fg_batch_autotune", "fg_ctx_set_solver_tiling", "fg_ctx_set_preconditioner",
    "fg_ctx_preconditioner", "fg_batch_substep", "fg_server_cancel", "fg_ctx_work_read", "fg_scene_init_objects", "fg_scene_set_emitters",
    "fg_scene_max_velocity", "fg_poisson_solve_matrix",
    "fg_poisson_solve_noalloc", "fg_ctx_set_solver_graphs", "fg_batch_autotune_tiling",
    "fg_ctx_set_vcycle_fused",
]
MAX_LANES = 4
REQ_QUEUED, REQ_RESIDENT, REQ_DONE, REQ_REJECTED, REQ_RELEASED, REQ_FAILED, REQ_CANCELLED = range(7)
E_INVALID, E_DEVICE, E_BUDGET, E_CAPACITY, E_STATE, E_NUMERIC, E_CANCELLED = -1, -2, -3, -4, -5, -6, -7
(K_ADVECT, K_SORT, K_P2G, K_SDF, K_LEVELSET, K_EXTRAPOLATE, K_PCG_A, K_PCG_B, K_G2P, K_RHS, K_APPLY,
 K_VCYCLE, K_MG_RESID, K_MG_COARSE, K_MG_DOWN0, K_MG_UP0, K_PARTICLES, K_MG_PP, K_NB_EULER, K_VISC) = range(20)
K_NUM = 20
PRECOND_MG, PRECOND_IPP = 0, 1


class FgError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class SceneParams(C.Structure):
    _fields_ = [
        ("size_i", C.c_int32), ("size_j", C.c_int32),
        ("dx", C.c_double), ("fluid_density", C.c_double),
        ("gravity_i", C.c_float), ("gravity_j", C.c_float),
        ("pic_ratio", C.c_float), ("cfl_number", C.c_float), ("particle_scale", C.c_float),
        ("fps", C.c_int32), ("max_substeps", C.c_int32), ("particles_per_cell", C.c_int32),
        ("pcg_iter_limit", C.c_int32), ("density_mode", C.c_int32),
        ("solver_tolerance", C.c_double), ("max_particles", C.c_int64),
        ("sim_method", C.c_int32), ("viscosity_enabled", C.c_int32),
    ]


class SceneState(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int64),
        ("px", C.c_void_p), ("py", C.c_void_p), ("vx", C.c_void_p), ("vy", C.c_void_p), ("pid", C.c_void_p),
        ("u", C.c_void_p), ("v", C.c_void_p), ("u_valid", C.c_void_p), ("v_valid", C.c_void_p),
        ("material", C.c_void_p), ("solid_sdf", C.c_void_p), ("fluid_sdf", C.c_void_p),
        ("solid_id", C.c_void_p), ("particle_counts", C.c_void_p),
        ("saved_u", C.c_void_p), ("saved_v", C.c_void_p), ("pressure", C.c_void_p), ("rhs", C.c_void_p),
        ("emitter_id", C.c_void_p), ("rng_epoch", C.c_uint32), ("next_pid", C.c_uint32),
        ("stencil", C.c_void_p), ("work_u", C.c_void_p), ("work_v", C.c_void_p),
        ("particle_viscosity", C.c_void_p), ("viscosity", C.c_void_p),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("substeps", C.c_int32), ("pressure_iters", C.c_int32), ("pressure_sweeps", C.c_int32),
        ("density_iters", C.c_int32), ("density_applied", C.c_int32), ("converged", C.c_int32),
        ("pressure_rel_resid", C.c_double),
        ("particles", C.c_int64), ("particle_substeps", C.c_int64),
        ("particles_killed", C.c_int64), ("particles_erased", C.c_int64),
        ("last_dt", C.c_float), ("max_velocity", C.c_float),
        ("stage_ms", C.c_double * 12), ("frame_ms", C.c_double),
        ("status", C.c_int32), ("viscosity_iters", C.c_int32), ("particles_reseeded", C.c_int64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "stage_ms"}
        d["stage_ms"] = {STAGE_NAMES[i]: self.stage_ms[i] for i in range(12)}
        return d


class Polygons(C.Structure):
    _fields_ = [("n", C.c_int32), ("nverts", C.c_void_p), ("verts", C.c_void_p)]


class SceneObjects(C.Structure):
    _fields_ = [("solids", Polygons), ("friction", C.c_void_p), ("sources", Polygons),
                ("source_velocity", C.c_void_p), ("source_transfer", C.c_void_p), ("sinks", Polygons),
                ("fluids", Polygons), ("seed", C.c_uint32), ("source_viscosity", C.c_void_p),
                ("fluid_viscosity", C.c_void_p)]


class Request(C.Structure):
    _fields_ = [("params", SceneParams), ("n_solids", C.c_int32), ("solid_nverts", C.c_void_p),
                ("solid_verts", C.c_void_p), ("friction", C.c_void_p), ("n_fluids", C.c_int32),
                ("fluid_nverts", C.c_void_p), ("fluid_verts", C.c_void_p), ("seed", C.c_uint32),
                ("frames", C.c_int32), ("keep", C.c_int32), ("fluid_viscosity", C.c_void_p)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("frames_done", C.c_int32), ("particles", C.c_int64),
                ("particle_substeps", C.c_int64), ("submit_ms", C.c_double), ("admit_ms", C.c_double),
                ("finish_ms", C.c_double), ("max_frame_ms", C.c_double), ("error", C.c_int32),
                ("reserved", C.c_int32)]


class TickStats(C.Structure):
    _fields_ = [("admitted", C.c_int32), ("stepped", C.c_int32), ("retired", C.c_int32), ("queued", C.c_int32),
                ("resident", C.c_int32), ("particle_substeps", C.c_int64), ("tick_ms", C.c_double),
                ("admit_ms", C.c_double), ("frame_ms_max", C.c_double), ("resident_bytes", C.c_uint64)]


class ServerSummary(C.Structure):
    _fields_ = [("frames", C.c_int64), ("particle_substeps", C.c_int64), ("busy_ms", C.c_double),
                ("p50_frame_ms", C.c_double), ("p99_frame_ms", C.c_double), ("max_frame_ms", C.c_double),
                ("p50_admit_wait_ms", C.c_double), ("p99_admit_wait_ms", C.c_double),
                ("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("peak_resident_bytes", C.c_uint64), ("peak_resident", C.c_int32), ("reserved", C.c_int32),
                ("failed", C.c_int64), ("cancelled", C.c_int64), ("budget_waits", C.c_int64),
                ("cap_waits", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("sweeps", C.c_int32), ("converged", C.c_int32),
                ("rel_residual", C.c_double), ("solve_ms", C.c_double)]


_lib = None


def lib():
    """Load libflipgpu.so (built by flipsolver2d_amd.build); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FgError(f"{LIB_PATH} is missing: run `python -m flipsolver2d_amd.build` first")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f32, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double
        sig = {
            "fg_last_error": (C.c_char_p, []),
            "fg_abi_version": (i32, []),
            "fg_ctx_create": (C.c_int, [i32, C.c_uint64, C.POINTER(vp)]),
            "fg_ctx_destroy": (C.c_int, [vp]),
            "fg_ctx_bytes_in_use": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
            "fg_ctx_set_timing": (C.c_int, [vp, i32]),
            "fg_scene_bytes": (C.c_int, [C.POINTER(SceneParams), C.POINTER(C.c_uint64)]),
            "fg_scene_create": (C.c_int, [vp, C.POINTER(SceneParams), C.POINTER(vp)]),
            "fg_scene_destroy": (C.c_int, [vp]),
            "fg_scene_init_geometry": (C.c_int, [vp, i32, vp, vp, vp, i32, vp, vp, C.c_uint32]),
            "fg_scene_set_friction": (C.c_int, [vp, i32, vp]),
            "fg_scene_init_objects": (C.c_int, [vp, C.POINTER(SceneObjects)]),
            "fg_scene_set_emitters": (C.c_int, [vp, i32, vp, vp, C.c_uint32]),
            "fg_scene_set_state": (C.c_int, [vp, C.POINTER(SceneState)]),
            "fg_scene_get_state": (C.c_int, [vp, C.POINTER(SceneState)]),
            "fg_scene_particle_count": (C.c_int, [vp, C.POINTER(i64)]),
            "fg_scene_max_velocity": (C.c_int, [vp, C.POINTER(f32)]),
            "fg_poisson_solve_matrix": (C.c_int, [vp, i32, i32, vp, f64, vp, vp, f64, i32, C.POINTER(SolveStats)]),
            "fg_scene_substep": (C.c_int, [vp, f32, C.POINTER(Stats)]),
            "fg_scene_step_frame": (C.c_int, [vp, C.POINTER(Stats)]),
            "fg_scene_stage": (C.c_int, [vp, i32, f32]),
            "fg_batch_step_frame": (C.c_int, [vp, vp, i32, vp]),
            "fg_batch_substep": (C.c_int, [vp, vp, i32, vp, vp]),
            "fg_poisson_solve": (C.c_int, [vp, i32, i32, vp, f64, vp, vp, f64, i32, C.POINTER(SolveStats)]),
            "fg_poisson_solve_noalloc": (C.c_int, [vp, i32, i32, vp, f64, vp, vp, f64, i32, C.POINTER(SolveStats)]),
            "fg_ctx_profile_kernel": (C.c_int, [vp, i32]),
            "fg_ctx_work_read": (C.c_int, [vp, vp, i32]),
            "fg_ctx_profile_read": (C.c_int, [vp, C.POINTER(f64), C.POINTER(f64), C.POINTER(i64), C.POINTER(i64)]),
            "fg_ctx_bandwidth_probe": (C.c_int, [vp, C.c_uint64, i32, C.POINTER(f64), C.POINTER(f64)]),
            "fg_ctx_set_batch_lanes": (C.c_int, [vp, i32]),
            "fg_ctx_batch_lanes": (C.c_int, [vp, C.POINTER(i32)]),
            "fg_ctx_set_solver_tiling": (C.c_int, [vp, i32, i32]),
            "fg_ctx_set_preconditioner": (C.c_int, [vp, i32]),
            "fg_ctx_set_solver_graphs": (C.c_int, [vp, i32]),
            "fg_ctx_set_vcycle_fused": (C.c_int, [vp, i32]),
            "fg_ctx_preconditioner": (C.c_int, [vp, C.POINTER(i32)]),
            "fg_batch_autotune": (C.c_int, [vp, vp, i32, i32, C.POINTER(i32), vp]),
            "fg_batch_autotune_tiling": (C.c_int, [vp, vp, i32, i32, vp, i32, C.POINTER(i32), vp]),
            "fg_server_create": (C.c_int, [vp, C.c_uint64, i32, C.POINTER(vp)]),
            "fg_server_destroy": (C.c_int, [vp]),
            "fg_server_submit": (C.c_int, [vp, C.POINTER(Request), C.POINTER(i64)]),
            "fg_server_tick": (C.c_int, [vp, C.POINTER(TickStats)]),
            "fg_server_status": (C.c_int, [vp, i64, C.POINTER(RequestStatus)]),
            "fg_server_scene": (C.c_int, [vp, i64, C.POINTER(vp)]),
            "fg_server_release": (C.c_int, [vp, i64]),
            "fg_server_cancel": (C.c_int, [vp, i64]),
            "fg_server_summary_get": (C.c_int, [vp, C.POINTER(ServerSummary)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != FG_OK:
        raise FgError(f"flipgpu error {rc}: {lib().fg_last_error().decode(errors='replace')}", rc)


def _ptr(a):
    return None if a is None else a.ctypes.data


def make_params(n_i, n_j=None, *, dx=None, ppc=4, density=1000.0, gravity=(9.81, 0.0), pic=0.03, cfl=5.0,
                particle_scale=0.8, fps=30, max_substeps=10, pcg_iter_limit=2000, density_mode=DENSITY_OFF,
                tol=1e-6, max_particles=None, sim=SIM_FLIP, viscosity=False):
    """FlipSolverParameters as fg_scene_params; sim: SIM_FLIP (FlipSolver) or
    SIM_NBFLIP (NBFlipSolver, whose band cells fill to ppc and crowded cells
    keep up to 2 ppc: its default capacity is twice FLIP's); viscosity:
    viscosityEnabled (the light model, FlipSolver only)."""
    n_j = n_i if n_j is None else n_j
    p = SceneParams()
    p.size_i, p.size_j = n_i, n_j
    p.dx = (1.0 / max(n_i, n_j)) if dx is None else dx
    p.fluid_density = density
    p.gravity_i, p.gravity_j = gravity
    p.pic_ratio, p.cfl_number, p.particle_scale = pic, cfl, particle_scale
    p.fps, p.max_substeps, p.particles_per_cell = fps, max_substeps, ppc
    p.pcg_iter_limit, p.density_mode, p.solver_tolerance = pcg_iter_limit, density_mode, tol
    p.max_particles = n_i * n_j * ppc * (2 if sim == SIM_NBFLIP else 1) if max_particles is None else max_particles
    p.sim_method = sim
    p.viscosity_enabled = 1 if viscosity else 0
    return p


class Context:
    def __init__(self, device=0, hbm_budget=0):
        self.h = C.c_void_p()
        _check(lib().fg_ctx_create(device, hbm_budget, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().fg_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def set_timing(self, on=True):
        _check(lib().fg_ctx_set_timing(self.h, 1 if on else 0))

    def set_batch_lanes(self, lanes):
        """Split batched frames over `lanes` streams (size classes), one host thread each."""
        _check(lib().fg_ctx_set_batch_lanes(self.h, lanes))

    def set_solver_tiling(self, target_blocks=1024, min_chunk=1024):
        _check(lib().fg_ctx_set_solver_tiling(self.h, target_blocks, min_chunk))

    def set_preconditioner(self, kind):
        """PRECOND_MG (multigrid V-cycle, default) or PRECOND_IPP for later pressure solves."""
        _check(lib().fg_ctx_set_preconditioner(self.h, kind))

    def set_vcycle_fused(self, on=True):
        """Fused V-cycle levels (default) or the per-level kernels (bit-identical solves)."""
        _check(lib().fg_ctx_set_vcycle_fused(self.h, 1 if on else 0))

    def set_solver_graphs(self, on=True):
        """Graph replay of PCG iterations is switched off: on=True raises (FG_E_INVALID)."""
        _check(lib().fg_ctx_set_solver_graphs(self.h, 1 if on else 0))

    def preconditioner(self):
        v = C.c_int32()
        _check(lib().fg_ctx_preconditioner(self.h, C.byref(v)))
        return v.value

    def batch_lanes(self):
        v = C.c_int32()
        _check(lib().fg_ctx_batch_lanes(self.h, C.byref(v)))
        return v.value

    def bytes_in_use(self):
        v = C.c_uint64()
        _check(lib().fg_ctx_bytes_in_use(self.h, C.byref(v)))
        return v.value

    def profile_kernel(self, kernel):
        _check(lib().fg_ctx_profile_kernel(self.h, kernel))

    def profile_read(self):
        """(summed launch ms, busy ms = union of launch intervals, launches, work units)"""
        t, b, n, u = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
        _check(lib().fg_ctx_profile_read(self.h, C.byref(t), C.byref(b), C.byref(n), C.byref(u)))
        return t.value, b.value, n.value, u.value

    def work_read(self, reset=False):
        """Work units of every kernel family (fg_kernel_id order) since the
        last reset, counted for every launch."""
        out = np.zeros(K_NUM, np.int64)
        _check(lib().fg_ctx_work_read(self.h, out.ctypes.data, 1 if reset else 0))
        return out

    def bandwidth_probe(self, nbytes=1 << 30, reps=5):
        g, ms = C.c_double(), C.c_double()
        _check(lib().fg_ctx_bandwidth_probe(self.h, nbytes, reps, C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def poisson_solve(self, material, scale, b, tol=1e-6, max_iters=0, noalloc=False):
        """fg_poisson_solve, or with noalloc fg_poisson_solve_noalloc (no
        device allocation in the call; systems up to its scratch size)."""
        material = np.ascontiguousarray(material, np.int8)
        b = np.ascontiguousarray(b, np.float64)
        I, J = material.shape
        x = np.zeros(I * J, np.float64)
        st = SolveStats()
        fn = lib().fg_poisson_solve_noalloc if noalloc else lib().fg_poisson_solve
        _check(fn(self.h, I, J, material.ctypes.data, scale, b.ctypes.data, x.ctypes.data, tol, max_iters,
                  C.byref(st)))
        return x, st

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Scene:
    def __init__(self, ctx: Context, params: SceneParams):
        self.ctx, self.params = ctx, params
        self.I, self.J = params.size_i, params.size_j
        self.h = C.c_void_p()
        _check(lib().fg_scene_create(ctx.h, C.byref(params), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().fg_scene_destroy(self.h)
            self.h = C.c_void_p()

    def init_geometry(self, solids, fluids, friction=None, seed=42):
        snv = np.array([len(w) for w in solids] or [0], np.int32)
        sv = np.ascontiguousarray(np.concatenate(solids).ravel() if solids else np.zeros(2), np.float32)
        fnv = np.array([len(f) for f in fluids] or [0], np.int32)
        fv = np.ascontiguousarray(np.concatenate(fluids).ravel() if fluids else np.zeros(2), np.float32)
        fr = None if friction is None else np.ascontiguousarray(friction, np.float32)
        _check(lib().fg_scene_init_geometry(self.h, len(solids), snv.ctypes.data, sv.ctypes.data, _ptr(fr),
                                            len(fluids), fnv.ctypes.data, fv.ctypes.data, seed))

    def init_objects(self, solids=(), sources=(), sinks=(), fluids=(), friction=None, emitters=(), seed=42,
                     fluid_viscosity=()):
        """First-frame init on the device (fg_scene_init_objects); emitters:
        (vx, vy, transfer[, viscosity]) per source polygon, velocity in m/s;
        fluid_viscosity: per initial-fluid polygon."""
        keep = []

        def polys(ps):
            nv = np.array([len(q) for q in ps] or [0], np.int32)
            v = np.ascontiguousarray(np.concatenate(ps).ravel() if len(ps) else np.zeros(2), np.float32)
            keep.extend([nv, v])
            return Polygons(len(ps), nv.ctypes.data, v.ctypes.data)

        o = SceneObjects()
        o.solids, o.sources, o.sinks, o.fluids = polys(solids), polys(sources), polys(sinks), polys(fluids)
        if friction is not None:
            fr = np.ascontiguousarray(friction, np.float32)
            keep.append(fr)
            o.friction = fr.ctypes.data
        if len(emitters):
            ev = np.ascontiguousarray([[e[0], e[1]] for e in emitters], np.float32).ravel()
            et = np.ascontiguousarray([1 if e[2] else 0 for e in emitters], np.uint8)
            evis = np.ascontiguousarray([e[3] if len(e) > 3 else 0.0 for e in emitters], np.float32)
            keep.extend([ev, et, evis])
            o.source_velocity, o.source_transfer = ev.ctypes.data, et.ctypes.data
            o.source_viscosity = evis.ctypes.data
        if len(fluid_viscosity):
            fvis = np.ascontiguousarray(fluid_viscosity, np.float32)
            keep.append(fvis)
            o.fluid_viscosity = fvis.ctypes.data
        o.seed = seed
        _check(lib().fg_scene_init_objects(self.h, C.byref(o)))

    def set_emitters(self, emitters, seed=0):
        ev = np.ascontiguousarray([[e[0], e[1]] for e in emitters] or [[0.0, 0.0]], np.float32).ravel()
        et = np.ascontiguousarray([1 if e[2] else 0 for e in emitters] or [0], np.uint8)
        _check(lib().fg_scene_set_emitters(self.h, len(emitters), ev.ctypes.data, et.ctypes.data, seed))

    def set_friction(self, friction):
        fr = np.ascontiguousarray(friction, np.float32)
        _check(lib().fg_scene_set_friction(self.h, len(fr), fr.ctypes.data))

    def set_state(self, st: dict):
        s = SceneState()
        keep = {}
        for name, _ in SceneState._fields_:
            if name == "n_particles":
                continue
            if name in st and st[name] is not None:
                if name in ("rng_epoch", "next_pid"):
                    setattr(s, name, int(st[name]))
                    continue
                dt = {"px": np.float32, "py": np.float32, "vx": np.float32, "vy": np.float32, "pid": np.uint32,
                      "u": np.float32, "v": np.float32, "u_valid": np.uint8, "v_valid": np.uint8,
                      "material": np.int8, "solid_sdf": np.float32, "fluid_sdf": np.float32,
                      "solid_id": np.int32, "emitter_id": np.int32, "particle_counts": np.int32,
                      "particle_viscosity": np.float32, "viscosity": np.float32}.get(name)
                if dt is None:
                    continue
                a = np.ascontiguousarray(st[name], dt)
                keep[name] = a
                setattr(s, name, a.ctypes.data)
        s.n_particles = len(keep.get("px", ()))
        _check(lib().fg_scene_set_state(self.h, C.byref(s)))

    def get_state(self):
        I, J = self.I, self.J
        n = self.particle_capacity_hint()
        out = {
            "px": np.zeros(n, np.float32), "py": np.zeros(n, np.float32), "vx": np.zeros(n, np.float32),
            "vy": np.zeros(n, np.float32), "pid": np.zeros(n, np.uint32),
            "u": np.zeros((I + 1) * J, np.float32), "v": np.zeros(I * (J + 1), np.float32),
            "u_valid": np.zeros((I + 1) * J, np.uint8), "v_valid": np.zeros(I * (J + 1), np.uint8),
            "material": np.zeros(I * J, np.int8), "solid_sdf": np.zeros(I * J, np.float32),
            "fluid_sdf": np.zeros(I * J, np.float32), "solid_id": np.zeros(I * J, np.int32),
            "particle_counts": np.zeros(I * J, np.int32), "saved_u": np.zeros((I + 1) * J, np.float32),
            "saved_v": np.zeros(I * (J + 1), np.float32), "pressure": np.zeros(I * J, np.float64),
            "rhs": np.zeros(I * J, np.float64), "stencil": np.zeros(I * J, np.uint8),
            "emitter_id": np.zeros(I * J, np.int32),
            "work_u": np.zeros((I + 1) * J, np.float32), "work_v": np.zeros(I * (J + 1), np.float32),
            "particle_viscosity": np.zeros(n, np.float32), "viscosity": np.zeros(I * J, np.float32),
        }
        s = SceneState()
        for k, a in out.items():
            setattr(s, k, a.ctypes.data)
        _check(lib().fg_scene_get_state(self.h, C.byref(s)))
        n = s.n_particles
        for k in ("px", "py", "vx", "vy", "pid", "particle_viscosity"):
            out[k] = out[k][:n]
        out["rng_epoch"], out["next_pid"] = s.rng_epoch, s.next_pid
        return out

    def particle_capacity_hint(self):
        return int(self.params.max_particles)

    def max_velocity(self):
        v = C.c_float()
        _check(lib().fg_scene_max_velocity(self.h, C.byref(v)))
        return v.value

    def particle_count(self):
        v = C.c_int64()
        _check(lib().fg_scene_particle_count(self.h, C.byref(v)))
        return v.value

    def substep(self, dt):
        st = Stats()
        _check(lib().fg_scene_substep(self.h, dt, C.byref(st)))
        return st

    def step_frame(self):
        st = Stats()
        _check(lib().fg_scene_step_frame(self.h, C.byref(st)))
        return st

    def stage(self, stage, dt=0.0):
        _check(lib().fg_scene_stage(self.h, stage, dt))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def box(x0, x1, y0, y1):
    return np.array([[x0, y0], [x0, y1], [x1, y1], [x1, y0]], np.float32)


def dam_break_geometry(n, fluid_box=(0.3, 1.0, 0.0, 0.4)):
    """Dam break in a 1x1 m box (x along I): one-cell solid walls, fluid rectangle clipped to the interior."""
    dx = 1.0 / n
    walls = [box(0.0, dx, 0.0, 1.0), box(1.0 - dx, 1.0, 0.0, 1.0), box(0.0, 1.0, 0.0, dx), box(0.0, 1.0, 1.0 - dx, 1.0)]
    x0, x1, y0, y1 = fluid_box
    fluid = [box(max(x0, dx), min(x1, 1.0 - dx), max(y0, dx), min(y1, 1.0 - dx))]
    return walls, fluid


def dam_break_scene(ctx, n, ppc=4, seed=42, fluid_box=(0.3, 1.0, 0.0, 0.4), **kw):
    p = make_params(n, ppc=ppc, **kw)
    sc = Scene(ctx, p)
    walls, fluid = dam_break_geometry(n, fluid_box)
    sc.init_geometry(walls, fluid, friction=np.zeros(len(walls), np.float32), seed=seed)
    return sc


def nb_dam_break_scene(ctx, n, ppc=8, seed=0, fluid_box=(0.3, 1.0, 0.0, 0.4), **kw):
    """The dam break run by NBFlipSolver (the solver the reference's demo
    scene file names): first frame on the device, no particles yet; seed is
    the reseed RNG's."""
    p = make_params(n, ppc=ppc, sim=SIM_NBFLIP, **kw)
    sc = Scene(ctx, p)
    walls, fluid = dam_break_geometry(n, fluid_box)
    sc.init_objects(solids=walls, fluids=fluid, friction=np.zeros(len(walls), np.float32), seed=seed)
    return sc


def batch_step_frame(ctx, scenes, allow_quarantine=False):
    """One FlipSolver::stepFrame for every scene, advanced together (one
    launch per phase over all scenes still inside their frame).  With
    allow_quarantine, a quarantined scene (stats[k].status) does not raise."""
    n = len(scenes)
    arr = (C.c_void_p * n)(*[s.h.value for s in scenes])
    stats = (Stats * n)()
    rc = lib().fg_batch_step_frame(ctx.h, arr, n, stats)
    if not (allow_quarantine and rc == E_NUMERIC):
        _check(rc)
    return list(stats)


def batch_substep(ctx, scenes, dts, allow_quarantine=False):
    """One FlipSolver::step of every scene, scene k with dts[k], advanced
    together (one launch per phase)."""
    n = len(scenes)
    arr = (C.c_void_p * n)(*[s.h.value for s in scenes])
    d = np.ascontiguousarray(dts, np.float32)
    assert len(d) == n
    stats = (Stats * n)()
    rc = lib().fg_batch_substep(ctx.h, arr, n, d.ctypes.data, stats)
    if not (allow_quarantine and rc == E_NUMERIC):
        _check(rc)
    return list(stats)


def stencil_matrix(stencil):
    """Split the stencil bytes of get_state() into the (n, mask, fluid)
    arrays of the pressure matrix (bits 4-6, bits 0-3, bit 7)."""
    st = np.asarray(stencil, np.uint8)
    return ((st >> 4) & 7).astype(np.uint8), (st & 15).astype(np.uint8), (st >> 7).astype(np.uint8)


TILING_CANDIDATES = ((1024, 256), (1024, 512), (1024, 1024), (1024, 2048), (1024, 4096))


def batch_autotune_tiling(ctx, scenes, rounds=1, candidates=TILING_CANDIDATES):
    """Measure each (target_blocks, min_chunk) solver tiling on `rounds`
    frames of the batch (the scenes advance, and their FP64 sums follow each
    tiling's order: tune on a copy), set the fastest on the context;
    returns (index, rates)."""
    n = len(scenes)
    arr = (C.c_void_p * n)(*[s.h.value for s in scenes])
    cand = np.asarray(candidates, np.int32).reshape(-1)
    best = C.c_int32()
    rates = np.zeros(len(candidates), np.float64)
    _check(lib().fg_batch_autotune_tiling(ctx.h, arr, n, rounds, cand.ctypes.data, len(candidates), C.byref(best),
                                          rates.ctypes.data))
    return best.value, rates


def batch_autotune(ctx, scenes, rounds=1):
    """Measure each lane count on `rounds` frames of the batch (the scenes
    advance), set the fastest on the context; returns (lanes, rates)."""
    n = len(scenes)
    arr = (C.c_void_p * n)(*[s.h.value for s in scenes])
    best = C.c_int32()
    rates = np.zeros(MAX_LANES, np.float64)
    _check(lib().fg_batch_autotune(ctx.h, arr, n, rounds, C.byref(best), rates.ctypes.data))
    return best.value, rates


MIXED_RESOLUTIONS = (64, 96, 128, 192, 256)


def mixed_scene_specs(count=256, master_seed=2024, resolutions=MIXED_RESOLUTIONS):
    """Independent dam-break scenes of uneven size: resolution from
    `resolutions`, fluid-rectangle corners uniform in [0.1, 0.6] of the box on
    each axis, gravity 9.81 m/s^2 at an angle uniform in +-30 degrees from +x,
    4-8 particles per cell, one seed per scene drawn from the master seed."""
    rng = np.random.default_rng(master_seed)
    specs = []
    for k in range(count):
        n = int(rng.choice(resolutions))
        xs = np.sort(rng.uniform(0.1, 0.6, 2))
        ys = np.sort(rng.uniform(0.1, 0.6, 2))
        ang = np.deg2rad(rng.uniform(-30.0, 30.0))
        ppc = int(rng.integers(4, 9))
        seed = int(rng.integers(0, 2**31 - 1))
        specs.append({"n": n, "fluid_box": (float(xs[0]), float(xs[1]), float(ys[0]), float(ys[1])),
                      "gravity": (float(9.81 * np.cos(ang)), float(9.81 * np.sin(ang))), "ppc": ppc, "seed": seed})
    return specs


class _SceneView(Scene):
    """A scene owned by a server (not destroyed by close())."""

    def __init__(self, handle, params):
        self.ctx, self.params = None, params
        self.I, self.J = params.size_i, params.size_j
        self.h = C.c_void_p(handle)

    def close(self):
        self.h = C.c_void_p()


class Server:
    """Continuous batching of scene simulations (fg_server_*): requests queue
    in arrival order, are admitted while their memory fits `hbm_budget` and a
    slot is free, and every tick steps one frame of all resident scenes in
    one batched call."""

    def __init__(self, ctx: Context, hbm_budget: int, max_resident: int = 256):
        self.ctx = ctx
        self.h = C.c_void_p()
        self._params = {}
        _check(lib().fg_server_create(ctx.h, hbm_budget, max_resident, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().fg_server_destroy(self.h)
            self.h = C.c_void_p()

    def submit(self, params, solids, fluids, frames, seed=42, friction=None, keep=False, fluid_viscosity=None):
        snv = np.array([len(w) for w in solids] or [0], np.int32)
        sv = np.ascontiguousarray(np.concatenate(solids).ravel() if solids else np.zeros(2), np.float32)
        fnv = np.array([len(f) for f in fluids] or [0], np.int32)
        fv = np.ascontiguousarray(np.concatenate(fluids).ravel() if fluids else np.zeros(2), np.float32)
        fr = None if friction is None else np.ascontiguousarray(friction, np.float32)
        r = Request()
        r.params = params
        r.n_solids, r.solid_nverts, r.solid_verts, r.friction = len(solids), snv.ctypes.data, sv.ctypes.data, _ptr(fr)
        r.n_fluids, r.fluid_nverts, r.fluid_verts = len(fluids), fnv.ctypes.data, fv.ctypes.data
        r.seed, r.frames, r.keep = seed, frames, 1 if keep else 0
        fvis = None if fluid_viscosity is None else np.ascontiguousarray(fluid_viscosity, np.float32)
        r.fluid_viscosity = _ptr(fvis)
        rid = C.c_int64()
        _check(lib().fg_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._params[rid.value] = params
        return rid.value

    def submit_spec(self, spec, frames, keep=False, **kw):
        p = make_params(spec["n"], ppc=spec["ppc"], gravity=spec["gravity"], **kw)
        walls, fluid = dam_break_geometry(spec["n"], spec["fluid_box"])
        return self.submit(p, walls, fluid, frames, seed=spec["seed"], friction=np.zeros(len(walls), np.float32),
                           keep=keep)

    def tick(self):
        t = TickStats()
        _check(lib().fg_server_tick(self.h, C.byref(t)))
        return t

    def status(self, rid):
        s = RequestStatus()
        _check(lib().fg_server_status(self.h, rid, C.byref(s)))
        return s

    def scene(self, rid):
        h = C.c_void_p()
        _check(lib().fg_server_scene(self.h, rid, C.byref(h)))
        return _SceneView(h.value, self._params[rid])

    def release(self, rid):
        _check(lib().fg_server_release(self.h, rid))

    def cancel(self, rid):
        _check(lib().fg_server_cancel(self.h, rid))

    def summary(self):
        s = ServerSummary()
        _check(lib().fg_server_summary_get(self.h, C.byref(s)))
        return s

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def served_request_specs(rate=200.0, duration=120.0, master_seed=77, res_range=(64, 1024),
                         frame_range=(30, 300)):
    """A Poisson stream of scene requests: arrival times at `rate` per second
    over `duration` seconds; per request a resolution log-uniform in
    `res_range`, the dam-break geometry, gravity angle and 4-8 particles per
    cell of `mixed_scene_specs`, and a frame count uniform in `frame_range`."""
    rng = np.random.default_rng(master_seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n = int(round(np.exp(rng.uniform(np.log(res_range[0]), np.log(res_range[1])))))
        xs = np.sort(rng.uniform(0.1, 0.6, 2))
        ys = np.sort(rng.uniform(0.1, 0.6, 2))
        ang = np.deg2rad(rng.uniform(-30.0, 30.0))
        out.append({"t": t, "n": n, "fluid_box": (float(xs[0]), float(xs[1]), float(ys[0]), float(ys[1])),
                    "gravity": (float(9.81 * np.cos(ang)), float(9.81 * np.sin(ang))),
                    "ppc": int(rng.integers(4, 9)), "seed": int(rng.integers(0, 2**31 - 1)),
                    "frames": int(rng.integers(frame_range[0], frame_range[1] + 1))})
    return out


def isolated_projection_system(n=4096, seed=1, dt=1.0 / 30, density=1000.0):
    """The isolated pressure projection: an n x n MAC grid whose liquid mask
    is uniform noise smoothed by a Gaussian (sigma 8 cells, reflecting edges)
    and thresholded at its 40th percentile (~60 % liquid, the rest air), plus
    a one-cell solid border and 2 % random solid cells; face velocities
    uniform in [-1, 1] (FP32); the right-hand side is calcPressureRhs
    (flipsolver2d.cpp:962-991) of those faces.  Returns (material[n, n] int8,
    u, v, b, scale) with scale = dt / (rho dx^2), dx = 1/n."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    smooth = gaussian_filter(rng.random((n, n), dtype=np.float32), 8.0)
    mat = np.where(smooth > np.percentile(smooth, 40), FLUID, EMPTY).astype(np.int8)
    mat[0, :] = mat[-1, :] = mat[:, 0] = mat[:, -1] = SOLID
    mat[rng.random((n, n)) < 0.02] = SOLID
    u = rng.uniform(-1.0, 1.0, (n + 1, n)).astype(np.float32)
    v = rng.uniform(-1.0, 1.0, (n, n + 1)).astype(np.float32)
    dx = 1.0 / n
    return mat, u, v, pressure_rhs(mat, u, v, dx), dt / (density * dx * dx)


def pressure_rhs(mat, u, v, dx):
    """FlipSolver::calcPressureRhs + divergenceAt (flipsolver2d.cpp:759-764,
    962-991) on host arrays: the divergence in float, widened, with the solid
    face terms (solid velocity 0) in the same order; neighbour materials
    clamp at the edges."""
    scale = 1.0 / dx
    u0, u1 = u[:-1, :], u[1:, :]
    v0, v1 = v[:, :-1], v[:, 1:]
    div = (((u1 - u0) + v1) - v0) + np.float32(0.0)
    ext = np.pad(mat, 1, mode="edge")
    solid = (ext & SOLID) != 0
    s_im, s_ip = solid[:-2, 1:-1], solid[2:, 1:-1]
    s_jm, s_jp = solid[1:-1, :-2], solid[1:-1, 2:]
    val = -scale * div.astype(np.float64)
    val = val - s_im * scale * u0.astype(np.float64)
    val = val + s_ip * scale * u1.astype(np.float64)
    val = val - s_jm * scale * v0.astype(np.float64)
    val = val + s_jp * scale * v1.astype(np.float64)
    return np.where((mat & FLUID) != 0, val, 0.0).ravel()


def make_scene(ctx, spec, **kw):
    """A dam-break scene from one `mixed_scene_specs` entry."""
    return dam_break_scene(ctx, spec["n"], ppc=spec["ppc"], seed=spec["seed"], fluid_box=spec["fluid_box"],
                           gravity=spec["gravity"], **kw)
