"""ctypes bindings for the CPU oracle and the reference building blocks.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg.  The product package never imports this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfsref.so")

SOLVER_REF_REPLICA = 0
SOLVER_CONVERGED = 1
DENSITY_OFF, DENSITY_GATED, DENSITY_ON = 0, 1, 2
SIM_FLIP, SIM_NBFLIP = 0, 1
FLUID, SOURCE, SOLID, SINK, EMPTY = 0x40, 0x41, 0x20, 0x12, 0x10

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


class OrcParams(C.Structure):
    _fields_ = [
        ("size_i", C.c_int32), ("size_j", C.c_int32),
        ("particles_per_cell", C.c_int32), ("pcg_iter_limit", C.c_int32),
        ("dx", C.c_double), ("fluid_density", C.c_double),
        ("gravity_i", C.c_float), ("gravity_j", C.c_float),
        ("pic_ratio", C.c_float), ("cfl_number", C.c_float), ("particle_scale", C.c_float),
        ("fps", C.c_int32), ("max_substeps", C.c_int32),
        ("solver_mode", C.c_int32), ("density_mode", C.c_int32), ("threads", C.c_int32),
        ("project_tol", C.c_double),
        ("n_obstacles", C.c_int32), ("friction", C.POINTER(C.c_float)),
        ("n_emitters", C.c_int32), ("emitter_velocity", C.c_void_p), ("emitter_transfer", C.c_void_p),
        ("reseed_seed", C.c_uint32),
        ("sim_method", C.c_int32), ("source_sdf", C.c_void_p),
        ("viscosity_enabled", C.c_int32), ("emitter_viscosity", C.c_void_p), ("fluid_viscosity", C.c_void_p),
        ("source_id", C.c_void_p),
    ]


class OrcState(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int64), ("capacity", C.c_int64),
        ("px", C.c_void_p), ("py", C.c_void_p), ("vx", C.c_void_p), ("vy", C.c_void_p),
        ("pid", C.c_void_p),
        ("u", C.c_void_p), ("v", C.c_void_p),
        ("u_valid", C.c_void_p), ("v_valid", C.c_void_p),
        ("material", C.c_void_p), ("solid_sdf", C.c_void_p), ("fluid_sdf", C.c_void_p),
        ("solid_id", C.c_void_p), ("particle_counts", C.c_void_p),
        ("saved_u", C.c_void_p), ("saved_v", C.c_void_p),
        ("pressure", C.c_void_p), ("rhs", C.c_void_p), ("density", C.c_void_p),
        ("emitter_id", C.c_void_p), ("rng_epoch", C.c_uint32), ("next_pid", C.c_uint32),
        ("pvisc", C.c_void_p), ("viscosity", C.c_void_p),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("pressure_iters", C.c_int32), ("density_iters", C.c_int32),
        ("density_applied", C.c_int32), ("substeps", C.c_int32),
        ("pressure_rel_resid", C.c_double), ("density_rel_resid", C.c_double),
        ("particles_killed", C.c_int64), ("particles_erased", C.c_int64),
        ("last_dt", C.c_float), ("max_velocity", C.c_float),
        ("stage_ms", C.c_double * 12),
    ]


def build(force: bool = False, ref: bool = True) -> None:
    """Compile the oracle (and, where the reference tree exists, the reference blocks)."""
    targets = ["oracle"]
    if ref and os.path.isdir("/root/reference/FlipSolver2dLib"):
        targets.append("ref")
    if force:
        subprocess.run(["make", "-C", HERE, "-B"] + targets, check=True, capture_output=True)
    else:
        subprocess.run(["make", "-C", HERE] + targets, check=True, capture_output=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        _lib = C.CDLL(ORACLE_SO)
        _declare_oracle(_lib)
    return _lib


def ref_lib():
    """The reference building blocks, or None where they were not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = C.CDLL(REF_SO)
        _declare_ref(_ref)
    return _ref


def _declare_oracle(L):
    P, S, T = C.POINTER(OrcParams), C.POINTER(OrcState), C.POINTER(OrcStats)
    sig = {
        "orc_bspline": (C.c_float, [C.c_float]),
        "orc_quadratic_bspline": (C.c_float, [C.c_float, C.c_float]),
        "orc_linear_hat": (C.c_float, [C.c_float]),
        "orc_lerp": (None, [_f32p, C.c_int32, C.c_int32, C.c_float, C.c_float, _f32p, _f32p, C.c_int64, _f32p]),
        "orc_velocity_at": (None, [_f32p, _f32p, C.c_int32, C.c_int32, _f32p, _f32p, C.c_int64, _f32p, _f32p]),
        "orc_rk4": (None, [_f32p, _f32p, C.c_int32, C.c_int32, C.c_float, _f32p, _f32p, C.c_int64, _f32p, _f32p]),
        "orc_closest_surface_point": (None, [_f32p, C.c_int32, C.c_int32, _f32p, _f32p, C.c_int64, _f32p, _f32p]),
        "orc_polygon_sdf": (C.c_float, [_f32p, C.c_int32, C.c_float, C.c_float]),
        "orc_advect": (None, [P, S, C.c_float, _u8p]),
        "orc_sort_particles": (None, [P, S, C.c_void_p, C.c_void_p]),
        "orc_cell_start": (None, [P, S, _i32p]),
        "orc_max_particle_velocity": (C.c_float, [S]),
        "orc_count_particles": (C.c_int32, [P, S, _i32p]),
        "orc_g2p": (None, [P, S]),
        "orc_p2g": (None, [P, S, _i32p]),
        "orc_fluid_sdf": (None, [P, S, _i32p]),
        "orc_update_materials": (None, [P, S]),
        "orc_extrapolate_levelset_inside": (None, [_f32p, C.c_int32, C.c_int32]),
        "orc_extrapolate_levelset_outside": (None, [_f32p, C.c_int32, C.c_int32]),
        "orc_euler_advect": (None, [_f32p, _f32p, C.c_int32, C.c_int32, C.c_float, _f32p, C.c_int32, C.c_int32,
                                    C.c_float, C.c_float, _f32p]),
        "orc_nb_prune": (None, [P, S, _u8p]),
        "orc_viscosity_grid": (None, [P, S, _i32p]),
        "orc_viscosity_matrix": (None, [_i8p, _f32p, C.c_int32, C.c_int32, C.c_float, _f64p, _f64p, _f64p]),
        "orc_viscosity_apply": (C.c_int32, [P, S, C.c_float, C.c_double]),
        "orc_nb_combine": (None, [P, S, _f32p, _f32p, _f32p]),
        "orc_nb_reseed": (C.c_int64, [P, S]),
        "orc_nb_substep": (C.c_int32, [P, S, C.c_float, T]),
        "orc_nb_init_objects": (None, [P, S, C.c_int32, _i32p, _f32p, C.c_int32, _i32p, _f32p, C.c_int32, _i32p,
                                       _f32p, C.c_int32, _i32p, _f32p, _f32p, _i32p]),
        "orc_nb_viscosity_grid": (None, [P, S, _i32p]),
        "orc_nb_combine_viscosity": (None, [P, S, _f32p]),
        "orc_extrapolate": (None, [_f32p, _u8p, C.c_int32, C.c_int32, C.c_int32]),
        "orc_body_forces": (None, [P, S, C.c_float]),
        "orc_pressure_rhs": (None, [P, S, _f64p]),
        "orc_apply_pressure": (None, [P, S, C.c_float, _f64p]),
        "orc_velocity_from_solids": (None, [P, S]),
        "orc_density_grid": (None, [P, S, _i32p]),
        "orc_density_rhs": (None, [P, S, C.c_float, _f64p]),
        "orc_adjust_particles": (None, [P, S, C.c_float, _f64p]),
        "orc_build_matrix": (None, [_i8p, C.c_int32, C.c_int32, _u8p, _u8p, _u8p]),
        "orc_spmv": (None, [_u8p, _u8p, _u8p, C.c_int32, C.c_int32, C.c_double, _f64p, _f64p]),
        "orc_ipp_ref_apply": (None, [_i8p, _u8p, C.c_int32, C.c_int32, C.c_double, _f64p, _f64p]),
        "orc_ipp_apply": (None, [_u8p, _u8p, _u8p, C.c_int32, C.c_int32, _f64p, _f64p]),
        "orc_pcg_replica": (C.c_int32, [_i8p, C.c_int32, C.c_int32, C.c_double, _f64p, _f64p, C.c_int32,
                                        C.c_double, C.c_int32, C.c_void_p]),
        "orc_pcg_converged": (C.c_int32, [_i8p, C.c_int32, C.c_int32, C.c_double, _f64p, _f64p, C.c_int32,
                                          C.c_double, C.POINTER(C.c_double)]),
        "orc_rel_residual": (C.c_double, [_i8p, C.c_int32, C.c_int32, C.c_double, _f64p, _f64p]),
        "orc_split_range": (C.c_int32, [C.c_int64, C.c_int64, C.c_int64, C.c_int32, _i64p, _i64p]),
        "orc_material_bits": (C.c_int32, [C.c_int8]),
        "orc_nonsolid_count_lin": (C.c_int32, [_i8p, C.c_int32, C.c_int32, C.c_int64]),
        "orc_maxabs_replica": (C.c_double, [_f64p, C.c_int64, C.c_int32]),
        "orc_dot_replica": (C.c_double, [_f64p, _f64p, C.c_int64, C.c_int32]),
        "orc_frame_dt": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_int32)]),
        "orc_substep": (C.c_int32, [P, S, C.c_float, T]),
        "orc_step_frame": (C.c_int32, [P, S, T]),
        "orc_init_geometry": (None, [P, S, C.c_int32, _i32p, _f32p, C.c_int32, _i32p, _f32p]),
        "orc_seed_particles": (C.c_int64, [P, S, C.c_uint32]),
        "orc_jitter": (C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_init_objects": (None, [P, S, C.c_int32, _i32p, _f32p, C.c_int32, _i32p, _f32p, C.c_int32, _i32p, _f32p,
                                    C.c_int32, _i32p, _f32p]),
        "orc_after_transfer": (None, [P, S]),
        "orc_reseed": (C.c_int64, [P, S]),
        "orc_reseed_jitter": (C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _declare_ref(L):
    sig = {
        "ref_bspline": (C.c_float, [C.c_float]),
        "ref_quadratic_bspline": (C.c_float, [C.c_float, C.c_float]),
        "ref_linear_hat": (C.c_float, [C.c_float]),
        "ref_bilinear_hat": (C.c_float, [C.c_float, C.c_float]),
        "ref_frac": (C.c_float, [C.c_float]),
        "ref_integr": (C.c_int, [C.c_float]),
        "ref_thread_count": (C.c_int32, []),
        "ref_split_range": (C.c_int32, [C.c_int64, C.c_int64, C.c_int64, _i64p, _i64p]),
        "ref_lerp": (None, [_f32p, C.c_int32, C.c_int32, C.c_float, C.c_float, _f32p, _f32p, C.c_int64, _f32p]),
        "ref_velocity_at": (None, [_f32p, _f32p, C.c_int32, C.c_int32, _f32p, _f32p, C.c_int64, _f32p, _f32p]),
        "ref_closest_surface_point": (None, [_f32p, C.c_int32, C.c_int32, _f32p, _f32p, C.c_int64, _f32p, _f32p]),
        "ref_extrapolate": (None, [_f32p, _u8p, C.c_int32, C.c_int32, C.c_int32]),
        "ref_polygon_sdf": (C.c_float, [_f32p, C.c_int32, C.c_float, C.c_float]),
        "ref_material_bits": (C.c_int32, [C.c_int8]),
        "ref_nonsolid_count_lin": (C.c_int32, [_i8p, C.c_int32, C.c_int32, C.c_int64]),
        "ref_sample_affected_by_solid": (None, [_i8p, C.c_int32, C.c_int32, _u8p, _u8p]),
        "ref_bins_for_cell": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _i64p]),
        "ref_vops_maxabs": (C.c_double, [_f64p, C.c_int64]),
        "ref_vops_dot": (C.c_double, [_f64p, _f64p, C.c_int64]),
        "ref_pcg": (C.c_int32, [_i8p, _u8p, _u8p, _u8p, C.c_int32, C.c_int32, C.c_double, _f64p, _f64p,
                                C.c_int32, C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


# ---------------------------------------------------------------------------
# Scene helpers
# ---------------------------------------------------------------------------

def make_params(n_i, n_j=None, *, dx=None, ppc=4, density=1000.0, gravity=(9.81, 0.0), pic=0.03, cfl=5.0,
                particle_scale=0.8, fps=30, max_substeps=10, pcg_iter_limit=200,
                solver=SOLVER_CONVERGED, density_mode=DENSITY_OFF, threads=1, tol=None, friction=(0.0,),
                emitters=(), reseed_seed=0, sim=SIM_FLIP, viscosity=False, fluid_viscosity=()):
    """emitters: (vx, vy, transfer[, viscosity]) per source, velocity in m/s;
    sim: SIM_FLIP (FlipSolver) or SIM_NBFLIP (NBFlipSolver); viscosity:
    viscosityEnabled (the light model); fluid_viscosity: per initial-fluid
    polygon."""
    n_j = n_i if n_j is None else n_j
    p = OrcParams()
    p.size_i, p.size_j = n_i, n_j
    p.particles_per_cell = ppc
    p.pcg_iter_limit = pcg_iter_limit
    p.dx = (1.0 / max(n_i, n_j)) if dx is None else dx
    p.fluid_density = density
    p.gravity_i, p.gravity_j = gravity
    p.pic_ratio, p.cfl_number, p.particle_scale = pic, cfl, particle_scale
    p.fps, p.max_substeps = fps, max_substeps
    p.solver_mode, p.density_mode, p.threads = solver, density_mode, threads
    if tol is None:
        tol = 1e-2 if solver == SOLVER_REF_REPLICA else 1e-12
    p.project_tol = tol
    fr = np.ascontiguousarray(friction, dtype=np.float32)
    p._friction_keep = fr
    p.n_obstacles = len(fr)
    p.friction = fr.ctypes.data_as(C.POINTER(C.c_float))
    ev = np.ascontiguousarray([[e[0], e[1]] for e in emitters] or [[0.0, 0.0]], np.float32).ravel()
    et = np.ascontiguousarray([1 if e[2] else 0 for e in emitters] or [0], np.uint8)
    p._emitter_keep = (ev, et)
    p.n_emitters = len(emitters)
    p.emitter_velocity = ev.ctypes.data
    p.emitter_transfer = et.ctypes.data
    p.reseed_seed = reseed_seed
    p.sim_method = sim
    p.source_sdf = None
    p.viscosity_enabled = 1 if viscosity else 0
    evis = np.ascontiguousarray([e[3] if len(e) > 3 else 0.0 for e in emitters] or [0.0], np.float32)
    fvis = np.ascontiguousarray(list(fluid_viscosity) or [0.0], np.float32)
    p._visc_keep = (evis, fvis)
    p.emitter_viscosity = evis.ctypes.data
    p.fluid_viscosity = fvis.ctypes.data
    p.source_id = None
    return p


class State:
    """numpy-backed OrcState."""

    def __init__(self, n_i, n_j, capacity):
        I, J = n_i, n_j
        self.I, self.J = I, J
        self.px = np.zeros(capacity, np.float32)
        self.py = np.zeros(capacity, np.float32)
        self.vx = np.zeros(capacity, np.float32)
        self.vy = np.zeros(capacity, np.float32)
        self.pid = np.zeros(capacity, np.uint32)
        self.u = np.zeros((I + 1) * J, np.float32)
        self.v = np.zeros(I * (J + 1), np.float32)
        self.u_valid = np.zeros((I + 1) * J, np.uint8)
        self.v_valid = np.zeros(I * (J + 1), np.uint8)
        self.material = np.full(I * J, EMPTY, np.int8)
        self.solid_sdf = np.zeros(I * J, np.float32)
        self.fluid_sdf = np.zeros(I * J, np.float32)
        self.solid_id = np.full(I * J, -1, np.int32)
        self.particle_counts = np.zeros(I * J, np.int32)
        self.saved_u = np.zeros((I + 1) * J, np.float32)
        self.saved_v = np.zeros(I * (J + 1), np.float32)
        self.pressure = np.zeros(I * J, np.float64)
        self.rhs = np.zeros(I * J, np.float64)
        self.density = np.zeros(I * J, np.float32)
        self.emitter_id = np.full(I * J, -1, np.int32)
        self.pvisc = np.zeros(capacity, np.float32)
        self.viscosity = np.zeros(I * J, np.float32)
        self.c = OrcState()
        self.c.capacity = capacity
        self.c.n_particles = 0
        for name in ("px", "py", "vx", "vy", "pid", "u", "v", "u_valid", "v_valid", "material", "solid_sdf",
                     "fluid_sdf", "solid_id", "particle_counts", "saved_u", "saved_v", "pressure", "rhs",
                     "density", "emitter_id", "pvisc", "viscosity"):
            setattr(self.c, name, getattr(self, name).ctypes.data)

    @property
    def n(self):
        return int(self.c.n_particles)

    @n.setter
    def n(self, value):
        self.c.n_particles = int(value)

    def ref(self):
        return C.byref(self.c)

    def copy(self):
        o = State(self.I, self.J, len(self.px))
        for name in ("px", "py", "vx", "vy", "pid", "u", "v", "u_valid", "v_valid", "material", "solid_sdf",
                     "fluid_sdf", "solid_id", "particle_counts", "saved_u", "saved_v", "pressure", "rhs",
                     "density", "emitter_id", "pvisc", "viscosity"):
            getattr(o, name)[...] = getattr(self, name)
        o.n = self.n
        o.c.rng_epoch, o.c.next_pid = self.c.rng_epoch, self.c.next_pid
        return o


def box_polys(x0, x1, y0, y1):
    return np.array([[x0, y0], [x0, y1], [x1, y1], [x1, y0]], np.float32)


def dam_break_geometry(n, fluid_box=(0.3, 1.0, 0.0, 0.4)):
    """Dam break in a 1x1 m box: 4 one-cell solid walls, fluid rectangle clipped to the interior."""
    dx = 1.0 / n
    walls = [box_polys(0.0, dx, 0.0, 1.0), box_polys(1.0 - dx, 1.0, 0.0, 1.0),
             box_polys(0.0, 1.0, 0.0, dx), box_polys(0.0, 1.0, 1.0 - dx, 1.0)]
    x0, x1, y0, y1 = fluid_box
    fluid = [box_polys(max(x0, dx), min(x1, 1.0 - dx), max(y0, dx), min(y1, 1.0 - dx))]
    return walls, fluid


def init_scene(params, walls, fluids, seed=42, capacity=None):
    """First-frame init + quadrant-jittered seeding through the oracle."""
    L = lib()
    I, J = params.size_i, params.size_j
    if capacity is None:
        capacity = I * J * params.particles_per_cell
    st = State(I, J, capacity)
    snv = np.array([len(w) for w in walls], np.int32)
    sv = np.ascontiguousarray(np.concatenate(walls).ravel(), np.float32) if walls else np.zeros(2, np.float32)
    fnv = np.array([len(f) for f in fluids], np.int32)
    fv = np.ascontiguousarray(np.concatenate(fluids).ravel(), np.float32) if fluids else np.zeros(2, np.float32)
    if len(snv) == 0:
        snv = np.zeros(1, np.int32)
    if len(fnv) == 0:
        fnv = np.zeros(1, np.int32)
    L.orc_init_geometry(C.byref(params), st.ref(), len(walls), snv, sv, len(fluids), fnv, fv)
    n = L.orc_seed_particles(C.byref(params), st.ref(), seed)
    assert n >= 0, "particle capacity exceeded"
    st.c.next_pid = n
    return st


def _polys(ps):
    nv = np.array([len(q) for q in ps] or [0], np.int32)
    v = np.ascontiguousarray(np.concatenate(ps).ravel() if len(ps) else np.zeros(2), np.float32)
    return len(ps), nv, v


def init_scene_objects(params, solids=(), sources=(), sinks=(), fluids=(), seed=42, capacity=None):
    """firstFrameInit (sinks, sources, solids, initial fluid) + seeding through the oracle."""
    L = lib()
    I, J = params.size_i, params.size_j
    if capacity is None:
        capacity = I * J * params.particles_per_cell
    st = State(I, J, capacity)
    L.orc_init_objects(C.byref(params), st.ref(), *_polys(solids), *_polys(sources), *_polys(sinks), *_polys(fluids))
    n = L.orc_seed_particles(C.byref(params), st.ref(), seed)
    assert n >= 0, "particle capacity exceeded"
    st.c.next_pid = n
    return st


def dam_break(n, ppc=4, seed=42, **kw):
    params = make_params(n, ppc=ppc, **kw)
    walls, fluids = dam_break_geometry(n)
    st = init_scene(params, walls, fluids, seed=seed)
    return params, st


def init_nb_scene(params, solids=(), sources=(), sinks=(), fluids=(), capacity=None):
    """NBFlipSolver::firstFrameInit through the oracle: materials, the fluid
    SDF of the initial fluid and no particles.  Sets params.source_sdf (kept
    alive on params) to the per-cell source distance the substeps read."""
    L = lib()
    I, J = params.size_i, params.size_j
    if capacity is None:
        capacity = I * J * params.particles_per_cell * 2
    st = State(I, J, capacity)
    src = np.zeros(I * J, np.float32)
    sid = np.zeros(I * J, np.int32)
    L.orc_nb_init_objects(C.byref(params), st.ref(), *_polys(solids), *_polys(sources), *_polys(sinks),
                          *_polys(fluids), src, sid)
    params._source_sdf_keep = (src, sid)
    params.source_sdf = src.ctypes.data
    params.source_id = sid.ctypes.data
    return st


def nb_dam_break(n, ppc=8, **kw):
    """The dam break run by NBFlipSolver (the demo scene's solver)."""
    params = make_params(n, ppc=ppc, sim=SIM_NBFLIP, **kw)
    walls, fluids = dam_break_geometry(n)
    st = init_nb_scene(params, solids=walls, fluids=fluids)
    return params, st
