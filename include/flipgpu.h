/*
 * flipgpu.h -- C-ABI of the MI355X-native FLIP substep (libflipgpu.so).
 *
 * Plain C types only: no C++ or torch types cross this boundary.  Every
 * function returns FG_OK (0) or a negative FG_E* code; fg_last_error()
 * returns the calling thread's last message.  No C++ exception crosses.
 * Host buffers passed in are caller-owned and only read/written during the
 * call; device state is owned by the context.  A context is driven from one
 * host thread; all device work is ordered on the scene's stream and every
 * fg_*_get_* call synchronises that stream.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   FlipSolver::FlipSolver(const FlipSolverParameters*)  FlipSolver2dLib/flipsolver2d.cpp:27-86
 *       -> fg_scene_create
 *   FlipSolver::firstFrameInit / updateSinks / updateSources / updateSolids /
 *       updateInitialFluid / seedInitialFluid, FlipSolver::addGeometry / addSource / addSink /
 *       addInitialFluid  FlipSolver2dLib/flipsolver2d.cpp:502-625, 682-707, 788-795
 *       -> fg_scene_init_objects (fg_scene_init_geometry: solids and fluid only)
 *   FlipSolver::stepFrame()          FlipSolver2dLib/flipsolver2d.cpp:464-500 -> fg_scene_step_frame
 *   FlipSolver::step() (protected)   FlipSolver2dLib/flipsolver2d.cpp:412-462 -> fg_scene_substep
 *   NBFlipSolver (constructor, step, firstFrameInit)  FlipSolver2dLib/nbflipsolver.cpp:9-64, 203-211
 *       -> fg_scene_params.sim_method = FG_SIM_NBFLIP, then the same entry points
 *   FlipSolver::advect/particleToGrid/updateSdf/project/particleUpdate/...
 *       (protected virtual hooks, FlipSolver2dLib/flipsolver2d.h:280-367) -> fg_scene_stage
 *   FlipSolver accessors markerParticles/materialGrid/fluidVelocityGrid/fluidSdf/solidSdf
 *       FlipSolver2dLib/flipsolver2d.h:226-236 -> fg_scene_get_state
 *   FlipSolver::timeStats()          FlipSolver2dLib/flipsolver2d.h:88-181,238 -> fg_stats
 *   LinearSolver::solve(A, M, x, b, iterLimit, tol)  FlipSolver2dLib/linearsolver.h:15-20
 *       -> fg_poisson_solve
 */
#ifndef FLIPGPU_H
#define FLIPGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FG_ABI_VERSION 8

enum fg_status {
    FG_OK = 0,
    FG_E_INVALID = -1,     /* bad argument */
    FG_E_DEVICE = -2,      /* HIP runtime error */
    FG_E_BUDGET = -3,      /* HBM budget would be exceeded */
    FG_E_CAPACITY = -4,    /* particle capacity exceeded */
    FG_E_STATE = -5,       /* call not valid in the current state */
    FG_E_NUMERIC = -6,     /* NaN/Inf detected: the scene is quarantined (see below) */
    FG_E_CANCELLED = -7    /* the request was cancelled */
};

/* Quarantine.  A scene whose particle velocities or pressure solve turn
 * non-finite is quarantined: it leaves the launch group at once (the other
 * scenes of a batch finish their frame exactly as they would without it),
 * its fg_stats.status is FG_E_NUMERIC, and every later step call skips it
 * and reports FG_E_NUMERIC for it until fg_scene_set_state or
 * fg_scene_init_geometry gives it a new state.  The reference has no
 * counterpart (FlipSolver::project only prints when its solve fails,
 * flipsolver2d.cpp:107-111). */

/* Material codes, FlipSolver2dLib/materialgrid.h:6-13 */
enum fg_material {
    FG_FLUID = 0x40, FG_SOURCE = 0x41, FG_SOLID = 0x20, FG_SINK = 0x12, FG_EMPTY = 0x10
};

/* Density correction (FlipSolver::densityCorrection, flipsolver2d.cpp:164-186).
 * The reference applies it only when its PCG converges within pcgIterLimit,
 * which its preconditioner never achieves, so FG_DENSITY_OFF reproduces the
 * reference's results.  GATED applies it when our solve reaches its
 * tolerance; ON always applies it. */
enum fg_density_mode { FG_DENSITY_OFF = 0, FG_DENSITY_GATED = 1, FG_DENSITY_ON = 2 };

/* Which solver class steps the scene (SimulationMethod SIMULATION_LIQUID /
 * SIMULATION_NBFLIP, flipsolver2d.h; chosen by a scene file's simType,
 * Utils/jsonscenereader.cpp:50-62, 166-186).  FG_SIM_NBFLIP is
 * NBFlipSolver::step (nbflipsolver.cpp:26-64): particles only in a band
 * below the surface, the deep interior carried by semi-Lagrangian advection
 * of the grids and the level set; its first frame seeds no particles
 * (firstFrameInit, 203-211) and it has no density correction (density_mode
 * must be OFF).  Its constructor sets the solve tolerance to 1e-6 (22):
 * pass that as solver_tolerance to follow it. */
enum fg_sim_method { FG_SIM_FLIP = 0, FG_SIM_NBFLIP = 1 };

/* Stages, in FlipSolver::step order (flipsolver2d.cpp:412-462). */
enum fg_stage {
    FG_STAGE_ADVECT = 0,          /* advect(): RK4 + solid projection + kill   */
    FG_STAGE_REBIN = 1,           /* matrix snapshot, prune + sort on cell key */
    FG_STAGE_DENSITY = 2,         /* densityCorrection()                       */
    FG_STAGE_P2G = 3,             /* particleVelocityToGrid()                  */
    FG_STAGE_SDF = 4,             /* updateSdf() + updateMaterials()           */
    FG_STAGE_LEVELSET = 5,        /* extrapolateLevelsetInside()               */
    FG_STAGE_EXTRAPOLATE = 6,     /* m_fluidVelocityGrid.extrapolate(10)       */
    FG_STAGE_FORCES = 7,          /* saved grid copy + applyBodyForces()       */
    FG_STAGE_PROJECT = 8,         /* calcPressureRhs + solve + apply pressure
                                   * (viscosity_enabled: then updateVelocityFromSolids,
                                   * applyViscosity and project() again, 446-453) */
    FG_STAGE_SOLIDS = 9,          /* updateVelocityFromSolids()                */
    FG_STAGE_G2P = 10,            /* particleUpdate()                          */
    FG_STAGE_COUNT = 11,          /* countParticles() (+ reseed)               */
    FG_STAGE_NUM = 12
};

typedef struct fg_scene_params {
    int32_t size_i, size_j;       /* gridSizeI/J */
    double  dx;                   /* cell size (m) */
    double  fluid_density;
    float   gravity_i, gravity_j; /* globalAcceleration (m/s^2) */
    float   pic_ratio;
    float   cfl_number;
    float   particle_scale;
    int32_t fps;
    int32_t max_substeps;
    int32_t particles_per_cell;
    int32_t pcg_iter_limit;       /* cap on inner PCG iterations per refinement sweep */
    int32_t density_mode;         /* fg_density_mode */
    double  solver_tolerance;     /* FP64 relative residual ||b-Ax||2/||b||2 target */
    int64_t max_particles;        /* particle capacity */
    int32_t sim_method;           /* fg_sim_method */
    int32_t viscosity_enabled;    /* viscosityEnabled (flipsolver2d.h:52): the light viscosity model
                                   * (LightViscosityModel, viscositymodel.cpp:4-140) after the
                                   * projection, then a second projection (flipsolver2d.cpp:446-453).
                                   * FG_SIM_FLIP only, with density_mode OFF (step() skips density
                                   * correction with viscosity, 425); the heavy model is not built. */
} fg_scene_params;

/* Host-side view of a scene's state.  Arrays are row-major, index i*J+j for
 * cells; U is (I+1) x J, V is I x (J+1) (staggeredvelocitygrid.cpp:6-14).
 * Particle positions/velocities are in cell units (cells, cells/s).
 * Any pointer may be NULL on get (skipped).  On set, particles, u/v, material
 * and solid_sdf are required, and for an FG_SIM_NBFLIP scene fluid_sdf too
 * (its next substep advects the level set); u_valid/v_valid, solid_id and
 * emitter_id default as noted. */
typedef struct fg_scene_state {
    int64_t   n_particles;
    float    *px, *py, *vx, *vy;  /* [n_particles] */
    uint32_t *pid;                /* [n_particles] stable particle id */
    float    *u, *v;
    uint8_t  *u_valid, *v_valid;
    int8_t   *material;           /* I*J */
    float    *solid_sdf;          /* I*J */
    float    *fluid_sdf;          /* I*J */
    int32_t  *solid_id;           /* I*J, -1 = none */
    int32_t  *particle_counts;    /* I*J: countParticles' per-cell counts (at most 2*ppc).  On set
                                   * optional (NULL: zero, as firstFrameInit leaves them): an
                                   * NB-FLIP substep's band prune reads the previous ones. */
    float    *saved_u, *saved_v;  /* get only: grid after extrapolation, before forces */
    double   *pressure;           /* I*J, get only: last pressure solve */
    double   *rhs;                /* I*J, get only: last pressure right-hand side */
    int32_t  *emitter_id;         /* I*J, -1 = none: the source (index into the emitter table)
                                   * of each SOURCE cell; NULL on set: none */
    uint32_t  rng_epoch;          /* substeps taken: the reseed RNG's counter */
    uint32_t  next_pid;           /* id of the next reseeded particle (0 on set: max id + 1) */
    uint8_t  *stencil;            /* I*J, get only: the pressure matrix of the last solve
                                   * (getPressureProjectionMatrix, flipsolver2d.cpp:797-887), one
                                   * byte per cell: bits 0-3 fluid neighbour at i-1, i+1, j-1, j+1;
                                   * bits 4-6 nonsolid neighbour count n; bit 7 fluid row.  Row c of
                                   * A is s*(n x_c - sum of the flagged neighbours) on fluid rows
                                   * and x_c elsewhere, s = dt/(rho dx^2) of that substep. */
    float    *work_u, *work_v;    /* get only: the grid fluidVelocityGrid holds after the last
                                   * fg_scene_stage of an unfinished substep: after FG_STAGE_P2G the
                                   * transferred grid, after FG_STAGE_EXTRAPOLATE the extrapolated one,
                                   * after FG_STAGE_FORCES the one the projection reads, after
                                   * FG_STAGE_PROJECT the projected one (updateVelocityFromSolids
                                   * applied too: the two stages share a kernel); otherwise, like u
                                   * and v above, the substep's final, extrapolated grid. */
    float    *particle_viscosity; /* [n_particles] the particles' viscosity property (a scene with
                                   * viscosity_enabled; on set optional: NULL gives 0) */
    float    *viscosity;          /* I*J: the viscosity grid (centeredParamsToGrid,
                                   * flipsolver2d.cpp:1379-1431) of the last substep; on set
                                   * optional (NULL: 0): an NB-FLIP substep advects it */
} fg_scene_state;

typedef struct fg_stats {
    int32_t substeps;
    int32_t pressure_iters;       /* max over substeps of total inner PCG iterations */
    int32_t pressure_sweeps;      /* max over substeps of FP64 refinement sweeps */
    int32_t density_iters;
    int32_t density_applied;
    int32_t converged;            /* 1 if every solve reached solver_tolerance */
    double  pressure_rel_resid;   /* last solve's FP64 ||b-Ax||2/||b||2 */
    int64_t particles;            /* live particles after the call */
    int64_t particle_substeps;    /* sum over substeps of live particles at substep start */
    int64_t particles_killed;     /* advect: out of bounds / sink */
    int64_t particles_erased;     /* countParticles: beyond 2*ppc */
    float   last_dt;
    float   max_velocity;
    double  stage_ms[FG_STAGE_NUM]; /* filled when timing is enabled */
    double  frame_ms;
    int32_t status;               /* FG_OK, or FG_E_NUMERIC: the scene is quarantined */
    int32_t viscosity_iters;      /* max over substeps of the viscosity CG's iterations (U or V) */
    int64_t particles_reseeded;   /* reseedParticles: added in SOURCE cells */
} fg_stats;

typedef struct fg_ctx fg_ctx;
typedef struct fg_scene fg_scene;

const char* fg_last_error(void);
int32_t fg_abi_version(void);

int fg_ctx_create(int32_t device, uint64_t hbm_budget_bytes, fg_ctx** out);
int fg_ctx_destroy(fg_ctx* ctx);
int fg_ctx_bytes_in_use(fg_ctx* ctx, uint64_t* out);
/* Per-stage HIP-event timing for later calls (adds synchronisation). */
int fg_ctx_set_timing(fg_ctx* ctx, int32_t enabled);

/* fg_batch_step_frame may split a batch over up to FG_MAX_LANES lanes:
 * streams of this context, each stepped by its own host thread, with the
 * scenes cut by size into contiguous classes of about equal solver work.
 * Small scenes then overlap the latency-bound solve tails of large ones.
 * Results do not depend on the number of lanes.  Default 1.  No counterpart
 * in the reference (its thread pool steps one solver at a time). */
#define FG_MAX_LANES 4
int fg_ctx_set_batch_lanes(fg_ctx* ctx, int32_t lanes);
int fg_ctx_batch_lanes(fg_ctx* ctx, int32_t* lanes);
/* Solver tiling: each scene's solver kernels use at most target_blocks
 * blocks (1..4096) of at least min_chunk cells (a multiple of 256, >= 256).
 * Default 1024 / 1024.  Results are deterministic for a given tiling; the
 * FP64 reductions' summation order depends on it. */
int fg_ctx_set_solver_tiling(fg_ctx* ctx, int32_t target_blocks, int32_t min_chunk);
/* Preconditioner of the pressure PCG, for later solves.  FG_PRECOND_MG
 * (default): a symmetric multigrid V-cycle (Galerkin coarse operators,
 * bilinear transfers, L1-Jacobi smoothing), about 11-17 iterations to 1e-6
 * on dam-break systems from 128^2 to 512^2 cells.  FG_PRECOND_IPP: the
 * reference's inverse-Poisson preconditioner (inversepoissonpreconditioner.h:45-79)
 * on A's 5-point pattern, 140-565 iterations on the same systems.  Both
 * solve the same system to the same FP64 tolerance.  A scene whose grid
 * would need more than 12 levels uses the IPP. */
#define FG_PRECOND_MG 0
#define FG_PRECOND_IPP 1
int fg_ctx_set_preconditioner(fg_ctx* ctx, int32_t kind);
int fg_ctx_preconditioner(fg_ctx* ctx, int32_t* kind);
/* How later solves run the V-cycle's levels between level 1 and the coarse
 * block.  Fused (default, enabled != 0): per level one launch for the
 * residual and its restriction and one for the prolongation and the
 * post-smoothing, each over LDS tiles; otherwise two per-cell launches for
 * each.  The values and their arithmetic are the same, so solves are
 * bit-identical either way; the per-level form is kept for that check.  No
 * counterpart in the reference (its preconditioner is the IPP). */
int fg_ctx_set_vcycle_fused(fg_ctx* ctx, int32_t enabled);
/* Graph replay of PCG iterations: switched off.  enabled = 0 returns FG_OK;
 * any other value returns FG_E_INVALID.  Replaying each PCG iteration's
 * launches (the direction update, the residual update and the V-cycle,
 * about 20 kernels) from a captured HIP graph gave bit-identical results
 * but frames 2-3 % slower than direct launches on the batched and 1024^2
 * dam-break workloads, and a run under the profiler ended in a host-side
 * fault inside it whose cause was not found.  No counterpart in the
 * reference (its PCG loop runs on the host, linearsolver.cpp:25-73). */
int fg_ctx_set_solver_graphs(fg_ctx* ctx, int32_t enabled);
/* Autotuner for that schedule: steps `rounds` frames of the batch with each
 * lane count 1..FG_MAX_LANES (order alternating per round; the simulation
 * advances normally), measures particle-substeps per second for each, and
 * sets the fastest on the context (the most lanes among those within 3 % of
 * the fastest rate: each count runs on different frames).  *lanes_out gets it; rates_out (may be
 * NULL) gets FG_MAX_LANES measured rates. */
int fg_batch_autotune(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, int32_t rounds, int32_t* lanes_out,
                      double* rates_out);
/* Autotuner for the solver tiling (fg_ctx_set_solver_tiling): steps `rounds`
 * frames of the batch with each of n_cand (target_blocks, min_chunk) pairs
 * (candidates: 2 * n_cand int32; NULL: the built-in set FG_TILING_CANDIDATES,
 * min_chunk 256 .. 4096 at 1024 target blocks), order alternating per round,
 * measures particle-substeps per second for each and sets the fastest on
 * the context.  *best_out gets its index; rates_out (may be NULL) gets the
 * n_cand rates.  The tiling sets the FP64 reductions' summation order, so
 * these frames round differently from frames under one fixed tiling: tune
 * on a copy of the batch where results are compared.  No counterpart in
 * the reference (its thread pool splits loops by splitRange,
 * threading/threadpool.cpp:36-70). */
#define FG_TILING_CANDIDATES 5
int fg_batch_autotune_tiling(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, int32_t rounds,
                             const int32_t* candidates, int32_t n_cand, int32_t* best_out, double* rates_out);

/* Per-launch HIP-event timing of one kernel family, recorded on the stream
 * the kernel runs on (the launching lane's stream).  kernel = fg_kernel_id,
 * or -1 to switch off.  fg_ctx_profile_read returns the summed event time,
 * the number of timed launches and the summed work units of launches that
 * did work, then resets the counters.  Work units per family: [B] the
 * family's algorithmic bytes (counted by the library from the scenes'
 * particle, cell and face counts), [C] cells, [P] particles.  busy_ms is the
 * union of the launch intervals: the time some timed launch was running
 * (launches on several lanes overlap).  Algorithmic bytes (FP32 = 4 B, flag
 * = 1 B; P particles, C cells, F = F_U + F_V faces, C1 level-1 cells):
 *   P2G          16 P + 5 F            (positions, velocities; faces + validity)
 *   SDF          8 P + 6 C             (positions; fluid SDF + material)
 *   PARTICLES    36 P + 8 F + 4 C      (k_advect + k_g2p_count: positions and
 *                                       velocities read and written, keys; U, V,
 *                                       saved U, saved V; solid SDF)
 *   MG_DOWN0     5 C + 5 C1 per V-cycle (stencil code, r; level-1 activity, r1)
 *   MG_UP0       9 C + 4 C1 per V-cycle (code, r, z; level-1 y)
 *   MG_RESID     per V-cycle, over every level l between level 1 and the
 *                coarse block: 9 C_l + 100 S_l (r, dinv; r_{l+1} written,
 *                1 B a level-l cell; S_l the rows that read stored
 *                coefficients).  Per-level kernels: 12 C1 + 100 S1 (the
 *                level-1 residual launch: r, dinv, y)
 *   MG_PP        per V-cycle, over the same levels: 14 C_l + 100 S_l (act,
 *                r, dinv, y_{l+1}: 1 B a level-l cell; y written)
 *   MG_COARSE    per scene and V-cycle: 4 B per cell of the block's levels
 *                (dinv), its input (r_1 or the finer level's y), 4 B per
 *                top-level cell (y written), 100 B per row above the bottom
 *                that reads stored coefficients, 100 B per bottom row. */
enum fg_kernel_id {
    FG_K_ADVECT = 0,      /* [P] */
    FG_K_SORT = 1,        /* [P] the counting sort's launches */
    FG_K_P2G = 2,         /* [B] */
    FG_K_SDF = 3,         /* [B] */
    FG_K_LEVELSET = 4,    /* [C] */
    FG_K_EXTRAPOLATE = 5, /* [C] (faces) */
    FG_K_PCG_A = 6,       /* [C] per iteration (17 B a cell) */
    FG_K_PCG_B = 7,       /* [C] per iteration (32 B a cell) */
    FG_K_G2P = 8,         /* [P] */
    FG_K_RHS = 9,         /* [C] */
    FG_K_APPLY = 10,      /* [C] */
    FG_K_VCYCLE = 11,     /* [C] one multigrid V-cycle (all its launches) */
    FG_K_MG_RESID = 12,   /* [B] the V-cycle's residual + restriction launches, k_mg_rr of
                           * every level between level 1 and the coarse block (per-level
                           * kernels, see fg_ctx_set_vcycle_fused: the level-1 residual) */
    FG_K_MG_COARSE = 13,  /* [B] the V-cycle's coarse-levels launch (one block per scene) */
    FG_K_MG_DOWN0 = 14,   /* [B] level-0 smoothing, residual and restriction */
    FG_K_MG_UP0 = 15,     /* [B] prolongation to level 0 and post-smoothing */
    FG_K_PARTICLES = 16,  /* [B] k_advect and k_g2p_count together (G2P + advection) */
    FG_K_MG_PP = 17,      /* [B] the V-cycle's prolongation + post-smoothing launches, k_mg_pp
                           * of the same levels (not timed with the per-level kernels) */
    FG_K_NB_EULER = 18,   /* [B] NB-FLIP's Euler pass (k_nb_euler): U, V and the level set
                           * read once and their advected copies written, 8 (F_U + F_V + C) */
    FG_K_VISC = 19,       /* [B] the viscosity CG's iterations (k_visc_spmv, k_visc_update,
                           * k_visc_dir, k_visc_red; timed as one stretch with its polls):
                           * 216 B a cell and iteration for the U and V systems together */
    FG_K_NUM = 20
};
int fg_ctx_profile_kernel(fg_ctx* ctx, int32_t kernel);
/* Work units (as above) of every kernel family, summed over all launches
 * since the context was created or last reset (reset != 0 clears them),
 * whether or not a family is being timed: units[FG_K_NUM].  Divided into a
 * profiler's byte counters over the same launches, they give the HBM traffic
 * per algorithmic byte of each kernel. */
int fg_ctx_work_read(fg_ctx* ctx, int64_t* units, int32_t reset);
int fg_ctx_profile_read(fg_ctx* ctx, double* total_ms, double* busy_ms, int64_t* launches, int64_t* work_units);
/* Device-to-device copy bandwidth (read + write bytes / time) of a plain
 * copy kernel over `bytes` per buffer, best of `reps`. */
int fg_ctx_bandwidth_probe(fg_ctx* ctx, uint64_t bytes, int32_t reps, double* gbps, double* ms_per_copy);

/* Bytes of device memory a scene with these parameters needs. */
int fg_scene_bytes(const fg_scene_params* p, uint64_t* out);
int fg_scene_create(fg_ctx* ctx, const fg_scene_params* p, fg_scene** out);
int fg_scene_destroy(fg_scene* s);

/* Polygons: polygon k has nverts[k] >= 3 vertices, (x, y) pairs in metres
 * (x along I), concatenated in verts. */
typedef struct fg_polygons {
    int32_t n;
    const int32_t* nverts;
    const float* verts;
} fg_polygons;

/* A scene's objects (Obstacle, Emitter, Sink, initial fluid; the reference's
 * addGeometry / addSource / addSink / addInitialFluid). */
typedef struct fg_scene_objects {
    fg_polygons solids;
    const float* friction;        /* [solids.n] or NULL: obstacle friction */
    fg_polygons sources;
    const float* source_velocity; /* [2 * sources.n] m/s or NULL: Emitter::velocity */
    const uint8_t* source_transfer; /* [sources.n] or NULL: Emitter::velocityTransfer */
    fg_polygons sinks;
    fg_polygons fluids;
    uint32_t seed;                /* seeding and reseeding RNG */
    const float* source_viscosity; /* [sources.n] or NULL: Emitter::viscosity (reseeded particles,
                                    * SOURCE cells' viscosity) */
    const float* fluid_viscosity; /* [fluids.n] or NULL: Emitter::viscosity of the initial fluid
                                   * (updateInitialFluid's grid, which seeding lerps) */
} fg_scene_objects;

/* First-frame init on the device (FlipSolver::firstFrameInit): the cell
 * passes updateSinks, updateSources, updateSolids, updateInitialFluid in the
 * reference's order (materials, solid SDF and ids, emitter ids), then
 * seedInitialFluid: ppc particles per strict-fluid cell, particle k jittered
 * in quadrant k % 4 by a counter-based RNG, ids in (cell, k) order.  Resets
 * the grids, the quarantine and the reseed counter.  FG_E_CAPACITY if the
 * seeding exceeds max_particles (the scene is then empty).  An FG_SIM_NBFLIP
 * scene runs NBFlipSolver::firstFrameInit instead (nbflipsolver.cpp:203-211,
 * 291-349): sinks, sources, solids, then the fluid SDF of the initial fluid
 * (the cell centre's distance over dx, FLUID where negative); no particles;
 * the per-cell source distance its substeps take is computed here. */
int fg_scene_init_objects(fg_scene* s, const fg_scene_objects* objects);
/* fg_scene_init_objects with solids and initial fluid only. */
int fg_scene_init_geometry(fg_scene* s,
                           int32_t n_solids, const int32_t* solid_nverts, const float* solid_verts,
                           const float* friction,
                           int32_t n_fluids, const int32_t* fluid_nverts, const float* fluid_verts,
                           uint32_t seed);
/* The emitter table of the SOURCE cells (afterTransfer, flipsolver2d.cpp:390-410;
 * reseedParticles, 627-680): per emitter its velocity (m/s, 2 floats) and
 * whether it is transferred; the reseed RNG seed.  At most 256 emitters. */
int fg_scene_set_emitters(fg_scene* s, int32_t n, const float* velocity, const uint8_t* transfer, uint32_t seed);
/* Obstacle friction table (updateVelocityFromSolids, flipsolver2d.cpp:1077-1104). */
int fg_scene_set_friction(fg_scene* s, int32_t n, const float* friction);
int fg_scene_set_state(fg_scene* s, const fg_scene_state* st);
int fg_scene_get_state(fg_scene* s, fg_scene_state* st);
int fg_scene_particle_count(fg_scene* s, int64_t* out);
/* FlipSolver::maxParticleVelocity (flipsolver2d.cpp:1560-1574): the largest
 * particle speed (cells/s), as the next substep's CFL split reads it. */
int fg_scene_max_velocity(fg_scene* s, float* out);

/* One FlipSolver::step() with the given dt.  FG_E_NUMERIC: the scene is (or
 * has just been) quarantined. */
int fg_scene_substep(fg_scene* s, float dt, fg_stats* stats);
/* One FlipSolver::stepFrame(): CFL-split substeps until the frame is done. */
int fg_scene_step_frame(fg_scene* s, fg_stats* stats);
/* Run a single stage (fg_stage) of step() with dt; stages must be called in
 * order for a full step.  For parity tests of the reference's hooks.
 * FlipSolver's stage order only: an FG_SIM_NBFLIP scene returns FG_E_INVALID. */
int fg_scene_stage(fg_scene* s, int32_t stage, float dt);

/* Step every scene one frame; all scenes advance together, one launch per
 * phase.  stats may be NULL or an array of n.  Returns FG_E_NUMERIC when a
 * scene of the batch is quarantined (stats[k].status names it); every other
 * scene has completed its frame. */
int fg_batch_step_frame(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, fg_stats* stats);
/* One FlipSolver::step() of every scene, scene k with dts[k]; the scenes
 * advance together, one launch per phase (the substep fg_batch_step_frame
 * repeats).  Quarantine as in fg_batch_step_frame. */
int fg_batch_substep(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, const float* dts, fg_stats* stats);

/* Isolated pressure projection (LinearSolver::solve replacement).  The matrix
 * is the reference's 5-point operator assembled from `material`
 * (flipsolver2d.cpp:797-887) with scale = dt/(rho*dx^2); solved to
 * rel_tol (FP64 relative residual) with FP32 IPP-PCG + FP64 refinement. */
typedef struct fg_solve_stats {
    int32_t iterations;           /* total inner PCG iterations */
    int32_t sweeps;               /* FP64 refinement sweeps */
    int32_t converged;
    double  rel_residual;
    double  solve_ms;
} fg_solve_stats;
int fg_poisson_solve(fg_ctx* ctx, int32_t size_i, int32_t size_j, const int8_t* material, double scale,
                     const double* b, double* x, double rel_tol, int32_t max_iters, fg_solve_stats* stats);
/* The same solve with the matrix given as its bytes, one per cell in the
 * stencil format of fg_scene_state (bits 0-3 fluid neighbours at i-1, i+1,
 * j-1, j+1 by linear offset; bits 4-6 nonsolid count; bit 7 fluid row): the
 * rows of the reference's IndexedPressureParameters unit list
 * (pressuredata.h:97-151), for LinearSolver::solve callers that hold one. */
int fg_poisson_solve_matrix(fg_ctx* ctx, int32_t size_i, int32_t size_j, const uint8_t* matrix, double scale,
                            const double* b, double* x, double rel_tol, int32_t max_iters, fg_solve_stats* stats);
/* The same solve with no allocation in the call.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The solver's arrays live in a fixed 48 MiB device scratch array that the
 * library holds on each device it runs on until the process ends (its
 * address is resolved when the first context is created on the device); a
 * system whose arrays do not fit (about 600,000 cells) is refused with
 * FG_E_CAPACITY.  Calls on one device share that scratch and are serialised;
 * the scratch is not counted against the context's HBM budget.  Results are
 * those of fg_poisson_solve bit for bit. */
int fg_poisson_solve_noalloc(fg_ctx* ctx, int32_t size_i, int32_t size_j, const int8_t* material, double scale,
                             const double* b, double* x, double rel_tol, int32_t max_iters,
                             fg_solve_stats* stats);

/* ------------------------------------------------------------------------
 * Serving: continuous batching of scene simulations on one device.
 *
 * Replaces the reference's application loop that owns one FlipSolver per
 * scene and calls stepFrame() on it (AutoBench/benchmarkrunnerapplication.cpp:88-92,
 * Liquid2dRender/liquidrenderapp.cpp:250-253), for many scenes at once.
 * Requests queue in arrival order; each tick admits queued requests while
 * their scene memory (fg_scene_bytes) fits the server's HBM budget and a
 * resident slot is free, steps one frame of every resident scene in one
 * batched call (fg_batch_step_frame), and retires scenes that finished their
 * frames.  State stays on the device between frames.  One host thread per
 * server; calls on one server must not overlap.
 * ---------------------------------------------------------------------- */
typedef struct fg_server fg_server;

typedef struct fg_request {
    fg_scene_params params;
    int32_t n_solids;             /* solids as in fg_scene_init_geometry */
    const int32_t* solid_nverts;
    const float* solid_verts;
    const float* friction;        /* [n_solids] or NULL */
    int32_t n_fluids;
    const int32_t* fluid_nverts;
    const float* fluid_verts;
    uint32_t seed;
    int32_t frames;               /* frames to simulate (>= 1) */
    int32_t keep;                 /* 1: keep the scene after its last frame until fg_server_release */
    const float* fluid_viscosity; /* [n_fluids] or NULL: the fluids' viscosity (params.viscosity_enabled);
                                   * params.sim_method chooses FlipSolver or NBFlipSolver */
} fg_request;

enum fg_request_state {
    FG_REQ_QUEUED = 0, FG_REQ_RESIDENT = 1, FG_REQ_DONE = 2, FG_REQ_REJECTED = 3, FG_REQ_RELEASED = 4,
    FG_REQ_FAILED = 5,            /* quarantined while resident (error = FG_E_NUMERIC) */
    FG_REQ_CANCELLED = 6          /* fg_server_cancel */
};

typedef struct fg_request_status {
    int32_t state;                /* fg_request_state */
    int32_t frames_done;
    int64_t particles;            /* live particles after the last frame */
    int64_t particle_substeps;
    double  submit_ms, admit_ms, finish_ms;  /* server clock: ms since fg_server_create */
    double  max_frame_ms;         /* slowest frame of this request */
    int32_t error;                /* REJECTED / FAILED: the FG_E_* code that ended it */
    int32_t reserved;
} fg_request_status;

typedef struct fg_tick_stats {
    int32_t admitted, stepped, retired, queued, resident;
    int64_t particle_substeps;
    double  tick_ms;              /* admission + batched frame + retirement */
    double  admit_ms;             /* of which: scene creation and first-frame init */
    double  frame_ms_max;         /* slowest scene frame of the tick */
    uint64_t resident_bytes;
} fg_tick_stats;

typedef struct fg_server_summary {
    int64_t frames;               /* scene frames stepped */
    int64_t particle_substeps;
    double  busy_ms;              /* sum of tick durations */
    double  p50_frame_ms, p99_frame_ms, max_frame_ms;   /* per scene frame */
    double  p50_admit_wait_ms, p99_admit_wait_ms;       /* submit -> admit */
    int64_t submitted, completed, rejected;
    uint64_t peak_resident_bytes;
    int32_t peak_resident;
    int32_t reserved;
    int64_t failed, cancelled;
    /* ticks whose admission stopped with requests still queued because the
     * head did not fit the HBM budget (budget_waits) or the resident cap
     * (cap_waits) */
    int64_t budget_waits, cap_waits;
} fg_server_summary;

int fg_server_create(fg_ctx* ctx, uint64_t hbm_budget_bytes, int32_t max_resident, fg_server** out);
int fg_server_destroy(fg_server* srv);
/* Queue a request (geometry is copied).  A request whose scene alone
 * exceeds the budget is marked FG_REQ_REJECTED at once; one whose admission
 * fails for another reason (its seeding exceeds max_particles, the device
 * cannot hold it while no other scene is resident) is marked REJECTED when
 * it reaches the head of the queue, and the queue moves on. */
int fg_server_submit(fg_server* srv, const fg_request* req, int64_t* id_out);
/* Admit, step one frame of every resident scene, retire.  A scene
 * quarantined during the frame ends its request as FG_REQ_FAILED (its scene
 * is freed unless the request keeps it); the others go on.  Returns an
 * error only for device or state errors. */
int fg_server_tick(fg_server* srv, fg_tick_stats* st);
int fg_server_status(fg_server* srv, int64_t id, fg_request_status* out);
/* The scene of a resident or kept request (for fg_scene_get_state). */
int fg_server_scene(fg_server* srv, int64_t id, fg_scene** out);
/* Free the scene of a finished (DONE or FAILED) request that kept it. */
int fg_server_release(fg_server* srv, int64_t id);
/* Cancel a queued or resident request: it leaves the queue or the batch
 * before the next tick and its scene is freed (state FG_REQ_CANCELLED).  A
 * finished request that kept its scene is released instead. */
int fg_server_cancel(fg_server* srv, int64_t id);
int fg_server_summary_get(fg_server* srv, fg_server_summary* out);

#ifdef __cplusplus
}
#endif
#endif
