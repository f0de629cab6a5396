/*
 * flip_oracle.h -- CPU restatement of FlipSolver2d's per-substep FLIP step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it.  The product library (flipsolver2d_amd) never links or calls it.
 *
 * Every function restates one reference function; the citation names the
 * reference file and lines it follows (paths relative to the reference repo
 * root, FlipSolver2dLib/ unless noted).  Floating-point expressions keep the
 * reference's operation order and are compiled with -ffp-contract=off so the
 * HIP kernels (built the same way) can be compared bit for bit.
 *
 * Deliberate, documented departures from the reference (all forced by the
 * reference's container artefacts, not by its arithmetic):
 *   - particles are held as one array sorted by (cell, stable id) instead of
 *     3x3-cell bins in insertion order; float sums over particles therefore
 *     run in that canonical order;
 *   - countParticles keeps, per cell, the 2*ppc particles with the smallest
 *     ids (the reference keeps the first 2*ppc in its bin storage order);
 *   - the density-correction solve is selectable (off / gated / on); the
 *     reference's own solve never converges, so its gate never opens and
 *     "off" is result-identical to the reference.
 */
#ifndef FLIP_ORACLE_H
#define FLIP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SOLVER_REF_REPLICA = 0, ORC_SOLVER_CONVERGED = 1 };
enum { ORC_DENSITY_OFF = 0, ORC_DENSITY_GATED = 1, ORC_DENSITY_ON = 2 };
enum { ORC_SIM_FLIP = 0, ORC_SIM_NBFLIP = 1 };

/* Material codes, materialgrid.h:6-13. */
enum {
    ORC_FLUID = 0x40, ORC_SOURCE = 0x41, ORC_SOLID = 0x20,
    ORC_SINK = 0x12, ORC_EMPTY = 0x10
};

typedef struct OrcParams {
    int32_t size_i, size_j;        /* grid cells, I along x (rows), J along y */
    int32_t particles_per_cell;
    int32_t pcg_iter_limit;
    double  dx;                    /* cell size (m) */
    double  fluid_density;         /* rho */
    float   gravity_i, gravity_j;  /* globalAcceleration (m/s^2) */
    float   pic_ratio;
    float   cfl_number;
    float   particle_scale;
    int32_t fps;
    int32_t max_substeps;
    int32_t solver_mode;           /* ORC_SOLVER_* */
    int32_t density_mode;          /* ORC_DENSITY_* */
    int32_t threads;               /* splitRange thread count (replica reductions, CPU pool) */
    double  project_tol;           /* replica: max-abs tol; converged: relative L2 residual */
    int32_t n_obstacles;
    const float* friction;         /* [n_obstacles] */
    /* sources (Emitter, emitter.h): velocity (m/s, x along I) and whether
     * it is transferred to the grid and to reseeded particles */
    int32_t n_emitters;
    const float* emitter_velocity; /* [2 * n_emitters] */
    const uint8_t* emitter_transfer; /* [n_emitters] */
    uint32_t reseed_seed;          /* counter-based RNG seed of reseedParticles */
    int32_t sim_method;            /* ORC_SIM_FLIP (FlipSolver) or ORC_SIM_NBFLIP (NBFlipSolver) */
    const float* source_sdf;       /* NB-FLIP: I*J distance (cells) of the nearest source at the cell
                                    * corner, FLT_MAX without sources (orc_nb_init_objects); may be NULL */
    int32_t viscosity_enabled;     /* viscosityEnabled: the light model (viscositymodel.cpp:4-140) */
    const float* emitter_viscosity; /* [n_emitters] Emitter::viscosity, or NULL (0) */
    const float* fluid_viscosity;  /* [initial fluids] Emitter::viscosity, or NULL (0) */
    const int32_t* source_id;      /* NB-FLIP: I*J the source whose corner distance is nearest and
                                    * negative, else -1 (orc_nb_init_objects); may be NULL */
} OrcParams;

typedef struct OrcState {
    int64_t   n_particles;         /* in/out */
    int64_t   capacity;
    float    *px, *py, *vx, *vy;   /* [capacity] */
    uint32_t *pid;                 /* [capacity] stable particle id */
    float    *u, *v;               /* (I+1)*J, I*(J+1) */
    uint8_t  *u_valid, *v_valid;
    int8_t   *material;            /* I*J */
    float    *solid_sdf, *fluid_sdf;
    int32_t  *solid_id;            /* I*J, -1 = none */
    int32_t  *particle_counts;     /* I*J (out) */
    float    *saved_u, *saved_v;   /* out: grid after extrapolation, before forces */
    double   *pressure;            /* I*J out: last pressure solve */
    double   *rhs;                 /* I*J out: last pressure rhs */
    float    *density;             /* I*J out: density grid (density mode != off) */
    int32_t  *emitter_id;          /* I*J, -1 = none: the source of a SOURCE cell (may be NULL) */
    uint32_t  rng_epoch;           /* substeps taken: the reseed RNG's counter */
    uint32_t  next_pid;            /* id of the next reseeded particle */
    float    *pvisc;               /* [capacity] the particles' viscosity property, or NULL */
    float    *viscosity;           /* I*J the viscosity grid (centeredParamsToGrid), or NULL */
} OrcState;

typedef struct OrcStats {
    int32_t pressure_iters;
    int32_t density_iters;
    int32_t density_applied;
    int32_t substeps;
    double  pressure_rel_resid;    /* ||b-Ax||2/||b||2 of the returned pressure (FP64) */
    double  density_rel_resid;
    int64_t particles_killed;      /* advect: out of bounds / sink */
    int64_t particles_erased;      /* countParticles: beyond 2*ppc */
    float   last_dt;
    float   max_velocity;
    double  stage_ms[12];          /* SolverStage order, flipsolver2d.h:58-72 */
} OrcStats;

/* ---- interpolation & particle motion (grid2d.h, mathfuncs.cpp, sdfgrid.cpp) ---- */
float orc_bspline(float v);
float orc_quadratic_bspline(float x, float y);
float orc_linear_hat(float v);
void  orc_lerp(const float* g, int32_t si, int32_t sj, float off_i, float off_j,
               const float* xi, const float* xj, int64_t n, float* out);
void  orc_velocity_at(const float* u, const float* v, int32_t I, int32_t J,
                      const float* xi, const float* xj, int64_t n, float* ou, float* ov);
void  orc_rk4(const float* u, const float* v, int32_t I, int32_t J, float dt,
              const float* xi, const float* xj, int64_t n, float* oi, float* oj);
void  orc_closest_surface_point(const float* sdf, int32_t I, int32_t J,
                                const float* xi, const float* xj, int64_t n,
                                float* oi, float* oj);
/* Polygon signed distance, geometry2d.cpp:54-73 (verts as x,y pairs). */
float orc_polygon_sdf(const float* verts, int32_t nverts, float x, float y);

/* ---- particles ---- */
void  orc_advect(const OrcParams* p, OrcState* s, float dt, uint8_t* alive_out);
void  orc_sort_particles(const OrcParams* p, OrcState* s, const uint8_t* alive, int32_t* cell_start);
void  orc_cell_start(const OrcParams* p, const OrcState* s, int32_t* cell_start);
float orc_max_particle_velocity(const OrcState* s);
int32_t orc_count_particles(const OrcParams* p, OrcState* s, const int32_t* cell_start);
void  orc_g2p(const OrcParams* p, OrcState* s);

/* ---- grid transfer & rebuild ---- */
void  orc_p2g(const OrcParams* p, OrcState* s, const int32_t* cell_start);
void  orc_fluid_sdf(const OrcParams* p, OrcState* s, const int32_t* cell_start);
void  orc_update_materials(const OrcParams* p, OrcState* s);
void  orc_extrapolate_levelset_inside(float* sdf, int32_t I, int32_t J);
void  orc_extrapolate_levelset_outside(float* sdf, int32_t I, int32_t J);
void  orc_extrapolate(float* g, uint8_t* valid, int32_t si, int32_t sj, int32_t radius);
void  orc_body_forces(const OrcParams* p, OrcState* s, float dt);
void  orc_pressure_rhs(const OrcParams* p, const OrcState* s, double* rhs);
void  orc_apply_pressure(const OrcParams* p, OrcState* s, float dt, const double* pr);
void  orc_velocity_from_solids(const OrcParams* p, OrcState* s);
void  orc_density_grid(const OrcParams* p, OrcState* s, const int32_t* cell_start);
void  orc_density_rhs(const OrcParams* p, const OrcState* s, float dt, double* rhs);
void  orc_adjust_particles(const OrcParams* p, OrcState* s, float dt, const double* pr);

/* ---- pressure matrix & solvers (flipsolver2d.cpp:797-945, pressuredata.h, linearsolver.cpp) ---- */
/* Per cell: nonsolid count n (0..4) and fluid-neighbour mask (bits i-,i+,j-,j+); fluid flag. */
void  orc_build_matrix(const int8_t* material, int32_t I, int32_t J,
                       uint8_t* n_out, uint8_t* mask_out, uint8_t* fluid_out);
void  orc_spmv(const uint8_t* n, const uint8_t* mask, const uint8_t* fluid,
               int32_t I, int32_t J, double scale, const double* x, double* y);
void  orc_ipp_ref_apply(const int8_t* material_a, const uint8_t* fluid, int32_t I, int32_t J,
                        double scale, const double* x, double* y);
void  orc_ipp_apply(const uint8_t* n, const uint8_t* mask, const uint8_t* fluid,
                    int32_t I, int32_t J, const double* x, double* y);
int32_t orc_pcg_replica(const int8_t* material_a, int32_t I, int32_t J, double scale,
                        const double* b, double* x, int32_t iter_limit, double tol,
                        int32_t threads, double* err_trace);
int32_t orc_pcg_converged(const int8_t* material_a, int32_t I, int32_t J, double scale,
                          const double* b, double* x, int32_t iter_limit, double rel_tol,
                          double* rel_resid_out);
double orc_rel_residual(const int8_t* material_a, int32_t I, int32_t J, double scale,
                        const double* b, const double* x);
/* Material predicates as bits: 1 fluid, 2 strict fluid, 4 empty, 8 solid,
 * 16 source, 32 sink (materialgrid.h:15-20). */
int32_t orc_material_bits(int8_t code);
/* MaterialGrid::nonsolidNeighborCount(ssize_t linIdx), materialgrid.cpp:129-140 */
int32_t orc_nonsolid_count_lin(const int8_t* material, int32_t I, int32_t J, int64_t lin);
/* VOps::maxAbs / VOps::dot as the replica solver evaluates them (vmath.cpp:25-45, 100-136) */
double orc_maxabs_replica(const double* v, int64_t n, int32_t threads);
double orc_dot_replica(const double* a, const double* b, int64_t n, int32_t threads);
/* splitRange, threading/threadpool.cpp:41-76: fills starts/ends, returns count. */
int32_t orc_split_range(int64_t length, int64_t min_size, int64_t jobs_per_thread,
                        int32_t threads, int64_t* starts, int64_t* ends);

/* ---- composition ---- */
float   orc_frame_dt(float vel, float cfl, float frame_dt, float substep_time,
                     int32_t substep_count, int32_t max_substeps, int32_t* finished);
int32_t orc_substep(const OrcParams* p, OrcState* s, float dt, OrcStats* st);
int32_t orc_step_frame(const OrcParams* p, OrcState* s, OrcStats* st);

/* ---- first-frame init restated for dam-break style scenes ---- */
/* Solid SDF / id / material from polygons (flipsolver2d.cpp:502-590). */
void  orc_init_geometry(const OrcParams* p, OrcState* s,
                        int32_t n_solids, const int32_t* solid_nverts, const float* solid_verts,
                        int32_t n_fluids, const int32_t* fluid_nverts, const float* fluid_verts);
/* FlipSolver::firstFrameInit's cell passes in its order (flipsolver2d.cpp:788-795):
 * updateSinks, updateSources, updateSolids, updateInitialFluid (502-590). */
void  orc_init_objects(const OrcParams* p, OrcState* s,
                       int32_t n_solids, const int32_t* solid_nverts, const float* solid_verts,
                       int32_t n_sources, const int32_t* source_nverts, const float* source_verts,
                       int32_t n_sinks, const int32_t* sink_nverts, const float* sink_verts,
                       int32_t n_fluids, const int32_t* fluid_nverts, const float* fluid_verts);
/* FlipSolver::afterTransfer (flipsolver2d.cpp:390-410): the source velocity
 * on the faces of SOURCE cells whose emitter transfers it. */
void  orc_after_transfer(const OrcParams* p, OrcState* s);
/* FlipSolver::reseedParticles (flipsolver2d.cpp:627-680): ppc/2 - count new
 * particles in each SOURCE cell, appended in cell order; returns how many. */
int64_t orc_reseed(const OrcParams* p, OrcState* s);
/* Uniform in [0,1) of the reseed RNG for substep `epoch` (replaces
 * jitteredPosInCell's std::mt19937 stream, flipsolver2d.cpp:1013-1019). */
float orc_reseed_jitter(uint32_t seed, uint32_t epoch, uint32_t cell, uint32_t k, uint32_t axis);
/* Quadrant-jittered seeding with the counter RNG shared with the product's init. */
int64_t orc_seed_particles(const OrcParams* p, OrcState* s, uint32_t seed);
float   orc_jitter(uint32_t seed, uint32_t cell, uint32_t k, uint32_t axis);

/* ---- viscosity, the light model (viscositymodel.cpp:4-140) ---- */
/* FlipSolver::centeredParamsToGrid's viscosity (flipsolver2d.cpp:1379-1431)
 * and afterTransfer's SOURCE cells (396-397): into s->viscosity. */
void  orc_viscosity_grid(const OrcParams* p, OrcState* s, const int32_t* cell_start);
/* LightViscosityModel::getMatrix (viscositymodel.cpp:63-134): per cell the
 * diagonal, and the symmetric couplings to (i+1, j) (east) and (i, j+1)
 * (north) as the assembly's last write leaves them. */
void  orc_viscosity_matrix(const int8_t* material, const float* viscosity, int32_t I, int32_t J, float dt,
                           double* diag, double* east, double* north);
/* LightViscosityModel::apply (4-61): U then V solved with the same matrix
 * (Jacobi-preconditioned CG in FP64 to relative residual rel_tol, x0 = 0);
 * returns the larger iteration count. */
int32_t orc_viscosity_apply(const OrcParams* p, OrcState* s, float dt, double rel_tol);

/* ---- NB-FLIP (NBFlipSolver, nbflipsolver.cpp) ---- */
/* FlipSolver::eulerAdvectionThread (flipsolver2d.cpp:340-351): out[i,j] =
 * in lerped (with its offset) at RK4(-dt) from the grid index (i, j). */
void  orc_euler_advect(const float* u, const float* v, int32_t I, int32_t J, float dt,
                       const float* in, int32_t si, int32_t sj, float off_i, float off_j, float* out);
/* NBFlipSolver::pruneNarrowBand (nbflipsolver.cpp:228-252) over the particles
 * advect left alive (positions already advected). */
void  orc_nb_prune(const OrcParams* p, const OrcState* s, uint8_t* alive);
/* updateGridFromSources' distance + combineAdvectedGrids (nbflipsolver.cpp:
 * 320-419): level set, then the faces by the combined level set. */
void  orc_nb_combine(const OrcParams* p, OrcState* s, const float* adv_u, const float* adv_v,
                     const float* adv_sdf);
/* NBFlipSolver::reseedParticles (nbflipsolver.cpp:115-201); returns how many. */
int64_t orc_nb_reseed(const OrcParams* p, OrcState* s);
/* NBFlipSolver::step (nbflipsolver.cpp:26-64). */
int32_t orc_nb_substep(const OrcParams* p, OrcState* s, float dt, OrcStats* st);
/* NBFlipSolver::firstFrameInit (nbflipsolver.cpp:203-211): sinks, sources,
 * solids, then the fluid SDF of the initial fluid (291-318); no particles.
 * source_sdf_out (I*J) gets updateGridFromSources' per-cell distance. */
void  orc_nb_init_objects(const OrcParams* p, OrcState* s,
                          int32_t n_solids, const int32_t* solid_nverts, const float* solid_verts,
                          int32_t n_sources, const int32_t* source_nverts, const float* source_verts,
                          int32_t n_sinks, const int32_t* sink_nverts, const float* sink_verts,
                          int32_t n_fluids, const int32_t* fluid_nverts, const float* fluid_verts,
                          float* source_sdf_out, int32_t* source_id_out);
/* NBFlipSolver::centeredParamsToGridThread (nbflipsolver.cpp:441-504): the
 * viscosity grid from the particles, |w| > 1e-6, 0 where no weight. */
void  orc_nb_viscosity_grid(const OrcParams* p, OrcState* s, const int32_t* cell_start);
/* afterTransfer's SOURCE cells, updateGridFromSources' viscosity (to both
 * grids) and combineCenteredGrids (nbflipsolver.cpp:338-346, 400-419). */
void  orc_nb_combine_viscosity(const OrcParams* p, OrcState* s, float* adv_visc);

#ifdef __cplusplus
}
#endif
#endif
