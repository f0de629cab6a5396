// fg_device.h -- scene descriptor and device helpers shared by the kernels.
//
// All kernels are compiled with -ffp-contract=off so each float expression
// rounds exactly as written; the helpers below keep the operation order of
// the reference functions they cite (paths relative to the reference repo).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

constexpr int kBlock = 256;             // threads per block (4 wave64s)
constexpr uint32_t kDead = 0xffffffffu;  // particle key of a removed particle
constexpr int kMaxLevel = 0x3fffffff;    // "unreached" BFS marker
constexpr int kMaxSweeps = 8;            // FP64 refinement sweeps per solve
constexpr int kMaxBands = 512;           // level-set bands of 16 levels
constexpr int kSlots = 5;                // FP64 reduction slots
constexpr int kSlotBB = 0, kSlotPQ = 1, kSlotCG = 2 /* and 3: r.z, r.r */, kSlotRR64 = 4;
constexpr int kPcgMaxBlocks = 1024;      // default solver blocks per scene (contiguous cell chunks)
constexpr int kPcgMinChunk = 1024;       // default cells per solver block at least (4 per thread)
constexpr int kPcgTilingMaxBlocks = 4096;  // finest tiling a context may select
constexpr int kPcgTilingMinChunk = 256;
constexpr int kSdfLanes = 4;             // lanes per 3x3-cell bin in the fluid SDF (two or three window rows each)
// multigrid preconditioner
constexpr int kPrecondMG = 0, kPrecondIPP = 1;
constexpr int kMgMaxLevels = 12;
constexpr int kMgBottomMax = 16;         // the bottom level has at most this many cells per side
constexpr int kD0Tile = 16;              // k_mg_down0: level-1 cells a block, per side
// blocks a scene of k_mg_down0: kD0Tile x kD0Tile tiles of the level-1 grid
__host__ __device__ inline int mg_d0_tiles(int I1, int J1)
{
    return ((I1 + kD0Tile - 1) / kD0Tile) * ((J1 + kD0Tile - 1) / kD0Tile);
}
constexpr int kVcTile = 16;              // k_mg_pp: level-l cells a block, per side
constexpr int kRRTile = 8;               // k_mg_rr: level-(l+1) cells a block, per side
// blocks of a fused V-cycle launch over a level of I x J cells in tiles of t x t
__host__ __device__ inline int mg_tiles(int I, int J, int t = kVcTile)
{
    return ((I + t - 1) / t) * ((J + t - 1) / t);
}
constexpr int kMgBottomSweeps = 8;       // L1-Jacobi sweeps on the bottom level
constexpr float kMgOmega = 1.5f;         // L1-Jacobi weight: x += w (r - A x) / sum_k |a_k| (SPD for w < 2)
constexpr int kMgStencil = 25;           // 5x5 Galerkin stencil of the coarse levels
constexpr int kMgRowStride = 32;         // floats per stored coefficient row: one 128-byte line, 25 used
// k_mg_coarse runs levels t .. n-1 of a scene in one block, a thread per
// cell: t is the first level >= 1 with at most kMgCoarseTop cells whose
// levels t .. n-1 hold at most kMgCoarseCells cells together (their r stays
// in LDS).  A larger top (4096 cells, four per thread) left the block
// waiting on one global-load round trip per cell in turn.
constexpr int kMgCoarseThreads = 1024;
constexpr int kMgCoarseTop = kMgCoarseThreads;
constexpr int kMgCoarseCells = 2 * kMgCoarseThreads;
constexpr int kFixupLongCap = 512;       // longest cell the wave-cooperative id sort takes

// Material codes and tests, FlipSolver2dLib/materialgrid.h:6-20.
enum : int8_t { M_FLUID = 0x40, M_SOURCE = 0x41, M_SOLID = 0x20, M_SINK = 0x12, M_EMPTY = 0x10 };
__device__ __forceinline__ bool is_fluid(int8_t m) { return (m & 0x40) != 0; }
__device__ __forceinline__ bool is_strict_fluid(int8_t m) { return m == M_FLUID; }
__device__ __forceinline__ bool is_empty(int8_t m) { return (m & 0x10) != 0; }
__device__ __forceinline__ bool is_solid(int8_t m) { return (m & 0x20) != 0; }
__device__ __forceinline__ bool is_sink(int8_t m) { return m == M_SINK; }

// Scalars a scene's kernels exchange on the device (reductions, flags).
struct SceneScalars {
    uint32_t max_v2_bits;   // max |v|^2 over live particles (float bits, >= 0)
    int32_t n_particles;    // live particles (sorted prefix of the particle arrays)
    int32_t killed;         // advect kills this substep
    int32_t erased;         // countParticles erasures this substep
    // solver
    double bb;              // ||b||^2
    double rr64;            // ||b - A x||^2 (FP64 residual of current x)
    double rz;              // inner r.z
    double rz_new;
    double pq;              // inner p.q
    double rr0;             // inner ||r||^2 at sweep start
    double rr;              // inner ||r||^2
    double gamma[2];        // r.z of iteration it (parity slots)
    double inner_eps2;      // inner stop: ||r||^2 <= eps2 * ||r0||^2
    int32_t inner_done;     // inner PCG stopped (converged or capped)
    int32_t outer_done;     // FP64 residual reached tolerance
    int32_t inner_iters;    // iterations of the current sweep
    int32_t total_iters;    // iterations of the current solve
    int32_t sweeps;
    int32_t numeric_bad;    // NaN/Inf seen in a reduction
    int32_t ls_bands;       // level-set BFS bands that ran this substep
    int32_t density_applied;  // density correction moved the particles this substep
    int32_t mg_stored[kMgMaxLevels];  // active rows per level that read stored coefficients
    int32_t reseeded;       // particles reseedParticles added this substep
    int32_t reseed_dropped; // reseeds beyond the particle capacity
    int32_t sdf_known;      // some cell's fluid SDF is positive (extrapolateLevelsetInside has a known cell)
};

// One level of the multigrid hierarchy (levels >= 1; level 0 is the
// pressure matrix itself, read from SceneDev::stencil).
struct MgLevel {
    int32_t I, J;
    uint8_t* act;     // the cell takes part: one of its 2x2 finer cells does
    float* coef;      // [I*J][kMgRowStride], row by row: entry k couples the cell to (i + k/5 - 2, j + k%5 - 2)
    float* dinv;      // kMgOmega / sum_k |coef_k| on active cells, +0 on inactive ones (-0 on an
                      // active cell whose row is all zero, so that a nonzero bit
                      // pattern marks exactly the active cells); negative on rows
                      // equal to the level's canonical interior stencil
    float* r;         // restricted residual
    float* x;         // pre-smoothed iterate plus the coarse correction
    float* y;         // residual on the way down, smoothed result on the way up
};
struct MgLevels {
    int32_t n;        // levels; level n-1 is the bottom
    int32_t t;        // first level of k_mg_coarse (1 <= t <= n-1)
    int32_t tcells;   // cells of levels t .. n-1 together
    // offset of level l >= t among the cells of levels t .. n-1: in
    // k_mg_coarse's LDS, and of its dinv from lv[t].dinv (the levels' dinv
    // arrays are contiguous in level order)
    int32_t roff[kMgMaxLevels];
    MgLevel lv[kMgMaxLevels];
    // the Galerkin row of a cell deep inside the liquid (all of its fine
    // support and coarse neighbourhood active), per level: such rows read
    // these instead of their stored coefficients
    float canon[kMgMaxLevels][kMgStencil];
};

// Per-scene device descriptor.  One array of these lives in device memory;
// kernels index it with blockIdx.y (the scene's slot in the launch).
struct SceneDev {
    int32_t I, J;           // cells
    int32_t ppc;
    int32_t n_friction;
    int32_t friction_any;   // some friction value is nonzero (else the friction factor is exactly 1)
    float pic, pscale;
    float g_i, g_j;         // gravity (m/s^2)
    double dx, rho;
    double tol;             // FP64 relative residual target
    int32_t max_inner;      // inner iteration cap per sweep
    int32_t active;         // 0: skipped by every kernel of a batched launch
    int32_t sdf_short;      // particleScale > sqrt(0.5): a cell holding a particle is inside
    int32_t density_mode;   // 0 off, 1 gated on convergence, 2 on
    float cell_vol;         // (float)(dx^3): updateDensityGridThread's cell volume
    float pmass;            // rho * cell_vol / ppc as a float
    float dens_scale;       // (float)(dt*dt / (rho dx^2)): adjustParticlesByDensity
    double dens_rhs_scale;  // 1.0 / dt: calcDensityCorrectionRhs
    int32_t max_blocks;     // stride of the FP64 partials array (>= pcg_blocks)
    int32_t pcg_blocks;     // solver blocks of this scene
    int32_t pcg_chunk;      // cells per solver block (multiple of kBlock)
    int32_t max_pblocks;    // stride of the particle partials arrays
    int32_t pblocks;        // particle blocks of this scene's current launches (<= max_pblocks)
    int32_t precond;        // kPrecondMG or kPrecondIPP
    MgLevels* mg;           // the scene's multigrid hierarchy (device)
    // per-substep
    float dt;
    float force;            // (float)(dt / dx), applyBodyForces
    double sA;              // dt / (rho dx^2): matrix scale
    double sApply;          // dt / (rho dx): applyPressure scale
    double sRhs;            // 1.f / dx: calcPressureRhs scale
    int64_t cap;            // particle capacity
    // particles: canonical (sorted) and advected (unsorted) buffers
    float2 *pos, *vel;
    uint32_t* pid;
    float2 *posA, *velA;
    uint32_t* key;
    uint32_t *tmpPid, *tmpSrc;
    // cells
    int32_t *cellCount, *cellStart;  // C, C+1
    int32_t* scanSums;      // per scan-tile totals
    int8_t* mat;
    uint8_t* stencil;       // matrix snapshot: bits0-3 fluid-nbr mask, 4-6 n, 7 fluid row
    float *ssdf, *fsdf;
    int32_t* sid;
    int32_t* pcount;
    int32_t* marker;        // level-set BFS levels
    int32_t* lsFlags;       // [kMaxBands] band reached its last level
    uint8_t* lsTile;        // [2][BFS tiles] by band parity: the tile reached its band's last level
                            // (band 0 reads the tiles holding a known cell)
    // faces: working (P2G/project), saved (extrapolated, pre-force), final
    float *u, *v, *su, *sv, *uf, *vf;
    uint8_t *uval, *vval, *uvalS, *vvalS, *uvalF, *vvalF;
    // solver
    double *b, *x64;
    // CG vectors: cr/cr2 residual (double-buffered), cu = z, cp/cs1 direction
    // (double-buffered), cw = q = A' p; the iterate is x64
    float *cr, *cr2, *cw, *cs1, *cu, *cp;
    double* partials;       // per-block partial sums [kSlots * max_blocks]
    int32_t* ipart;         // per-block particle counters [2 * max_pblocks]: killed, erased
    float* fpart;           // per-block max |v|^2 [max_pblocks]
    float* friction;
    SceneScalars* sc;
    // sources: emitter id per cell (-1 none), per emitter the source
    // velocity as face values (vx/dx, vy/dx) and its transfer flag
    int32_t* eid;
    float4* emitters;       // [kMaxEmitters]: (u, v, transfer, 0)
    int32_t n_emitters;
    int32_t has_sources;    // some cell is a SOURCE cell (fixed after set-up)
    uint32_t rseed;         // reseed RNG seed
    uint32_t epoch;         // substeps taken before this one (the reseed RNG's counter)
    uint32_t next_pid;      // id of this substep's first reseeded particle
    // NB-FLIP (NBFlipSolver): the step's semi-Lagrangian copies of the grid
    // and the level set, and per cell the distance (cells) of the nearest
    // source at the cell corner (FLT_MAX without sources)
    int32_t nb;
    float *adv_u, *adv_v, *adv_sdf, *src_sdf;
    // viscosity (the light model, viscositymodel.cpp:4-140): the particles'
    // viscosity (sorted, advected copy), the cell grid; the matrix (diagonal,
    // couplings to i+1 and j+1) and the CG vectors of the U and V systems
    // ([2][C]: x, r, p, q), per-block partials and the CG scalars
    int32_t visc_on;
    float *pvisc, *pviscA, *cvisc;
    double *vd, *ve, *vn, *vx, *vr, *vp, *vq;
    double* vpart;          // [kViscSlots][max cell blocks]
    double* vsc;            // kViscScalars
    // NB-FLIP with viscosity: the advected viscosity grid, and per cell the
    // source whose corner distance is nearest and negative (-1: none)
    float* adv_visc;
    int32_t* src_id;
};
// viscosity CG scalars (per system s = 0 (U), 1 (V)): bb, rz, pq, rr, alpha,
// beta, done at index 2 k + s; the iteration counts at 14, 15
constexpr int kViscSlots = 4;         // per-block partials: two dots x two systems
constexpr int kViscScalars = 16;
constexpr int kMaxEmitters = 256;
// NBFlipSolver's bands (nbflipsolver.cpp:18-20), in cells of level set
constexpr float kNbNarrowBand = -3.f, kNbCombinationBand = -2.f, kNbResamplingBand = -1.f;

// counter-based uniform in [0,1): 24 bits of a murmur3-finalised hash of
// (seed, cell, k, axis); the same function seeds and reseeds particles
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__host__ __device__ __forceinline__ float jitter(uint32_t seed, uint32_t cell, uint32_t k, uint32_t axis)
{
    const uint32_t h = fmix32(seed ^ fmix32(cell * 0x9e3779b1u ^ fmix32(k * 2u + axis + 0x632be5abu)));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
// the reseed RNG's seed for one substep
__host__ __device__ __forceinline__ uint32_t reseed_key(uint32_t seed, uint32_t epoch)
{
    return fmix32(seed ^ fmix32(epoch * 0x9e3779b1u + 0x85ebca77u));
}

// ---------------------------------------------------------------------------
// Arithmetic helpers (FlipSolver2dLib/mathfuncs.cpp, grid2d.h)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fclamp(float v, float lo, float hi)
{
    return (v < lo) ? lo : ((hi < v) ? hi : v);   // std::clamp
}
__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// mathfuncs.cpp:3-6 frac(v) = v - (long)v  (truncf is exact)
__device__ __forceinline__ float frac_(float v) { return v - truncf(v); }
// mathfuncs.cpp:8-11
__device__ __forceinline__ int integr(float v) { return (int)floorf(v); }
// mathfuncs.cpp:13-16
__device__ __forceinline__ float lerp1(float a, float b, float f) { return (a * (1.0f - f)) + (b * f); }

// Grid2d::getAt with OOB_EXTEND (grid2d.h:122-143)
__device__ __forceinline__ float get_ext(const float* g, int si, int sj, int i, int j)
{
    return g[(size_t)iclamp(i, 0, si - 1) * sj + iclamp(j, 0, sj - 1)];
}
__device__ __forceinline__ int8_t mat_ext(const int8_t* m, int I, int J, int i, int j)
{
    return m[(size_t)iclamp(i, 0, I - 1) * J + iclamp(j, 0, J - 1)];
}

// Grid2d::lerp, grid2d.h:186-216
__device__ __forceinline__ float lerp_grid(const float* g, int si, int sj, float oi, float oj, float i, float j)
{
    i += oi;
    j += oj;
    i = fclamp(i, 0.f, (float)(si - 1));
    j = fclamp(j, 0.f, (float)(sj - 1));
    const int ci = integr(i), cj = integr(j);
    const float fi = frac_(i), fj = frac_(j);
    const int nj = fj >= 0.5f ? cj + 1 : cj - 1;
    const int ni = fi >= 0.5f ? ci + 1 : ci - 1;
    const float iF = fi < 0.5f ? 0.5f - fi : fi - 0.5f;
    const float jF = fj < 0.5f ? 0.5f - fj : fj - 0.5f;
    const float v1 = lerp1(get_ext(g, si, sj, ci, cj), get_ext(g, si, sj, ni, cj), iF);
    const float v2 = lerp1(get_ext(g, si, sj, ci, nj), get_ext(g, si, sj, ni, nj), iF);
    return lerp1(v1, v2, jF);
}

// StaggeredVelocityGrid::velocityAt (staggeredvelocitygrid.cpp:158-167)
__device__ __forceinline__ float2 velocity_at(const float* u, const float* v, int I, int J, float x, float y)
{
    return make_float2(lerp_grid(u, I + 1, J, 0.5f, 0.f, x, y), lerp_grid(v, I, J + 1, 0.f, 0.5f, x, y));
}

// FlipSolver::rk4Integrate, flipsolver2d.cpp:1195-1203 (Vec3 arithmetic of
// geometry2d.cpp:3-27: a scalar times a vector scales each component)
__device__ __forceinline__ float2 rk4(const float* u, const float* v, int I, int J, float2 p, float dt)
{
    float2 a = velocity_at(u, v, I, J, p.x, p.y);
    const float k1x = a.x * dt, k1y = a.y * dt;
    a = velocity_at(u, v, I, J, p.x + k1x * 0.5f, p.y + k1y * 0.5f);
    const float k2x = a.x * dt, k2y = a.y * dt;
    a = velocity_at(u, v, I, J, p.x + k2x * 0.5f, p.y + k2y * 0.5f);
    const float k3x = a.x * dt, k3y = a.y * dt;
    a = velocity_at(u, v, I, J, p.x + k3x, p.y + k3y);
    const float k4x = a.x * dt, k4y = a.y * dt;
    const float sx = ((k1x + k2x * 2.f) + k3x * 2.f) + k4x;
    const float sy = ((k1y + k2y * 2.f) + k3y * 2.f) + k4y;
    const float sixth = 1.0f / 6.0f;
    return make_float2(p.x + sx * sixth, p.y + sy * sixth);
}

// simmath::bSpline / quadraticBSpline, mathfuncs.cpp:32-43.  The reference
// multiplies each piece by a 0/1 factor and adds them; exactly one piece
// survives (beyond 1.5 both are zero and the sum is +0), so selecting the
// piece gives the same float bit for bit for every finite |value| < 2^64
// (beyond, value^2 overflows and the reference's form gives NaN, as it does
// for NaN; the select gives +0 there).  The outer piece
// is (0.5 b) b with b = max(1.5 - a, 0): b = 1.5 - a below 1.5, and +0 from
// 1.5 up and for NaN (fmaxf returns the other operand), where the reference
// gives +0 as well; so one select remains.  Both pieces are made before it
// (the empty asm pins them): a select, not divergent branches.  Checked
// against the CPU restatement of the reference on the host.
// bSpline(0) = 0.75.
__device__ __forceinline__ float bspline(float value)
{
    const float a = fabsf(value);
    const float b = fmaxf(1.5f - a, 0.0f);
    float t1 = 0.75f - a * a;
    float t2 = (0.5f * b) * b;
    asm volatile("" : "+v"(t1), "+v"(t2));
    return a < 0.5f ? t1 : t2;
}
__device__ __forceinline__ float qbspline(float x, float y) { return (bspline(x) * bspline(y)) * 0.75f; }

// simmath::linearHat / bilinearHat, mathfuncs.cpp:61-77
__device__ __forceinline__ float linear_hat(float v)
{
    if (v >= 0 && v <= 1) return 1 - v;
    if (v < 0 && v >= -1) return 1 + v;
    return 0;
}
__device__ __forceinline__ float bilinear_hat(float x, float y)
{
    return (linear_hat(x) * linear_hat(y)) * linear_hat(0.f);
}

// Array pointers the kernels read from descriptors are generic (flat) to the
// compiler; cast to the global address space they load and store with
// global instructions (shorter latency, no LDS-counter waits).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> gbl(T* p)
{
    return (gptr<T>)p;
}

// p[i] of a float2 array as one 64-bit global load (a vector-type object
// cannot be copied out of the global address space directly)
__device__ __forceinline__ float2 ldg2(const float2* p, int i)
{
    const unsigned long long v = gbl<const unsigned long long>(reinterpret_cast<const unsigned long long*>(p))[i];
    return make_float2(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
}

// Read-only data of a launch (descriptors, the multigrid level table, the
// scalars earlier kernels wrote) read through the constant address space:
// uniform addresses become scalar loads, which issue together.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> cst(const T* p)
{
    return (cptr<T>)p;
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// run of tiles so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int full = nb & ~7;
    if (b >= full) return b;
    const int per = full >> 3;
    return (b & 7) * per + (b >> 3);
}

// The (scene slot, block within the scene) a block works on.  A launch with
// a block map runs a flat grid over only the blocks that have work: entry
// blockIdx.x of the map names them (the map is already in XCD-tiled order).
// Without one, the grid is (blocks, scenes) and the block index is remapped
// by xcd_tile when `tiled`.
// 1.0f or 0.0f, opaque to the optimizer.  A loaded value times mask01(ok)
// keeps its load unconditional: with a plain (ok ? 1 : 0) the compiler
// specialises the code on ok and guards each load by a branch and a wait.
__device__ __forceinline__ float mask01(bool ok)
{
    float m = ok ? 1.f : 0.f;
    asm volatile("" : "+v"(m));
    return m;
}

// The solver chunk a block of a (blocks, scenes) grid works on: chunks in
// XCD-tiled order, so the rows a chunk shares with its neighbours (the
// stencil's reach) are read by blocks of one L2.
__device__ __forceinline__ int solver_block() { return xcd_tile(blockIdx.x, gridDim.x); }

struct SceneBlock {
    int scene, block;
};
__device__ __forceinline__ SceneBlock scene_block(const int2* map, bool tiled)
{
    if (map) {
        const int2 e = map[blockIdx.x];
        return SceneBlock{e.x, e.y};
    }
    return SceneBlock{(int)blockIdx.y, tiled ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x};
}

// wave64 + block reduction of doubles; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += lds[k];
    return t;
}

__device__ __forceinline__ float block_maxf(float v, float* lds)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    float t = 0.f;
    if (threadIdx.x == 0) {
        t = lds[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) t = fmaxf(t, lds[k]);
    }
    return t;
}

__device__ __forceinline__ int block_sumi(int v, int* lds)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    int t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += lds[k];
    return t;
}

__device__ __forceinline__ bool finite_d(double x) { return x == x && x < 1e300 && x > -1e300; }

// Solver blocks own contiguous chunks of cells.
struct Chunk {
    int begin, end;
};
__device__ __forceinline__ Chunk my_chunk(const SceneDev& s, int blk)
{
    const int C = s.I * s.J;
    const int beg = blk * s.pcg_chunk;
    return Chunk{beg < C ? beg : C, min(C, beg + s.pcg_chunk)};
}

// Sum of the nb per-block partials of `slot`, in a fixed order (thread t
// takes partials t, t+256, ...; then a fixed shuffle/LDS tree), broadcast to
// every thread.  Every block computing it gets the identical value.
template <int NV>
__device__ __forceinline__ void sum_partials(const SceneDev& s, int slot, int nb, double (&out)[NV])
{
    __shared__ double red[kBlock / 64];
    __shared__ double bc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        gptr<const double> part = gbl<const double>(s.partials + (size_t)(slot + q) * s.max_blocks);
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += kBlock) acc += part[b];
        const double t = block_sum(acc, red);
        if (threadIdx.x == 0) bc[q] = t;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) out[q] = bc[q];
}

// Block-reduce NV doubles and store them as the partials of solver block blk.
template <int NV>
__device__ __forceinline__ void put_partials(const SceneDev& s, int slot, double (&v)[NV], int blk)
{
    __shared__ double red[kBlock / 64];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const double t = block_sum(v[q], red);
        if (threadIdx.x == 0) s.partials[(size_t)(slot + q) * s.max_blocks + blk] = t;
    }
}

}  // namespace fg
