// fg_api.hip -- host orchestration behind the C-ABI in include/flipgpu.h.
//
// A scene owns one device arena (a single hipMalloc carved into aligned
// buffers) and a device copy of its SceneDev descriptor.  A substep is the
// reference's FlipSolver::step (flipsolver2d.cpp:412-462) as a fixed
// sequence of kernel launches on the context's stream; the only host round
// trips are the per-substep read of max |v| (to form dt exactly as
// FlipSolver::stepFrame does, flipsolver2d.cpp:476-490) and the solver's
// convergence polls.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fg_kernels.h"
#include "flipgpu.h"

using namespace fg;

namespace fg {
LaunchEvents& next_launch_events();
}

namespace {

thread_local std::string g_err;
thread_local fg::LaunchEvents g_next_events;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(FG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

#define FGCHK(expr)               \
    do {                          \
        int r_ = (expr);          \
        if (r_ != FG_OK) return r_; \
    } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Block maps of the solve's per-iteration launches, by kind: 0 = solver
// chunks, 1 = level-1 tiles of kVcTile (k_mg_down0, and k_mg_pp of level 1);
// with the fused V-cycle l = level-l tiles of kRRTile (2 <= l <= t: k_mg_rr
// of level l-1) and kMgMaxLevels + l = level-l tiles of kVcTile (2 <= l < t:
// k_mg_pp of level l), else 1 + l = level-l cells one per thread (1 <= l < t).
constexpr int kMapKinds = 2 * kMgMaxLevels;
// the kind of k_mg_pp's launch over level l (fused)
inline int pp_kind(int l) { return l == 1 ? 1 : kMgMaxLevels + l; }
constexpr int kMapMaxScenes = 4096;   // scenes one map build handles (LDS prefix)
constexpr int kFlagWords = 4 + kMapKinds;

}  // namespace

// One HIP stream with the buffers its launch groups use.  A context owns
// lane 0 (its own stream); a batch may be split over more lanes, each
// stepped by its own host thread, so the scenes of one lane overlap on the
// device with those of the others.
struct Lane {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int bcap = 0;
    SceneDev* bdesc = nullptr;      // a group's descriptor array
    SceneDev* bdesc_host = nullptr;
    SceneDev* bsub = nullptr;       // a sub-group's descriptors (density correction)
    SceneDev* bsub_host = nullptr;
    SceneDev* bpack = nullptr;      // iterating scenes of a solve, packed
    SceneScalars* bsc = nullptr;    // gathered scalar blocks
    SceneScalars* bsc_host = nullptr;
    int32_t* bflags = nullptr;      // [kFlagWords]: poll flags / pack info, then map totals
    int32_t* bflags_host = nullptr;
    int2* bmap = nullptr;           // block maps of the packed scenes
    int64_t bmap_cap = 0;
    void* bgeo = nullptr;           // polygons of a scene set-up
    size_t bgeo_cap = 0;
    // per-launch kernel timing
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    int64_t prof_units = 0;
    // work units of every family since the last fg_ctx_work_read reset
    int64_t work[FG_K_NUM] = {};
};

struct fg_ctx {
    int device = 0;
    uint64_t budget = 0;
    uint64_t used = 0;
    hipStream_t stream = nullptr;   // lane 0's stream
    bool timing = false;
    int prof_kernel = -1;           // kernel family timed per launch, -1: none
    hipEvent_t prof_ref = nullptr;  // recorded when timing starts, before any timed launch
    int nlanes = 1;                 // lanes a batch is split over
    int pcg_target_blocks = kPcgMaxBlocks;   // solver tiling: blocks per scene at most
    int pcg_min_chunk = kPcgMinChunk;        // and cells per block at least
    int precond = kPrecondMG;                // pressure PCG preconditioner
    bool vcycle_fused = true;                // V-cycle levels as k_mg_rr / k_mg_pp (else per-level kernels)
    std::vector<Lane*> lanes;
};

struct fg_scene {
    fg_ctx* ctx = nullptr;
    fg_scene_params p{};
    bool solver_only = false;
    int I = 0, J = 0;
    int64_t C = 0, FU = 0, FV = 0, cap = 0;
    int max_blocks = 0;     // FP64 partials stride: solver blocks at the finest tiling
    int pcg_chunk = 0;      // current tiling (set per launch group from the context)
    int pcg_blocks = 0;
    int max_pblocks = 0;    // particle blocks at capacity
    void* arena = nullptr;
    uint64_t bytes = 0;
    SceneDev h{};                 // host copy of the descriptor
    MgLevels mgh{};               // host copy of the multigrid hierarchy (device pointers)
    bool mg_ok = false;           // the hierarchy fits kMgMaxLevels
    SceneDev* d = nullptr;        // device descriptor
    SceneDev* hpin = nullptr;     // pinned staging for descriptor uploads
    SceneScalars* hsc = nullptr;  // pinned readback of the scalars
    int iters_hint[2] = {0, 0};   // first-sweep PCG iterations of the last solve (pressure, density)
    int ls_hint = 0;              // level-set bands the last substep ran
    int ls_hint_out = 0;          // NB-FLIP: bands the last outside extrapolation launched
    int64_t n_bound = 0;          // upper bound of live particles (host mirror)
    int64_t n_live = 0;           // live particles at the last readback
    bool vmax_ready = false;
    int last_stage = -1;          // the last fg_scene_stage of an unfinished substep, else -1
    int next_stage = 0;
    int frame = 0;
    int status = FG_OK;           // FG_E_NUMERIC: quarantined until the next set_state
    uint32_t epoch = 0;           // substeps taken (the reseed RNG's counter)
    uint32_t next_pid = 0;        // id of the next reseeded particle
};

// Device scratch of fg_poisson_solve_noalloc: the solver-only scene's arena
// and its descriptor.  Each device that loads the library holds its own copy
// until the process ends; its address on a device comes from
// hipGetSymbolAddress (resolved once per device when a context is created
// there) and only that address is handed to kernels and copies.
constexpr uint64_t kNoAllocScratchBytes = 48ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocDevices = 64;
// Per device: the scratch's device address, the host-side scene that lives
// in it and the descriptor / scalar staging of its copies (ordinary host
// memory: nothing is allocated per call).  The mutex serialises the calls
// on one device.
struct NoAllocSlot {
    std::mutex mu;
    void* dev = nullptr;
    fg_scene scene;
    SceneDev desc_stage;
    SceneScalars scalars;
};
NoAllocSlot g_noalloc[kNoAllocDevices];

// Resolve the scratch's address on `device` (current) once.
int noalloc_resolve(int device)
{
    if (device < 0 || device >= kNoAllocDevices) return FG_OK;   // such a device has no slot
    NoAllocSlot& slot = g_noalloc[device];
    std::lock_guard<std::mutex> lock(slot.mu);
    if (slot.dev) return FG_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(FG_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    slot.dev = ptr;
    return FG_OK;
}

}  // namespace

namespace {

// --------------------------------------------------------------------------
// arena layout
// --------------------------------------------------------------------------
struct Carver {
    uint64_t off = 0;
    char* base = nullptr;
    template <typename T>
    T* take(int64_t count)
    {
        off = (off + 255) & ~uint64_t(255);
        T* ptr = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += sizeof(T) * (uint64_t)std::max<int64_t>(count, 1);
        return ptr;
    }
};

// The Galerkin operators of an all-liquid grid, level by level: the same
// products, in the same order, as k_mg_galerkin forms for a cell whose whole
// neighbourhood is active (level 0: the 5-point operator, n = 4).
const float (&mg_canonical())[kMgMaxLevels][kMgStencil]
{
    static float canon[kMgMaxLevels][kMgStencil];
    static bool done = false;
    if (done) return canon;
    auto half_floor = [](int t) { return t >= 0 ? t / 2 : -((1 - t) / 2); };
    auto w1 = [](int a) { return (a == 1 || a == 2) ? 0.75 : 0.25; };
    for (int k = 0; k < kMgStencil; k++) canon[0][k] = 0.f;
    canon[0][12] = 4.f;
    canon[0][7] = canon[0][17] = canon[0][11] = canon[0][13] = -1.f;
    for (int l = 1; l < kMgMaxLevels; l++) {
        double acc[kMgStencil] = {};
        for (int ab = 0; ab < 16; ab++) {
            const int a = ab / 4, b = ab % 4;
            const double wc = w1(a) * w1(b);
            auto spread = [&](int ei, int ej, double v) {
                const int ti = -1 + a + ei, tj = -1 + b + ej;
                const int pi0 = half_floor(ti), pj0 = half_floor(tj);
                const int pi1 = (ti & 1) ? pi0 + 1 : pi0 - 1, pj1 = (tj & 1) ? pj0 + 1 : pj0 - 1;
                const int pis[4] = {pi0, pi0, pi1, pi1}, pjs[4] = {pj0, pj1, pj0, pj1};
                const double ws[4] = {0.5625, 0.1875, 0.1875, 0.0625};
                for (int q = 0; q < 4; q++)
                    if (pis[q] >= -2 && pis[q] <= 2 && pjs[q] >= -2 && pjs[q] <= 2)
                        acc[(pis[q] + 2) * 5 + (pjs[q] + 2)] += (wc * ws[q]) * v;
            };
            if (l == 1) {
                spread(0, 0, 4.0);
                spread(-1, 0, -1.0);
                spread(1, 0, -1.0);
                spread(0, -1, -1.0);
                spread(0, 1, -1.0);
            } else {
                for (int e = 0; e < kMgStencil; e++) spread(e / 5 - 2, e % 5 - 2, (double)canon[l - 1][e]);
            }
        }
        for (int k = 0; k < kMgStencil; k++) canon[l][k] = (float)(0.25 * acc[k]);
    }
    done = true;
    return canon;
}

void carve(fg_scene* s, Carver& cv)
{
    SceneDev& h = s->h;
    const int64_t C = s->C, FU = s->FU, FV = s->FV, P = s->cap;
    const int64_t nb = s->max_blocks;
    // solver (always)
    h.mat = cv.take<int8_t>(C);
    h.stencil = cv.take<uint8_t>(C);
    h.b = cv.take<double>(C);
    h.x64 = cv.take<double>(C);
    h.cr = cv.take<float>(C);
    h.cr2 = cv.take<float>(C);
    h.cw = cv.take<float>(C);
    h.cs1 = cv.take<float>(C);
    h.cu = cv.take<float>(C);
    h.cp = cv.take<float>(C);
    h.partials = cv.take<double>(kSlots * nb);
    h.sc = cv.take<SceneScalars>(1);
    // multigrid: halve until the bottom fits kMgBottomMax per side (at
    // least one coarse level)
    {
        MgLevels& M = s->mgh;
        M = MgLevels{};
        M.lv[0].I = s->I;
        M.lv[0].J = s->J;
        int n = 1;
        while (n < kMgMaxLevels && (n == 1 || std::max(M.lv[n - 1].I, M.lv[n - 1].J) > kMgBottomMax)) {
            M.lv[n].I = (M.lv[n - 1].I + 1) / 2;
            M.lv[n].J = (M.lv[n - 1].J + 1) / 2;
            n++;
        }
        s->mg_ok = std::max(M.lv[n - 1].I, M.lv[n - 1].J) <= kMgBottomMax;
        M.n = n;
        // the levels k_mg_coarse runs in one block: from the first level
        // that fits, with its coarser levels, in LDS (the bottom always
        // does: at most kMgBottomMax^2 cells)
        auto cells = [&](int l) { return (int64_t)M.lv[l].I * M.lv[l].J; };
        M.t = std::max(1, n - 1);
        for (int t = 1; t < n - 1; t++) {
            int64_t sum = 0;
            for (int l = t; l < n; l++) sum += cells(l);
            if (cells(t) <= kMgCoarseTop && sum <= kMgCoarseCells) {
                M.t = t;
                break;
            }
        }
        int o = 0;
        for (int l = M.t; l < n; l++) {
            M.roff[l] = o;
            o += (int)cells(l);
        }
        M.tcells = o;
        std::memcpy(M.canon, mg_canonical(), sizeof(M.canon));
        h.mg = cv.take<MgLevels>(1);
        int64_t all = 0;
        for (int l = 1; l < n; l++) all += cells(l);
        float* const dinv = cv.take<float>(all);   // every level's dinv, contiguous in level order
        int64_t doff = 0;
        for (int l = 1; l < n; l++) {
            MgLevel& L = M.lv[l];
            const int64_t CL = (int64_t)L.I * L.J;
            L.act = cv.take<uint8_t>(CL);
            L.coef = cv.take<float>(kMgRowStride * CL);
            L.dinv = dinv ? dinv + doff : nullptr;
            doff += CL;
            L.r = cv.take<float>(CL);
            L.x = cv.take<float>(CL);
            L.y = cv.take<float>(CL);
        }
    }
    if (s->solver_only) return;
    h.pos = cv.take<float2>(P);
    h.vel = cv.take<float2>(P);
    h.posA = cv.take<float2>(P);
    h.velA = cv.take<float2>(P);
    h.pid = cv.take<uint32_t>(P);
    h.key = cv.take<uint32_t>(P);
    h.tmpPid = cv.take<uint32_t>(P);
    h.tmpSrc = cv.take<uint32_t>(P);
    h.ipart = cv.take<int32_t>(2 * (int64_t)s->max_pblocks);
    h.fpart = cv.take<float>(s->max_pblocks);
    h.cellCount = cv.take<int32_t>(C);
    h.cellStart = cv.take<int32_t>(C + 1);
    h.scanSums = cv.take<int32_t>(cdiv(C, kScanTileCells));
    h.ssdf = cv.take<float>(C);
    h.fsdf = cv.take<float>(C);
    h.sid = cv.take<int32_t>(C);
    h.pcount = cv.take<int32_t>(C);
    h.marker = cv.take<int32_t>(C);
    h.lsFlags = cv.take<int32_t>(kMaxBands);
    h.lsTile = cv.take<uint8_t>(2 * cdiv(s->I, kBfsTile) * cdiv(s->J, kBfsTile));
    h.u = cv.take<float>(FU);
    h.su = cv.take<float>(FU);
    h.uf = cv.take<float>(FU);
    h.v = cv.take<float>(FV);
    h.sv = cv.take<float>(FV);
    h.vf = cv.take<float>(FV);
    h.uval = cv.take<uint8_t>(FU);
    h.uvalS = cv.take<uint8_t>(FU);
    h.uvalF = cv.take<uint8_t>(FU);
    h.vval = cv.take<uint8_t>(FV);
    h.vvalS = cv.take<uint8_t>(FV);
    h.vvalF = cv.take<uint8_t>(FV);
    h.friction = cv.take<float>(256);
    h.eid = cv.take<int32_t>(C);
    h.emitters = cv.take<float4>(kMaxEmitters);
    if (s->p.viscosity_enabled) {
        const int64_t cb = cdiv(C, kBlock);
        h.pvisc = cv.take<float>(P);
        h.pviscA = cv.take<float>(P);
        h.cvisc = cv.take<float>(C);
        h.vd = cv.take<double>(C);
        h.ve = cv.take<double>(C);
        h.vn = cv.take<double>(C);
        h.vx = cv.take<double>(2 * C);
        h.vr = cv.take<double>(2 * C);
        h.vp = cv.take<double>(2 * C);
        h.vq = cv.take<double>(2 * C);
        h.vpart = cv.take<double>(kViscSlots * cb);
        h.vsc = cv.take<double>(kViscScalars);
    }
    if (s->p.sim_method == FG_SIM_NBFLIP) {
        h.adv_u = cv.take<float>(FU);
        h.adv_v = cv.take<float>(FV);
        h.adv_sdf = cv.take<float>(C);
        h.src_sdf = cv.take<float>(C);
        if (s->p.viscosity_enabled) {
            h.adv_visc = cv.take<float>(C);
            h.src_id = cv.take<int32_t>(C);
        }
    }
}

uint64_t arena_bytes(fg_scene* s)
{
    Carver cv;
    SceneDev keep = s->h;
    carve(s, cv);
    s->h = keep;
    return cv.off + 256;
}

int validate_params(const fg_scene_params* p)
{
    if (!p) return fail(FG_E_INVALID, "params is NULL");
    if (p->size_i < 2 || p->size_j < 2) return fail(FG_E_INVALID, "grid must be at least 2x2");
    if ((int64_t)p->size_i * p->size_j > (int64_t)1 << 30) return fail(FG_E_INVALID, "grid too large");
    if (!(p->dx > 0) || !(p->fluid_density > 0)) return fail(FG_E_INVALID, "dx and fluid_density must be > 0");
    if (p->fps <= 0) return fail(FG_E_INVALID, "fps must be > 0");
    if (p->max_particles < 0 || p->max_particles > ((int64_t)1 << 31) - 1)
        return fail(FG_E_INVALID, "max_particles out of range");
    if (p->density_mode < 0 || p->density_mode > 2) return fail(FG_E_INVALID, "bad density_mode");
    if (p->sim_method != FG_SIM_FLIP && p->sim_method != FG_SIM_NBFLIP) return fail(FG_E_INVALID, "bad sim_method");
    if (p->sim_method == FG_SIM_NBFLIP && p->density_mode != FG_DENSITY_OFF)
        return fail(FG_E_INVALID, "NBFlipSolver::step has no density correction: density_mode must be OFF");
    if (p->viscosity_enabled != 0 && p->viscosity_enabled != 1) return fail(FG_E_INVALID, "bad viscosity_enabled");
    if (p->viscosity_enabled && p->density_mode != FG_DENSITY_OFF)
        return fail(FG_E_INVALID, "FlipSolver::step skips density correction with viscosity: density_mode must be OFF");
    return FG_OK;
}

void init_geometry_fields(fg_scene* s)
{
    s->I = s->p.size_i;
    s->J = s->p.size_j;
    s->C = (int64_t)s->I * s->J;
    s->FU = (int64_t)(s->I + 1) * s->J;
    s->FV = (int64_t)s->I * (s->J + 1);
    s->cap = s->solver_only ? 0 : s->p.max_particles;
    // solver blocks own contiguous chunks of whole 256-cell groups
    const int64_t groups = cdiv(s->C, kBlock);
    // partials for the finest tiling the context may choose
    const int64_t want = std::min<int64_t>(groups, kPcgTilingMaxBlocks);
    const int64_t finest = std::max<int64_t>(kPcgTilingMinChunk, cdiv(groups, want) * kBlock);
    s->max_blocks = (int)cdiv(s->C, finest);
    s->pcg_chunk = (int)finest;
    s->pcg_blocks = s->max_blocks;
    s->max_pblocks = (int)std::max<int64_t>(1, cdiv(s->cap, kBlock));
}

int upload_desc(fg_scene* s)
{
    *s->hpin = s->h;
    HIPCHK(hipMemcpyAsync(s->d, s->hpin, sizeof(SceneDev), hipMemcpyHostToDevice, s->ctx->stream));
    return FG_OK;
}

int read_scalars(fg_scene* s)
{
    HIPCHK(hipMemcpyAsync(s->hsc, s->h.sc, sizeof(SceneScalars), hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return FG_OK;
}

// Carve a scene's arena (s->arena, s->bytes) and fill its descriptor; clear
// the arena and upload the hierarchy table and the descriptor (s->d) on the
// context's stream.  Allocates nothing.
int setup_scene_memory(fg_scene* s)
{
    fg_ctx* ctx = s->ctx;
    const fg_scene_params* p = &s->p;
    const bool solver_only = s->solver_only;
    Carver cv;
    cv.base = reinterpret_cast<char*>(s->arena);
    carve(s, cv);
    SceneDev& h = s->h;
    h.I = s->I;
    h.J = s->J;
    h.ppc = p->particles_per_cell;
    h.n_friction = 0;
    h.friction_any = 0;
    h.pic = p->pic_ratio;
    h.pscale = p->particle_scale;
    // every particle of a cell is within sqrt(0.5) = 0.70710677f of its
    // centre, so above that radius a cell holding a particle is inside
    h.sdf_short = p->particle_scale > 0.70710677f ? 1 : 0;
    h.g_i = p->gravity_i;
    h.g_j = p->gravity_j;
    h.dx = p->dx;
    h.rho = p->fluid_density;
    h.tol = p->solver_tolerance > 0 ? p->solver_tolerance : 1e-6;
    h.max_inner = p->pcg_iter_limit > 0 ? p->pcg_iter_limit : 2000;
    h.active = 1;
    h.density_mode = p->density_mode;
    h.nb = p->sim_method == FG_SIM_NBFLIP ? 1 : 0;
    h.visc_on = p->viscosity_enabled ? 1 : 0;
    // flipsolver2d.cpp:207-208
    h.cell_vol = (float)(p->dx * p->dx * p->dx);
    h.pmass = (float)((p->fluid_density * h.cell_vol) / (float)p->particles_per_cell);
    h.max_blocks = s->max_blocks;
    h.pcg_blocks = s->pcg_blocks;
    h.pcg_chunk = s->pcg_chunk;
    h.max_pblocks = s->max_pblocks;
    h.pblocks = 1;
    h.cap = s->cap;
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemsetAsync(s->arena, 0, s->bytes, st));
    if (!solver_only) {
        // empty, all-EMPTY scene until state arrives
        HIPCHK(hipMemsetAsync(h.mat, (int)(int8_t)FG_EMPTY, s->C, st));
        HIPCHK(hipMemsetAsync(h.sid, 0xff, sizeof(int32_t) * s->C, st));
        HIPCHK(hipMemsetAsync(h.eid, 0xff, sizeof(int32_t) * s->C, st));
        if (h.nb) HIPCHK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(h.src_sdf), 0x7f7fffff, s->C, st));
        if (h.nb && h.visc_on) HIPCHK(hipMemsetAsync(h.src_id, 0xff, sizeof(int32_t) * s->C, st));
    }
    HIPCHK(hipMemcpyAsync(h.mg, &s->mgh, sizeof(MgLevels), hipMemcpyHostToDevice, st));
    FGCHK(upload_desc(s));
    HIPCHK(hipStreamSynchronize(st));
    return FG_OK;
}

int create_scene(fg_ctx* ctx, const fg_scene_params* p, bool solver_only, fg_scene** out)
{
    FGCHK(validate_params(p));
    auto* s = new fg_scene();
    s->ctx = ctx;
    s->p = *p;
    s->solver_only = solver_only;
    init_geometry_fields(s);
    s->bytes = arena_bytes(s);
    if (ctx->budget && ctx->used + s->bytes > ctx->budget) {
        const uint64_t need = s->bytes;
        delete s;
        return fail(FG_E_BUDGET, "scene needs " + std::to_string(need) + " bytes beyond the HBM budget");
    }
    hipError_t e = hipMalloc(&s->arena, s->bytes);
    if (e != hipSuccess) {
        delete s;
        return fail(FG_E_DEVICE, std::string("hipMalloc arena: ") + hipGetErrorString(e));
    }
    e = hipMalloc(&s->d, sizeof(SceneDev));
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->hpin, sizeof(SceneDev), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->hsc, sizeof(SceneScalars), hipHostMallocDefault);
    if (e != hipSuccess) {
        hipFree(s->arena);
        if (s->d) hipFree(s->d);
        if (s->hpin) hipHostFree(s->hpin);
        delete s;
        return fail(FG_E_DEVICE, std::string("descriptor alloc: ") + hipGetErrorString(e));
    }
    ctx->used += s->bytes;
    const int rc = setup_scene_memory(s);
    if (rc != FG_OK) {
        fg_scene_destroy(s);
        return rc;
    }
    *out = s;
    return FG_OK;
}

void stats_reset(fg_stats* st)
{
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->converged = 1;
}

// --------------------------------------------------------------------------
// launch geometry
// --------------------------------------------------------------------------
struct Geo {
    dim3 cells, parts, scan, scan1, faces, extrap, ls, one, pcg, bins, p2g, nodes;
};

// x extents of one scene's launches (y is set by the launch group)
Geo geometry(const fg_scene* s)
{
    Geo g;
    g.cells = dim3((unsigned)cdiv(s->C, kBlock));
    g.parts = dim3((unsigned)std::max<int64_t>(1, cdiv(std::max<int64_t>(s->n_bound, 1), kBlock)));
    g.scan = dim3((unsigned)cdiv(s->C, kScanTileCells));
    g.scan1 = dim3(1);
    g.faces = dim3((unsigned)cdiv(std::max(s->FU, s->FV), kBlock));
    const int64_t tu = cdiv(s->I + 1, kBfsTile) * cdiv(s->J, kBfsTile);
    const int64_t tv = cdiv(s->I, kBfsTile) * cdiv(s->J + 1, kBfsTile);
    g.extrap = dim3((unsigned)std::max(tu, tv), 1, 2);
    g.ls = dim3((unsigned)(cdiv(s->I, kBfsTile) * cdiv(s->J, kBfsTile)));
    g.one = dim3(1);
    g.pcg = dim3((unsigned)s->pcg_blocks);
    g.bins = dim3((unsigned)cdiv(cdiv(s->I, 3) * cdiv(s->J, 3), kBlock / kSdfLanes));
    g.p2g = dim3((unsigned)(cdiv(s->I, 16) * cdiv(s->J, 16)));
    g.nodes = dim3((unsigned)cdiv((int64_t)(s->I + 1) * (s->J + 1), kBlock));
    return g;
}

void geo_max(Geo& a, const Geo& b)
{
    dim3* pa[] = {&a.cells, &a.parts, &a.scan, &a.scan1, &a.faces, &a.extrap,
                  &a.ls, &a.one, &a.pcg, &a.bins, &a.p2g, &a.nodes};
    const dim3* pb[] = {&b.cells, &b.parts, &b.scan, &b.scan1, &b.faces, &b.extrap,
                        &b.ls, &b.one, &b.pcg, &b.bins, &b.p2g, &b.nodes};
    for (int k = 0; k < 12; k++) pa[k]->x = std::max(pa[k]->x, pb[k]->x);
}

void geo_set_y(Geo& a, int n)
{
    dim3* pa[] = {&a.cells, &a.parts, &a.scan, &a.scan1, &a.faces, &a.extrap,
                  &a.ls, &a.one, &a.pcg, &a.bins, &a.p2g, &a.nodes};
    for (dim3* d : pa) d->y = (unsigned)n;
}

void set_substep_scalars(fg_scene* s, float dt)
{
    SceneDev& h = s->h;
    h.dt = dt;
    h.force = (float)(dt / s->p.dx);                                   // flipsolver2d.cpp:1224
    h.sA = dt / (s->p.fluid_density * s->p.dx * s->p.dx);             // flipsolver2d.cpp:799
    h.sApply = dt / (s->p.fluid_density * s->p.dx);                   // flipsolver2d.cpp:1135
    h.sRhs = 1.f / s->p.dx;                                            // flipsolver2d.cpp:964
    h.dens_scale = (float)((dt * dt) / (s->p.fluid_density * s->p.dx * s->p.dx));  // flipsolver2d.cpp:263
    h.dens_rhs_scale = 1.0 / dt;                                       // flipsolver2d.cpp:995
    h.epoch = s->epoch;
    h.next_pid = s->next_pid;
}

// ---------------------------------------------------------------------------
// tiny per-scene scalar kernels
// ---------------------------------------------------------------------------
__global__ void k_substep_begin(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (threadIdx.x != 0 || !s.active) return;
    s.sc->killed = 0;
    s.sc->erased = 0;
    s.sc->density_applied = 0;
    s.sc->reseeded = 0;
    s.sc->reseed_dropped = 0;
    s.sc->sdf_known = 0;
}
__global__ void k_set_count(const SceneDev* scenes, int n)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (threadIdx.x != 0) return;
    s.sc->n_particles = n;
}
// Copy every scene's scalar block into one array (one host read per group).
__global__ void k_gather_scalars(const SceneDev* scenes, int n, SceneScalars* out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = *scenes[k].sc;
}
// Group-wide flags for the host's polls:
//   out[0]: some scene's solve has not met its FP64 tolerance
//   out[1]: some scene's inner PCG is still iterating
//   out[2]: some scene's level-set band `band` reached its last level
__global__ void k_group_flags(const SceneDev* scenes, int n, int band, int32_t* out)
{
    __shared__ int any[3];
    if (threadIdx.x < 3) any[threadIdx.x] = 0;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const SceneDev& s = scenes[k];
        if (!s.active) continue;
        if (band < 0) {
            const SceneScalars* sc = s.sc;
            if (!sc->outer_done) {
                any[0] = 1;
                if (!sc->inner_done) any[1] = 1;
            }
        } else if (s.lsFlags[band]) {
            any[2] = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) out[threadIdx.x] = any[threadIdx.x];
}

// The descriptors of the scenes whose inner PCG still iterates, packed in
// slot order into `out` (one wave); info = {count, max solver blocks,
// scenes whose FP64 test has not passed}.
__global__ void k_pack_iterating(const SceneDev* in, int n, SceneDev* out, int32_t* info)
{
    const int lane = threadIdx.x;
    int base = 0, maxblk = 0, open = 0;
    for (int b0 = 0; b0 < n; b0 += 64) {
        const int k = b0 + lane;
        bool it = false, op = false;
        if (k < n) {
            const SceneDev& s = in[k];
            op = s.active && !s.sc->outer_done;
            it = op && !s.sc->inner_done;
        }
        open += __popcll(__ballot(op));
        const unsigned long long mask = __ballot(it);
        const int pos = __popcll(mask & ((1ull << lane) - 1ull));
        if (it) {
            out[base + pos] = in[k];
            maxblk = max(maxblk, in[k].pcg_blocks);
        }
        base += __popcll(mask);
    }
    for (int o = 32; o > 0; o >>= 1) maxblk = max(maxblk, __shfl_xor(maxblk, o, 64));
    if (lane == 0) {
        info[0] = base;
        info[1] = maxblk;
        info[2] = open;
    }
}

// Blocks of kind `kind` (see kMapKinds) a packed scene launches.
__device__ int map_blocks(const SceneDev& s, int kind, bool fused)
{
    if (kind == 0) return s.pcg_blocks;
    if (s.precond != kPrecondMG) return 0;
    const MgLevels& M = *s.mg;
    if (kind == 1) return mg_d0_tiles(M.lv[1].I, M.lv[1].J);
    if (fused) {
        if (kind <= kMgMaxLevels) return kind <= M.t ? mg_tiles(M.lv[kind].I, M.lv[kind].J, kRRTile) : 0;
        const int l = kind - kMgMaxLevels;
        return l < M.t ? mg_tiles(M.lv[l].I, M.lv[l].J) : 0;
    }
    const int l = kind - 1;
    if (l >= M.t) return 0;   // level l runs inside k_mg_coarse
    return (M.lv[l].I * M.lv[l].J + kBlock - 1) / kBlock;
}

// Blocks of map kind `kind` for one scene (the host's copy of map_blocks).
int64_t map_blocks_host(const fg_scene* s, int kind, bool fused)
{
    if (kind == 0) return s->pcg_blocks;
    if (s->h.precond != kPrecondMG) return 0;
    const MgLevels& M = s->mgh;
    if (kind == 1) return mg_d0_tiles(M.lv[1].I, M.lv[1].J);
    if (fused) {
        if (kind <= kMgMaxLevels) return kind <= M.t ? mg_tiles(M.lv[kind].I, M.lv[kind].J, kRRTile) : 0;
        const int l = kind - kMgMaxLevels;
        return l < M.t ? mg_tiles(M.lv[l].I, M.lv[l].J) : 0;
    }
    const int l = kind - 1;
    if (l >= M.t) return 0;
    return cdiv((int64_t)M.lv[l].I * M.lv[l].J, kBlock);
}

struct MapOffsets {
    int64_t off[kMapKinds];
};

// One block per map kind: the (slot, block) pairs of every packed scene's
// blocks, stored so that flat block e takes entry xcd_tile(e, total) of the
// natural order (consecutive tiles of a scene on one XCD).  info[0] = the
// packed count; info[4 + kind] gets the kind's total.
__global__ void __launch_bounds__(1024) k_build_maps(const SceneDev* packed, int32_t* info, int2* maps, MapOffsets offs,
                                                     int n_all, bool fused)
{
    __shared__ int pre[kMapMaxScenes + 1];
    __shared__ int wsum[1024 / 64];
    const int kind = blockIdx.x;
    const int n = n_all >= 0 ? n_all : info[0];   // n_all: every scene of the array
    const int t = threadIdx.x;
    // each thread sums a contiguous run of scenes, then a block scan of the runs
    const int per = (n + 1023) / 1024;
    const int b0 = min(n, t * per), b1 = min(n, b0 + per);
    int run = 0;
    for (int i = b0; i < b1; i++) {
        const int c = map_blocks(packed[i], kind, fused);
        pre[i + 1] = c;
        run += c;
    }
    int incl = run;
    const int lane = t & 63, w = t >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int woff = 0;
    for (int k = 0; k < w; k++) woff += wsum[k];
    int acc = woff + incl - run;   // exclusive prefix of this thread's run
    if (t == 0) pre[0] = 0;
    for (int i = b0; i < b1; i++) {
        acc += pre[i + 1];
        pre[i + 1] = acc;
    }
    __syncthreads();
    const int total = pre[n];
    int2* out = maps + offs.off[kind];
    for (int e = t; e < total; e += 1024) {
        const int nat = xcd_tile(e, total);
        int lo = 0, hi = n - 1;   // the last scene whose blocks start at or before nat
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= nat) lo = mid;
            else hi = mid - 1;
        }
        out[e] = make_int2(lo, nat - pre[lo]);
    }
    if (t == 0) info[4 + kind] = total;
}

// ---------------------------------------------------------------------------
// launch groups: the scenes of one launch, grid.y = scene slot
// ---------------------------------------------------------------------------
struct Group {
    fg_ctx* ctx = nullptr;
    Lane* L = nullptr;
    std::vector<fg_scene*> sc;
    std::vector<fg_stats*> st;     // per scene, entries may be NULL
    const SceneDev* d = nullptr;   // device descriptor array [n]
    int n = 0;
    Geo g{};
    int64_t cells = 0, faces = 0, parts = 0;
    int max_inner = 0;
    int max_bands = 0;
    int max_j = 0;                      // the widest scene's J
    bool sources = false;               // some scene has SOURCE cells (afterTransfer, reseeding)
    bool nb = false;                    // NB-FLIP scenes (a group holds one solver class)
    bool visc = false;                  // some scene has viscosity enabled
    int mg_levels = 0;                  // most multigrid levels of a scene using the V-cycle
    int map_kinds = 0;                  // block-map kinds the solve uses (0 when maps are off)
    int64_t map_off[kMapKinds]{};       // offset of each kind's map in the lane's map buffer
    int mg_top = 0;                     // largest first level of k_mg_coarse (levels below: one launch each)
    unsigned mg_grid[kMgMaxLevels]{};   // blocks per scene over level l, one cell per thread
    unsigned sym_grid[kMgMaxLevels]{};  // 16 x 16-row tiles per scene over level l (k_mg_sym)
    unsigned vc_grid[kMgMaxLevels]{};   // the same over the scenes whose V-cycle launches level l (l < t)
    unsigned rr_grid[kMgMaxLevels]{};   // kRRTile tiles of level l (2 <= l <= t): k_mg_rr's blocks a scene
    unsigned pp_grid[kMgMaxLevels]{};   // kVcTile tiles of level l (l < t): k_mg_pp's blocks a scene
    bool fused = true;                  // the context's vcycle_fused when the group was made
    unsigned mg_grid_d0 = 0;            // blocks per scene of k_mg_down0 (level-1 tiles)
};

void lane_free_buffers(Lane* L)
{
    if (L->bdesc) hipFree(L->bdesc);
    if (L->bdesc_host) hipHostFree(L->bdesc_host);
    if (L->bsub) hipFree(L->bsub);
    if (L->bsub_host) hipHostFree(L->bsub_host);
    if (L->bpack) hipFree(L->bpack);
    if (L->bsc) hipFree(L->bsc);
    if (L->bsc_host) hipHostFree(L->bsc_host);
    L->bdesc = L->bdesc_host = L->bsub = L->bsub_host = L->bpack = nullptr;
    L->bsc = L->bsc_host = nullptr;
    L->bcap = 0;
}

void lane_destroy(Lane* L)
{
    if (!L) return;
    if (L->stream) hipStreamSynchronize(L->stream);
    lane_free_buffers(L);
    if (L->bflags) hipFree(L->bflags);
    if (L->bflags_host) hipHostFree(L->bflags_host);
    if (L->bmap) hipFree(L->bmap);
    if (L->bgeo) hipFree(L->bgeo);
    for (hipEvent_t e : L->ev_pool) hipEventDestroy(e);
    if (L->own_stream && L->stream) hipStreamDestroy(L->stream);
    delete L;
}

int lane_create(hipStream_t stream, Lane** out)
{
    auto* L = new Lane();
    hipError_t e = hipSuccess;
    if (stream) {
        L->stream = stream;
    } else {
        e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
        L->own_stream = e == hipSuccess;
    }
    if (e == hipSuccess) e = hipMalloc(&L->bflags, sizeof(int32_t) * kFlagWords);
    if (e == hipSuccess) e = hipHostMalloc((void**)&L->bflags_host, sizeof(int32_t) * kFlagWords, hipHostMallocDefault);
    if (e != hipSuccess) {
        lane_destroy(L);
        return fail(FG_E_DEVICE, std::string("lane: ") + hipGetErrorString(e));
    }
    *out = L;
    return FG_OK;
}

// Solver blocks own contiguous chunks of whole 256-cell groups: at most
// pcg_target_blocks per scene, pcg_min_chunk cells each at least.
void apply_tiling(const fg_ctx* ctx, fg_scene* s)
{
    const int64_t groups = cdiv(s->C, kBlock);
    const int64_t want = std::max<int64_t>(1, std::min<int64_t>(groups, ctx->pcg_target_blocks));
    int64_t chunk = std::max<int64_t>(ctx->pcg_min_chunk, cdiv(groups, want) * kBlock);
    int64_t blocks = cdiv(s->C, chunk);
    if (blocks > s->max_blocks) {   // finer than the partials allow
        chunk = cdiv(s->C, s->max_blocks);
        chunk = cdiv(chunk, kBlock) * kBlock;
        blocks = cdiv(s->C, chunk);
    }
    s->pcg_chunk = (int)chunk;
    s->pcg_blocks = (int)blocks;
    s->h.pcg_chunk = s->pcg_chunk;
    s->h.pcg_blocks = s->pcg_blocks;
}

int ensure_group_buffers(Lane* L, int n)
{
    if (n <= L->bcap) return FG_OK;
    const int cap = std::max(n, 2 * L->bcap);
    HIPCHK(hipStreamSynchronize(L->stream));
    lane_free_buffers(L);
    HIPCHK(hipMalloc(&L->bdesc, sizeof(SceneDev) * cap));
    HIPCHK(hipHostMalloc((void**)&L->bdesc_host, sizeof(SceneDev) * cap, hipHostMallocDefault));
    HIPCHK(hipMalloc(&L->bsub, sizeof(SceneDev) * cap));
    HIPCHK(hipHostMalloc((void**)&L->bsub_host, sizeof(SceneDev) * cap, hipHostMallocDefault));
    HIPCHK(hipMalloc(&L->bpack, sizeof(SceneDev) * cap));
    HIPCHK(hipMalloc(&L->bsc, sizeof(SceneScalars) * cap));
    HIPCHK(hipHostMalloc((void**)&L->bsc_host, sizeof(SceneScalars) * cap, hipHostMallocDefault));
    L->bcap = cap;
    return FG_OK;
}

// Build a group and upload its descriptors.  A single scene uses its own
// device descriptor; several share the context's descriptor array.
int make_group(fg_ctx* ctx, const std::vector<fg_scene*>& scenes, const std::vector<fg_stats*>& stats, Group& G,
               bool sub = false, Lane* lane = nullptr)
{
    G = Group{};
    G.ctx = ctx;
    G.L = lane ? lane : ctx->lanes[0];
    Lane* L = G.L;
    G.sc = scenes;
    G.st = stats;
    G.st.resize(scenes.size(), nullptr);
    G.n = (int)scenes.size();
    if (G.n == 0) return FG_OK;
    G.fused = ctx->vcycle_fused;
    // the pinned staging buffers may still feed an earlier copy
    HIPCHK(hipStreamSynchronize(L->stream));
    bool first = true;
    for (fg_scene* s : scenes) {
        apply_tiling(ctx, s);
        s->h.active = 1;
        s->h.precond = (ctx->precond == kPrecondMG && s->mg_ok) ? kPrecondMG : kPrecondIPP;
        if (s->h.precond == kPrecondMG) {
            G.mg_levels = std::max(G.mg_levels, s->mgh.n);
            G.mg_top = std::max(G.mg_top, s->mgh.t);
            for (int l = 1; l < s->mgh.n; l++) {
                const unsigned b = (unsigned)cdiv((int64_t)s->mgh.lv[l].I * s->mgh.lv[l].J, kBlock);
                G.mg_grid[l] = std::max(G.mg_grid[l], b);
                G.sym_grid[l] = std::max(G.sym_grid[l], (unsigned)(cdiv(s->mgh.lv[l].I, 16) * cdiv(s->mgh.lv[l].J, 16)));
                if (l < s->mgh.t) G.vc_grid[l] = std::max(G.vc_grid[l], b);
                if (l <= s->mgh.t)
                    G.rr_grid[l] = std::max(G.rr_grid[l], (unsigned)mg_tiles(s->mgh.lv[l].I, s->mgh.lv[l].J, kRRTile));
                if (l < s->mgh.t)
                    G.pp_grid[l] = std::max(G.pp_grid[l], (unsigned)mg_tiles(s->mgh.lv[l].I, s->mgh.lv[l].J));
            }
            G.mg_grid_d0 = std::max(G.mg_grid_d0, (unsigned)mg_d0_tiles(s->mgh.lv[1].I, s->mgh.lv[1].J));
        }
        s->h.pblocks = (int)std::max<int64_t>(1, cdiv(std::max<int64_t>(s->n_bound, 1), kBlock));
        const Geo gs = geometry(s);
        if (first) G.g = gs;
        else geo_max(G.g, gs);
        first = false;
        G.cells += s->C;
        G.max_j = std::max(G.max_j, s->J);
        G.faces += s->FU + s->FV;
        G.parts += s->n_bound;
        G.max_inner = std::max(G.max_inner, s->h.max_inner);
        G.sources = G.sources || s->h.has_sources;
        G.nb = G.nb || s->h.nb;
        G.visc = G.visc || s->h.visc_on;
        G.max_bands = std::max(G.max_bands, (int)std::min<int64_t>(kMaxBands, cdiv(std::max(s->I, s->J), 16) + 1));
    }
    geo_set_y(G.g, G.n);
    // block maps for the solve's per-iteration launches (several scenes):
    // capacity = every scene's blocks, the bound for any packed subset
    if (G.n > 1 && G.n <= kMapMaxScenes) {
        // pcg, down0, and levels 1 .. t-1 (fused: through k_mg_pp of level t-1)
        G.map_kinds = G.fused ? kMgMaxLevels + G.mg_top : G.mg_top + 1;
        int64_t cap[kMapKinds] = {};
        for (fg_scene* s : scenes)
            for (int k = 0; k < G.map_kinds; k++) cap[k] += map_blocks_host(s, k, G.fused);
        int64_t total = 0;
        for (int k = 0; k < G.map_kinds; k++) {
            G.map_off[k] = total;
            total += cap[k];
        }
        if (total > L->bmap_cap) {
            HIPCHK(hipStreamSynchronize(L->stream));
            if (L->bmap) hipFree(L->bmap);
            L->bmap = nullptr;
            L->bmap_cap = 0;
            HIPCHK(hipMalloc(&L->bmap, sizeof(int2) * std::max<int64_t>(total, 1)));
            L->bmap_cap = total;
        }
    }
    if (G.n == 1) {
        fg_scene* s = scenes[0];
        *s->hpin = s->h;
        HIPCHK(hipMemcpyAsync(s->d, s->hpin, sizeof(SceneDev), hipMemcpyHostToDevice, L->stream));
        G.d = s->d;
        return FG_OK;
    }
    FGCHK(ensure_group_buffers(L, G.n));
    SceneDev* dev = sub ? L->bsub : L->bdesc;
    SceneDev* host = sub ? L->bsub_host : L->bdesc_host;
    for (int k = 0; k < G.n; k++) host[k] = scenes[k]->h;
    HIPCHK(hipMemcpyAsync(dev, host, sizeof(SceneDev) * G.n, hipMemcpyHostToDevice, L->stream));
    G.d = dev;
    return FG_OK;
}

// Queue the copy of every scene's scalar block to the host; the caller
// synchronises and then calls scalars_arrived.
int queue_group_scalars(Group& G)
{
    Lane* L = G.L;
    if (G.n == 1) {
        fg_scene* s = G.sc[0];
        HIPCHK(hipMemcpyAsync(s->hsc, s->h.sc, sizeof(SceneScalars), hipMemcpyDeviceToHost, L->stream));
        return FG_OK;
    }
    hipLaunchKernelGGL(k_gather_scalars, dim3((unsigned)cdiv(G.n, 64)), dim3(64), 0, L->stream, G.d, G.n, L->bsc);
    HIPCHK(hipMemcpyAsync(L->bsc_host, L->bsc, sizeof(SceneScalars) * G.n, hipMemcpyDeviceToHost, L->stream));
    return FG_OK;
}
void scalars_arrived(Group& G)
{
    if (G.n > 1)
        for (int k = 0; k < G.n; k++) *G.sc[k]->hsc = G.L->bsc_host[k];
}

// Read every scene's scalar block into its pinned host mirror (synchronises).
int read_group_scalars(Group& G)
{
    FGCHK(queue_group_scalars(G));
    HIPCHK(hipStreamSynchronize(G.L->stream));
    scalars_arrived(G);
    return FG_OK;
}

// with_scalars: the scenes' scalar blocks come back with the flags (one
// synchronisation for both)
int poll_group(Group& G, int band, int32_t* flags, bool with_scalars = false)
{
    Lane* L = G.L;
    hipLaunchKernelGGL(k_group_flags, dim3(1), dim3(256), 0, L->stream, G.d, G.n, band, L->bflags);
    HIPCHK(hipMemcpyAsync(L->bflags_host, L->bflags, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, L->stream));
    if (with_scalars) FGCHK(queue_group_scalars(G));
    HIPCHK(hipStreamSynchronize(L->stream));
    for (int k = 0; k < 3; k++) flags[k] = L->bflags_host[k];
    if (with_scalars) scalars_arrived(G);
    return FG_OK;
}

// How the solve's per-iteration kernels launch: over a descriptor array on
// (blocks, scenes) grids, or flat over block maps (kinds as in kMapKinds).
struct SolveLaunch {
    const SceneDev* d = nullptr;
    unsigned ny = 0;                    // scenes in d
    bool flat = false;
    const int2* map[kMapKinds] = {};
    unsigned blocks[kMapKinds] = {};    // flat: total blocks of the kind; else blocks per scene
    dim3 grid(int kind) const { return flat ? dim3(blocks[kind]) : dim3(blocks[kind], ny); }
};

// A group's own descriptors on (blocks, scenes) grids.
SolveLaunch group_launch(const Group& G, const SceneDev* d, unsigned ny)
{
    SolveLaunch S;
    S.d = d;
    S.ny = ny;
    S.blocks[0] = G.g.pcg.x;
    S.blocks[1] = G.mg_grid_d0;
    if (G.fused)
        for (int l = 2; l < kMgMaxLevels; l++) {
            S.blocks[l] = G.rr_grid[l];
            S.blocks[pp_kind(l)] = G.pp_grid[l];
        }
    else
        for (int l = 1; l < kMgMaxLevels; l++) S.blocks[1 + l] = G.vc_grid[l];
    return S;
}

// Pack the scenes whose inner PCG still iterates and build their block maps;
// returns their count, the count of scenes whose FP64 test has not passed,
// and the launch description; the scalar blocks come back with them
// (synchronises once).
int pack_iterating(Group& G, SolveLaunch* S, int* count, int* open)
{
    Lane* L = G.L;
    hipLaunchKernelGGL(k_pack_iterating, dim3(1), dim3(64), 0, L->stream, G.d, G.n, L->bpack, L->bflags);
    const int kinds = G.map_kinds;
    if (kinds > 0) {
        MapOffsets offs{};
        for (int k = 0; k < kinds; k++) offs.off[k] = G.map_off[k];
        hipLaunchKernelGGL(k_build_maps, dim3(kinds), dim3(1024), 0, L->stream, L->bpack, L->bflags, L->bmap, offs, -1,
                           G.fused);
    }
    HIPCHK(hipMemcpyAsync(L->bflags_host, L->bflags, sizeof(int32_t) * (4 + kinds), hipMemcpyDeviceToHost, L->stream));
    FGCHK(queue_group_scalars(G));
    HIPCHK(hipStreamSynchronize(L->stream));
    scalars_arrived(G);
    *count = L->bflags_host[0];
    *open = L->bflags_host[2];
    if (kinds > 0) {
        *S = SolveLaunch{};
        S->d = L->bpack;
        S->ny = (unsigned)std::max(1, *count);
        S->flat = true;
        for (int k = 0; k < kinds; k++) {
            S->map[k] = L->bmap + G.map_off[k];
            S->blocks[k] = (unsigned)L->bflags_host[4 + k];
        }
    } else {
        *S = group_launch(G, L->bpack, (unsigned)std::max(1, *count));
        S->blocks[0] = (unsigned)std::max(1, L->bflags_host[1]);
    }
    return FG_OK;
}


// Flat launches over every scene of the group (a solve's first sweep, where
// all of them iterate): the maps are built on the device from the group's
// own descriptors and their totals are known here, so nothing synchronises.
void all_scene_launch(Group& G, SolveLaunch* S)
{
    Lane* L = G.L;
    const int kinds = G.map_kinds;
    MapOffsets offs{};
    for (int k = 0; k < kinds; k++) offs.off[k] = G.map_off[k];
    hipLaunchKernelGGL(k_build_maps, dim3(kinds), dim3(1024), 0, L->stream, G.d, L->bflags, L->bmap, offs, G.n,
                       G.fused);
    *S = SolveLaunch{};
    S->d = G.d;
    S->ny = (unsigned)G.n;
    S->flat = true;
    for (int k = 0; k < kinds; k++) {
        int64_t total = 0;
        for (const fg_scene* s : G.sc) total += map_blocks_host(s, k, G.fused);
        S->map[k] = L->bmap + G.map_off[k];
        S->blocks[k] = (unsigned)total;
    }
}

// ---------------------------------------------------------------------------
// stage timing
// ---------------------------------------------------------------------------
struct Timer {
    fg_ctx* ctx;
    Group* G;
    hipEvent_t a = nullptr;
    Timer(fg_ctx* c, Group* g) : ctx(c), G(g)
    {
        if (ctx->timing) {
            hipEventCreate(&a);
            hipEventRecord(a, G->L->stream);
        }
    }
    // the group's time for the stage is charged to every scene in it
    void lap(int stage)
    {
        if (!ctx->timing) return;
        hipEvent_t b;
        hipEventCreate(&b);
        hipEventRecord(b, G->L->stream);
        hipEventSynchronize(b);
        float ms = 0;
        hipEventElapsedTime(&ms, a, b);
        for (fg_stats* st : G->st)
            if (st) st->stage_ms[stage] += ms;
        hipEventDestroy(a);
        a = b;
    }
    ~Timer()
    {
        if (a) hipEventDestroy(a);
    }
};

// ---------------------------------------------------------------------------
// per-launch kernel timing (events bracket each launch on its own stream)
// ---------------------------------------------------------------------------
// Work units of a kernel family: always counted; also the timed family's
// units when it is the one being timed.
inline void add_work(Lane* L, int timed, int family, int64_t units)
{
    L->work[family] += units;
    if (timed == family) L->prof_units += units;
}

// The timed family takes in the kernel: itself, or a family made of it.
inline bool prof_family_has(int family, int kernel)
{
    return family == kernel || (family == FG_K_PARTICLES && (kernel == FG_K_ADVECT || kernel == FG_K_G2P));
}

struct KProf {
    Lane* L;
    int timed, kernel;
    bool on, single;
    // single: the events are attached to the next kernel launch itself
    // (hipExtLaunchKernelGGL) and bracket exactly that kernel; otherwise they
    // are recorded around the whole multi-kernel stage.  Work units are
    // counted only when the kernel is the timed family itself (a composite
    // family counts its bytes once per substep).
    KProf(const Group& G, int kernel_, bool single_kernel = false)
        : L(G.L), timed(G.ctx->prof_kernel), kernel(kernel_), on(prof_family_has(timed, kernel_)),
          single(single_kernel)
    {
        if (!on) return;
        hipEvent_t e0 = take(), e1 = take();
        if (single) next_launch_events() = LaunchEvents{e0, e1};
        else hipEventRecord(e0, L->stream);
    }
    hipEvent_t take()
    {
        if (L->ev_used == L->ev_pool.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            L->ev_pool.push_back(e);
        }
        return L->ev_pool[L->ev_used++];
    }
    void done(int64_t units)
    {
        if (kernel >= 0) add_work(L, timed, kernel, units);
        if (!on) return;
        if (!single) hipEventRecord(L->ev_pool[L->ev_used - 1], L->stream);
    }
};

// ---------------------------------------------------------------------------
// the solve (shared by project and fg_poisson_solve); b is in place.  All
// scenes of the group run the same sweep and iteration index in lock step;
// a scene whose solve has stopped skips the launches that remain.
// ---------------------------------------------------------------------------
// Whether k_mg_up0 forms the neighbours' smoothing starts in an LDS window
// over the chunk's rows rather than per cell: for a single scene whose
// chunks span four rows or more (on one-row chunks and on batches of small
// scenes the per-cell kernel is faster).  The same window for k_pcg_a's
// direction (a 4096^2 projection) took 70 us a launch against 51 us per
// cell: there the neighbours' rows it saves come back from the last-level
// cache, and the window's prologue and barriers cost more.
bool ring_window(const Group& G) { return G.n == 1 && G.sc[0]->pcg_chunk >= 4 * (int64_t)G.sc[0]->J; }

// One V-cycle z = M r over the scenes of S whose solve uses it; which: the
// residual buffer (0: cr, 1: cr2).  Level-l launches cover the scenes that
// have level l; a kind with no blocks is not launched.
void launch_vcycle(const Group& G, hipStream_t st, const SolveLaunch& S, int which)
{
    // levels 1 .. t-1 one launch per step, levels t .. n-1 in k_mg_coarse
    // (t per scene; the maps and grids hold a level's blocks only for the
    // scenes that launch it)
    const int top = G.mg_top;
    if (S.blocks[1]) {
        KProf kp(G, FG_K_MG_DOWN0, true);
        launch_mg_down0(S.grid(1), st, S.d, S.map[1], which);
        kp.done(0);
    }
    if (G.fused) {
        // level l's residual restricted to level l+1 (written out for the
        // coarse block at l+1 = t): kRRTile tiles of level l+1
        for (int l = 1; l < top; l++) {
            if (!S.blocks[l + 1]) continue;
            KProf kp(G, FG_K_MG_RESID, true);
            launch_mg_rr(S.grid(l + 1), st, S.d, S.map[l + 1], l);
            kp.done(0);
        }
    } else {
        for (int l = 1; l < top; l++) {
            if (S.blocks[1 + l]) {
                // the profiling family times the largest residual launch (level 1)
                KProf kp(G, l == 1 ? FG_K_MG_RESID : -2, true);
                launch_mg_resid(S.grid(1 + l), st, S.d, S.map[1 + l], l);
                kp.done(0);
            }
            if (l + 1 < top && S.blocks[2 + l]) launch_mg_restrict(S.grid(2 + l), st, S.d, S.map[2 + l], l + 1);
        }
    }
    {
        KProf kp(G, FG_K_MG_COARSE, true);
        launch_mg_coarse(dim3(1, S.ny), st, S.d);
        kp.done(0);
    }
    if (G.fused) {
        // prolongation and post-smoothing of level l: tiles of level l
        for (int l = top - 1; l >= 1; l--) {
            if (!S.blocks[pp_kind(l)]) continue;
            KProf kp(G, FG_K_MG_PP, true);
            launch_mg_pp(S.grid(pp_kind(l)), st, S.d, S.map[pp_kind(l)], l);
            kp.done(0);
        }
    } else {
        for (int l = top - 1; l >= 1; l--) {
            if (!S.blocks[1 + l]) continue;
            launch_mg_prolong(S.grid(1 + l), st, S.d, S.map[1 + l], l);
            launch_mg_post(S.grid(1 + l), st, S.d, S.map[1 + l], l);
        }
    }
    KProf kp(G, FG_K_MG_UP0, true);
    launch_mg_up0(S.grid(0), st, S.d, S.map[0], which, G.max_j, ring_window(G));
    kp.done(0);
}

// One PCG iteration it: p, q = A'p and p.q; r, x; z = M r (the V-cycle, when
// the group uses it).
void launch_iteration(const Group& G, hipStream_t st, const SolveLaunch& S, int it)
{
    {
        KProf kp(G, FG_K_PCG_A, true);
        launch_pcg_a(S.grid(0), st, S.d, S.map[0], it);
        kp.done(0);   // units: at the end of the solve
    }
    {
        KProf kp(G, FG_K_PCG_B, true);
        launch_pcg_b(S.grid(0), st, S.d, S.map[0], it);
        kp.done(0);
    }
    if (G.mg_levels) {
        // the new residual: iteration it writes cr2 for even it, cr for odd
        KProf kp(G, FG_K_VCYCLE);
        launch_vcycle(G, st, S, (it & 1) ? 0 : 1);
        kp.done(0);
    }
}

// Algorithmic bytes of one k_mg_coarse block (one scene, one V-cycle): the
// dinv of its levels (4 B a cell), its input (r_1, or the finer level's y it
// restricts), the y it writes (4 B a level-t cell), once each the 25
// coefficients of the rows above the bottom that read stored coefficients,
// and the bottom's coefficients (100 B a row: its sweeps read every row's).
int64_t mg_coarse_bytes(const fg_scene* s)
{
    const MgLevels& M = s->mgh;
    const int t = M.t, nb = M.n - 1;
    auto cells = [&](int l) { return (int64_t)M.lv[l].I * M.lv[l].J; };
    int64_t b = 4 * (int64_t)M.tcells + 4 * cells(t) + 4 * cells(t == 1 ? 1 : t - 1);
    for (int l = t; l < nb; l++) b += 4 * kMgStencil * (int64_t)s->hsc->mg_stored[l];
    return b + 4 * kMgStencil * cells(nb);
}

int run_solve(Group& G, bool rhs_from_grid)
{
    fg_ctx* ctx = G.ctx;
    hipStream_t st = G.L->stream;
    const Geo& g = G.g;
    launch_solve_reset(g.one, st, G.d);
    if (rhs_from_grid) {
        KProf kp(G, FG_K_RHS, true);
        launch_rhs(g.pcg, st, G.d);
        kp.done(G.cells);
    } else {
        launch_solve_begin(g.pcg, st, G.d);
    }
    // multigrid hierarchy of the matrix snapshot (fixed for the whole solve)
    // per level: Galerkin product, then its symmetrisation together with
    // the next level's activity marks
    if (G.mg_levels > 1) launch_mg_mark(dim3(G.mg_grid[1], (unsigned)G.n), st, G.d, 1);
    for (int l = 1; l < G.mg_levels; l++) {
        const dim3 gl(G.mg_grid[l], (unsigned)G.n);
        launch_mg_galerkin(gl, st, G.d, l);
        if (l + 1 < G.mg_levels)
            launch_mg_symmark(dim3(G.sym_grid[l] + G.mg_grid[l + 1], (unsigned)G.n), st, G.d, l, (int)G.sym_grid[l]);
        else
            launch_mg_sym(dim3(G.sym_grid[l], (unsigned)G.n), st, G.d, l);
    }
    // A poll is a host round trip; launches after every scene of the group
    // has stopped run empty.  The first poll of a solve comes after as many
    // iterations as its scenes' first sweeps took the last time (plus the
    // launch that sees the stop), later ones every kPollNext.
    constexpr int kPollFirst = 16, kPollNext = 2, kPollMax = 64;
    const int kind = rhs_from_grid ? 0 : 1;
    int hint = 0;
    for (fg_scene* s : G.sc) hint = std::max(hint, s->iters_hint[kind]);
    // Polls bring the scalar blocks back with their flags.  After each chunk
    // of iterations the FP64 test runs on the scenes whose inner solve has
    // stopped (spec passes), so the poll that sees the last one stop also
    // sees whether every scene met its tolerance, and the solve ends there.
    int32_t flags[3];
    bool scalars_fresh = false;   // the last poll read the scalars, nothing ran since
    for (int sweep = 0; sweep <= kMaxSweeps; sweep++) {
        launch_refine_residual(g.pcg, st, G.d, 0);
        launch_cg_init(g.pcg, st, G.d, sweep, 0);
        scalars_fresh = false;
        // the scenes whose solve goes on: several are packed (and re-packed
        // at every poll as they stop); one keeps its own descriptor
        SolveLaunch S;
        int count = 0, open = 0;
        if (G.n > 1 && sweep == 0 && G.map_kinds > 0) {
            // every scene takes part (one whose FP64 test already passed
            // exits each kernel at once); the first poll repacks
            all_scene_launch(G, &S);
            count = G.n;
        } else if (G.n > 1) {
            FGCHK(pack_iterating(G, &S, &count, &open));
            scalars_fresh = true;
        } else if (sweep == 0) {
            // one scene: it takes part (its kernels exit at once if the FP64
            // test already passed); the first poll tells
            count = 1;
            S = group_launch(G, G.d, 1);
        } else {
            FGCHK(poll_group(G, -1, flags, true));
            scalars_fresh = true;
            count = flags[1];
            S = group_launch(G, G.d, 1);
        }
        if (count == 0) break;
        if (G.mg_levels) {
            KProf kp(G, FG_K_VCYCLE);
            launch_vcycle(G, st, S, 0);
            kp.done(0);
        }
        int it = 0;
        int chunk = sweep > 0 ? kPollNext : hint > 0 ? std::min(hint + 1, kPollMax) : kPollFirst;
        while (count > 0) {
            for (int k = 0; k < chunk; k++, it++) launch_iteration(G, st, S, it);
            launch_refine_residual(g.pcg, st, G.d, 1);
            launch_cg_init(g.pcg, st, G.d, sweep + 1, 1);
            if (G.n > 1) {
                FGCHK(pack_iterating(G, &S, &count, &open));
            } else {
                FGCHK(poll_group(G, -1, flags, true));
                count = flags[1];
                open = flags[0];
            }
            scalars_fresh = true;
            if (count > 0 && it > G.max_inner + kPollMax) return fail(FG_E_STATE, "inner solve did not stop at its cap");
            chunk = kPollNext;
        }
        if (open == 0) break;   // every scene passed its FP64 test
    }
    if (!scalars_fresh) FGCHK(read_group_scalars(G));
    for (fg_scene* s : G.sc) {
        const SceneScalars& h = *s->hsc;
        s->iters_hint[kind] = h.sweeps > 1 ? h.total_iters - h.inner_iters : h.total_iters;
    }
    // work units of the launches that iterated (the rest exited at once)
    const int timed = ctx->prof_kernel;
    for (fg_scene* s : G.sc) {
        const SceneScalars& h = *s->hsc;
        add_work(G.L, timed, FG_K_PCG_A, (int64_t)h.total_iters * s->C);
        add_work(G.L, timed, FG_K_PCG_B, (int64_t)h.total_iters * s->C);
        if (s->h.precond != kPrecondMG) continue;
        const int64_t cycles = (int64_t)h.total_iters + h.sweeps;   // one per iteration and per sweep start
        add_work(G.L, timed, FG_K_VCYCLE, cycles * s->C);
        const MgLevel& L1 = s->mgh.lv[1];
        const int64_t C1 = (int64_t)L1.I * L1.J;
        // the per-level launches (scenes with a level below the coarse
        // block's).  Fused, every level l < t: k_mg_rr reads r and dinv and
        // writes r_{l+1} (9 B a level-l cell), k_mg_pp reads act, r, dinv,
        // y_{l+1} and writes y (14 B), and each reads a stored row's 100 B.
        // Per-level kernels: the level-1 residual launch alone (r, dinv, y).
        if (s->mgh.t >= 2) {
            if (G.fused) {
                int64_t brr = 0, bpp = 0;
                for (int l = 1; l < s->mgh.t; l++) {
                    const int64_t cl = (int64_t)s->mgh.lv[l].I * s->mgh.lv[l].J;
                    const int64_t st = 4 * kMgStencil * (int64_t)h.mg_stored[l];
                    brr += 9 * cl + st;
                    bpp += 14 * cl + st;
                }
                add_work(G.L, timed, FG_K_MG_RESID, cycles * brr);
                add_work(G.L, timed, FG_K_MG_PP, cycles * bpp);
            } else {
                add_work(G.L, timed, FG_K_MG_RESID, cycles * (12 * C1 + 4 * kMgStencil * (int64_t)h.mg_stored[1]));
            }
        }
        add_work(G.L, timed, FG_K_MG_COARSE, cycles * mg_coarse_bytes(s));
        add_work(G.L, timed, FG_K_MG_DOWN0, cycles * (5 * s->C + 5 * C1));
        add_work(G.L, timed, FG_K_MG_UP0, cycles * (9 * s->C + 4 * C1));
    }
    // a scene whose solve saw NaN/Inf stopped iterating (its flags say so);
    // the caller quarantines it, the others are unaffected
    return FG_OK;
}

// ---------------------------------------------------------------------------
// stages of FlipSolver::step
// ---------------------------------------------------------------------------
// applyViscosity and the second project() of FlipSolver::step
// (flipsolver2d.cpp:446-453) for the scenes of G with viscosity, as a
// sub-group: the light model's matrix, its Jacobi-preconditioned CG on U and
// V together (polled every kViscPoll iterations), the result, then the
// pressure solve again with the substep's matrix and its plain application.
constexpr int kViscPoll = 8;
constexpr int kViscMaxIters = 20000;
constexpr double kViscTol = 1e-12;   // relative residual of each system (FP64)
int run_viscosity(Group& G)
{
    std::vector<fg_scene*> vs;
    std::vector<fg_stats*> vst;
    for (int k = 0; k < G.n; k++)
        if (G.sc[k]->h.visc_on) {
            vs.push_back(G.sc[k]);
            vst.push_back(G.st[k]);
        }
    Group V;
    FGCHK(make_group(G.ctx, vs, vst, V, true, G.L));
    hipStream_t st = V.L->stream;
    const dim3 one(1, V.n);
    const double tol2 = kViscTol * kViscTol;
    launch_visc_setup(V.g.cells, st, V.d);
    launch_visc_red(one, st, V.d, 0, tol2);
    int it = 0, open = 1;
    KProf kp(V, FG_K_VISC);
    while (open && it < kViscMaxIters) {
        for (int k = 0; k < kViscPoll; k++) {
            launch_visc_spmv(V.g.cells, st, V.d);
            launch_visc_red(one, st, V.d, 1, tol2);
            launch_visc_update(V.g.cells, st, V.d);
            launch_visc_red(one, st, V.d, 2, tol2);
            launch_visc_dir(V.g.cells, st, V.d);
        }
        it += kViscPoll;
        launch_visc_flags(st, V.d, V.n, V.L->bflags);
        HIPCHK(hipMemcpyAsync(V.L->bflags_host, V.L->bflags, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        open = V.L->bflags_host[0];
    }
    // per cell and iteration, U and V together: spmv 24 (d, e, n) + 16 (p)
    // + 16 (q); update 8 (d) + 16 x 4 read (p, x, r, q) + 16 x 2 written;
    // direction 8 + 16 x 2 read + 16 written
    kp.done(216 * V.cells * (int64_t)it);
    launch_visc_apply(V.g.cells, st, V.d);
    for (int k = 0; k < V.n; k++) {
        double sc[kViscScalars];
        HIPCHK(hipMemcpyAsync(sc, V.sc[k]->h.vsc, sizeof(sc), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (fg_stats* stats = V.st[k]) {
            stats->viscosity_iters = std::max(stats->viscosity_iters, (int32_t)std::max(sc[14], sc[15]));
            if (sc[12] == 0.0 || sc[13] == 0.0) stats->converged = 0;
        }
    }
    // project() again: rhs from the viscous grid, the same matrix snapshot
    const int r = run_solve(V, true);
    for (int k = 0; k < V.n; k++) {
        fg_stats* stats = V.st[k];
        if (!stats) continue;
        const SceneScalars& sc = *V.sc[k]->hsc;
        stats->pressure_iters = std::max(stats->pressure_iters, sc.total_iters);
        stats->pressure_sweeps = std::max(stats->pressure_sweeps, sc.sweeps);
        stats->pressure_rel_resid = sc.bb > 0 ? std::sqrt(sc.rr64 / sc.bb) : 0.0;
        if (stats->pressure_rel_resid > V.sc[k]->h.tol) stats->converged = 0;
    }
    FGCHK(r);
    launch_apply_solids(V.g.cells, st, V.d, 0);
    return FG_OK;
}

// The band launches of a level-set extrapolation whose markers
// k_levelset_init has set.  A band launch after the last one needed exits
// at once but still dispatches its blocks: the first poll comes after one
// band more than the group's scenes ran last substep, later ones every kChunk.
int levelset_bands(Group& G, bool outside)
{
    constexpr int kChunk = 8;
    // the inside chain's hint is the band count the last one reached; the
    // outside chain (NB-FLIP, far longer) keeps the count it launched
    int hint = 0;
    for (fg_scene* s : G.sc) hint = std::max(hint, outside ? s->ls_hint_out : (s->ls_hint > 0 ? s->ls_hint + 1 : 0));
    int chunk = hint > 0 ? hint : kChunk;
    int32_t flags[3];
    int launched = 0;
    for (int band = 0; band < G.max_bands; band += chunk, chunk = kChunk) {
        const int end = std::min(G.max_bands, band + chunk);
        for (int k = band; k < end; k++) launch_levelset_band(G.g.ls, G.L->stream, G.d, k, outside);
        launched = end;
        FGCHK(poll_group(G, end - 1, flags));
        if (!flags[2]) break;
    }
    if (outside)
        for (fg_scene* s : G.sc) s->ls_hint_out = launched;
    return FG_OK;
}

int run_stage(Group& G, int stage)
{
    fg_ctx* ctx = G.ctx;
    hipStream_t st = G.L->stream;
    const Geo& g = G.g;
    const SceneDev* d = G.d;
    switch (stage) {
    case FG_STAGE_ADVECT: {
        KProf kp(G, FG_K_ADVECT, true);
        launch_advect(g.parts, st, d, 1);
        kp.done(G.parts);
        break;
    }
    case FG_STAGE_REBIN: {
        launch_stencil(g.cells, st, d);          // matrix from the pre-update materials
        KProf kp(G, FG_K_SORT);
        launch_scan_local(g.scan, st, d);
        launch_scan_top(g.scan1, st, d);
        launch_scan_add(g.scan, st, d);
        launch_scatter(g.parts, st, d);
        launch_fixup(g.cells, st, d);
        launch_gather(g.parts, st, d);
        launch_finalize(g.one, st, d, 1);
        kp.done(G.parts);
        break;
    }
    case FG_STAGE_DENSITY: {
        // FlipSolver::densityCorrection (flipsolver2d.cpp:164-186) for the
        // scenes that run it, as a sub-group with its own descriptors
        std::vector<fg_scene*> ds;
        std::vector<fg_stats*> dst;
        for (int k = 0; k < G.n; k++)
            if (G.sc[k]->p.density_mode != FG_DENSITY_OFF) {
                ds.push_back(G.sc[k]);
                dst.push_back(G.st[k]);
            }
        if (ds.empty()) break;
        Group D;
        FGCHK(make_group(ctx, ds, dst, D, true, G.L));
        launch_density_rhs(D.g.cells, st, D.d);
        FGCHK(run_solve(D, false));
        for (int k = 0; k < D.n; k++)
            if (D.st[k]) D.st[k]->density_iters = std::max(D.st[k]->density_iters, D.sc[k]->hsc->total_iters);
        // move the particles (where the gate opens) and restore (cell, id) order
        launch_density_adjust(D.g.parts, st, D.d);
        launch_scan_local(D.g.scan, st, D.d);
        launch_scan_top(D.g.scan1, st, D.d);
        launch_scan_add(D.g.scan, st, D.d);
        launch_scatter(D.g.parts, st, D.d);
        launch_fixup(D.g.cells, st, D.d);
        launch_gather(D.g.parts, st, D.d);
        launch_finalize(D.g.one, st, D.d, 4);
        break;
    }
    case FG_STAGE_P2G: {
        {
            KProf kp(G, FG_K_P2G, true);
            launch_p2g(g.p2g, st, d);
            kp.done(0);   // bytes: stage_bytes, once the substep's counts are back
        }
        // particleToGrid's centeredParamsToGrid: the viscosity grid
        if (G.visc) launch_visc_grid(g.cells, st, d);
        break;
    }
    case FG_STAGE_SDF: {
        {
            KProf kp(G, FG_K_SDF, true);
            launch_sdf_materials(g.bins, st, d);
            kp.done(0);
        }
        // afterTransfer (flipsolver2d.cpp:390-410), before the level set and
        // the extrapolation
        if (G.sources) launch_after_transfer(g.cells, st, d);
        break;
    }
    case FG_STAGE_LEVELSET: {
        KProf kp(G, FG_K_LEVELSET);
        launch_levelset_init(g.ls, st, d, false);
        launch_sdf_exact(g.bins, st, d);   // only where no cell is outside
        FGCHK(levelset_bands(G, false));
        kp.done(G.cells);
        break;
    }
    case FG_STAGE_EXTRAPOLATE: {
        KProf kp(G, FG_K_EXTRAPOLATE);
        launch_extrapolate(g.extrap, st, d, 0);  // u -> su (saved)
        kp.done(G.faces);
        break;
    }
    case FG_STAGE_FORCES:
        launch_forces(g.faces, st, d);
        break;
    case FG_STAGE_PROJECT: {
        const int r = run_solve(G, true);
        for (int k = 0; k < G.n; k++) {
            fg_stats* stats = G.st[k];
            if (!stats) continue;
            const SceneScalars& sc = *G.sc[k]->hsc;
            stats->pressure_iters = std::max(stats->pressure_iters, sc.total_iters);
            stats->pressure_sweeps = std::max(stats->pressure_sweeps, sc.sweeps);
            stats->pressure_rel_resid = sc.bb > 0 ? std::sqrt(sc.rr64 / sc.bb) : 0.0;
            if (stats->pressure_rel_resid > G.sc[k]->h.tol) stats->converged = 0;
        }
        FGCHK(r);
        {
            KProf kp(G, FG_K_APPLY, true);
            launch_apply_solids(g.cells, st, d, 1);
            kp.done(G.cells);
        }
        if (G.visc) FGCHK(run_viscosity(G));
        break;
    }
    case FG_STAGE_SOLIDS:
        // fused into the pressure application kernel (k_apply_solids)
        break;
    case FG_STAGE_G2P: {
        {
            KProf kp(G, FG_K_EXTRAPOLATE);
            launch_extrapolate(g.extrap, st, d, 1);  // u -> uf (final)
            kp.done(G.faces);
        }
        KProf kp(G, FG_K_G2P, true);
        launch_g2p_count(g.parts, st, d);
        kp.done(G.parts);
        launch_finalize(g.one, st, d, 2);
        break;
    }
    case FG_STAGE_COUNT:
        // countParticles runs inside k_g2p_count; reseedParticles
        // (flipsolver2d.cpp:627-680) fills the SOURCE cells of the scenes
        // that have them, NB-FLIP's (nbflipsolver.cpp:115-201) its band
        // cells: per-cell additions, their scan, the writes
        if (!G.sources && !G.nb) break;
        launch_reseed_count(g.cells, st, d);
        launch_scan_local(g.scan, st, d);
        launch_scan_top(g.scan1, st, d);
        launch_scan_add(g.scan, st, d);
        launch_reseed_write(g.cells, st, d);
        launch_reseed_commit(g.one, st, d);
        break;
    default:
        return fail(FG_E_INVALID, "bad stage");
    }
    return FG_OK;
}

// Algorithmic bytes of the byte-counted particle-stage families for one
// substep of every scene of G (P = the particles this substep sorted, the
// input of P2G, the fluid SDF and G2P; see fg_kernel_id).
void stage_bytes(Group& G)
{
    const int timed = G.ctx->prof_kernel;
    for (fg_scene* s : G.sc) {
        const int64_t P = s->hsc->n_particles, F = s->FU + s->FV, C = s->C;
        add_work(G.L, timed, FG_K_P2G, 16 * P + 5 * F);
        add_work(G.L, timed, FG_K_SDF, 8 * P + 6 * C);
        add_work(G.L, timed, FG_K_PARTICLES, 36 * P + 8 * F + 4 * C);
    }
}

// Take a scene out of stepping until set_state gives it a new state.
void quarantine(fg_scene* s, fg_stats* st)
{
    s->status = FG_E_NUMERIC;
    s->vmax_ready = false;
    if (st) {
        st->status = FG_E_NUMERIC;
        st->converged = 0;
    }
}

// After a substep's last stage: commit the counts the kernels produced;
// a scene whose solve saw NaN/Inf is quarantined.
void commit_substep(Group& G, float* dts)
{
    for (int k = 0; k < G.n; k++) {
        fg_scene* s = G.sc[k];
        const SceneScalars& sc = *s->hsc;
        if (sc.numeric_bad) quarantine(s, G.st[k]);
        const int64_t n_start = s->n_live;
        s->epoch++;
        s->next_pid += (uint32_t)sc.reseeded;
        s->ls_hint = sc.ls_bands;
        s->n_live = sc.n_particles - sc.erased;
        s->n_bound = sc.n_particles;
        s->vmax_ready = true;
        if (fg_stats* stats = G.st[k]) {
            stats->substeps += 1;
            stats->particle_substeps += n_start;
            stats->particles = s->n_live;
            stats->particles_killed += sc.killed;
            stats->particles_erased += sc.erased;
            stats->density_applied += sc.density_applied;
            stats->particles_reseeded += sc.reseeded;
            stats->last_dt = dts[k];
        }
    }
}

// NBFlipSolver::step (nbflipsolver.cpp:26-64) for a group of NB-FLIP
// scenes.  Stage times are charged to the FlipSolver stage of the same work.
int run_nb_substep(Group& G, Timer& tm)
{
    hipStream_t st = G.L->stream;
    const Geo& g = G.g;
    const SceneDev* d = G.d;
    // advect(): FlipSolver::advect with the band prune, then the Euler pass
    // over the grid and the level set as they are now (66-113)
    {
        KProf kp(G, FG_K_ADVECT, true);
        launch_advect(g.parts, st, d, 1);
        kp.done(G.parts);
    }
    {
        KProf kp(G, FG_K_NB_EULER, true);
        launch_nb_euler(g.nodes, st, d);
        int64_t bytes = 0;
        for (fg_scene* s : G.sc) bytes += 8 * (s->FU + s->FV + s->C);
        kp.done(bytes);
    }
    tm.lap(FG_STAGE_ADVECT);
    {   // pruneParticles + rebin
        KProf kp(G, FG_K_SORT);
        launch_scan_local(g.scan, st, d);
        launch_scan_top(g.scan1, st, d);
        launch_scan_add(g.scan, st, d);
        launch_scatter(g.parts, st, d);
        launch_fixup(g.cells, st, d);
        launch_gather(g.parts, st, d);
        launch_finalize(g.one, st, d, 1);
        kp.done(G.parts);
    }
    tm.lap(FG_STAGE_REBIN);
    {   // particleToGrid (with NBFlipSolver's centeredParamsToGrid), extrapolate(10): the saved grid
        KProf kp(G, FG_K_P2G, true);
        launch_p2g(g.p2g, st, d);
        kp.done(0);
    }
    if (G.visc) launch_visc_grid(g.cells, st, d);
    tm.lap(FG_STAGE_P2G);
    {
        KProf kp(G, FG_K_EXTRAPOLATE);
        launch_extrapolate(g.extrap, st, d, 0);
        kp.done(G.faces);
    }
    tm.lap(FG_STAGE_EXTRAPOLATE);
    // gridUpdate (213-226): updateSdf, extrapolateLevelsetOutside,
    // afterTransfer (the source faces, updateGridFromSources, the
    // combination), extrapolateLevelsetInside, updateMaterials, body forces
    {
        KProf kp(G, FG_K_SDF, true);
        launch_sdf_materials(g.bins, st, d);
        kp.done(0);
    }
    tm.lap(FG_STAGE_SDF);
    {
        KProf kp(G, FG_K_LEVELSET);
        launch_levelset_init(g.ls, st, d, true);
        FGCHK(levelset_bands(G, true));
        launch_nb_combine_sdf(g.cells, st, d);
        launch_nb_combine_vel(g.faces, st, d);
        launch_levelset_init(g.ls, st, d, false);
        FGCHK(levelset_bands(G, false));
        launch_nb_materials(g.cells, st, d);
        kp.done(2 * G.cells);
    }
    tm.lap(FG_STAGE_LEVELSET);
    // the pressure matrix from the materials as they are now (45-47)
    launch_stencil(g.cells, st, d);
    tm.lap(FG_STAGE_FORCES);
    for (int stage : {FG_STAGE_PROJECT, FG_STAGE_SOLIDS, FG_STAGE_G2P, FG_STAGE_COUNT}) {
        FGCHK(run_stage(G, stage));
        tm.lap(stage);
    }
    return FG_OK;
}

// One step for every scene of `scenes` (one solver class), each with its own dt.
int substep_one_class(fg_ctx* ctx, Lane* L, const std::vector<fg_scene*>& scenes, const std::vector<float>& dts,
                      const std::vector<fg_stats*>& stats)
{
    for (size_t k = 0; k < scenes.size(); k++) {
        set_substep_scalars(scenes[k], dts[k]);
        scenes[k]->last_stage = -1;
    }
    Group G;
    FGCHK(make_group(ctx, scenes, stats, G, false, L));
    if (G.n == 0) return FG_OK;
    hipLaunchKernelGGL(k_substep_begin, G.g.one, dim3(64), 0, G.L->stream, G.d);
    Timer tm(ctx, &G);
    if (G.nb) {
        FGCHK(run_nb_substep(G, tm));
    } else {
        static const int order[] = {FG_STAGE_ADVECT, FG_STAGE_REBIN, FG_STAGE_DENSITY, FG_STAGE_P2G,
                                    FG_STAGE_SDF, FG_STAGE_LEVELSET, FG_STAGE_EXTRAPOLATE, FG_STAGE_FORCES,
                                    FG_STAGE_PROJECT, FG_STAGE_SOLIDS, FG_STAGE_G2P, FG_STAGE_COUNT};
        for (int stage : order) {
            FGCHK(run_stage(G, stage));
            tm.lap(stage);
        }
    }
    HIPCHK(hipGetLastError());
    FGCHK(read_group_scalars(G));
    stage_bytes(G);
    std::vector<float> d(dts);
    commit_substep(G, d.data());
    return FG_OK;
}

// One FlipSolver::step (NBFlipSolver::step for NB-FLIP scenes) for every
// scene of `scenes`, each with its own dt; the two solver classes run as
// separate groups, one after the other on the lane.
int substep_group(fg_ctx* ctx, Lane* L, const std::vector<fg_scene*>& scenes, const std::vector<float>& dts,
                  const std::vector<fg_stats*>& stats)
{
    bool mixed = false;
    for (fg_scene* s : scenes) mixed = mixed || (s->h.nb != scenes[0]->h.nb);
    if (!mixed) return substep_one_class(ctx, L, scenes, dts, stats);
    for (int cls = 0; cls < 2; cls++) {
        std::vector<fg_scene*> sc;
        std::vector<float> dt;
        std::vector<fg_stats*> st;
        for (size_t k = 0; k < scenes.size(); k++)
            if (scenes[k]->h.nb == cls) {
                sc.push_back(scenes[k]);
                dt.push_back(dts[k]);
                st.push_back(k < stats.size() ? stats[k] : nullptr);
            }
        FGCHK(substep_one_class(ctx, L, sc, dt, st));
    }
    return FG_OK;
}

// max |v|^2 for scenes whose last substep did not leave it on the device
int ensure_vmax(fg_ctx* ctx, Lane* L, const std::vector<fg_scene*>& scenes)
{
    std::vector<fg_scene*> need;
    for (fg_scene* s : scenes)
        if (!s->vmax_ready) need.push_back(s);
    if (need.empty()) return FG_OK;
    Group G;
    FGCHK(make_group(ctx, need, {}, G, false, L));
    launch_max_v2(G.g.parts, G.L->stream, G.d);
    launch_finalize(G.g.one, G.L->stream, G.d, 2);
    FGCHK(read_group_scalars(G));
    for (fg_scene* s : need) s->vmax_ready = true;
    return FG_OK;
}

// FlipSolver::stepFrame (flipsolver2d.cpp:464-500) for every scene: each
// scene splits its frame by its own CFL dt; the scenes still inside their
// frame advance together on lane L, one launch per phase.
int step_frames_lane(fg_ctx* ctx, Lane* L, const std::vector<fg_scene*>& scenes, const std::vector<fg_stats*>& stats,
                     std::chrono::steady_clock::time_point t0)
{
    const size_t n = scenes.size();
    std::vector<float> t(n, 0.f);
    std::vector<int> count(n, 0);
    std::vector<char> finished(n, 0);
    for (size_t k = 0; k < n; k++)
        if (scenes[k]->status != FG_OK) {   // quarantined earlier: skipped
            finished[k] = 1;
            if (stats[k]) {
                stats[k]->status = scenes[k]->status;
                stats[k]->converged = 0;
            }
        }
    for (;;) {
        std::vector<fg_scene*> need;
        for (size_t k = 0; k < n; k++)
            if (!finished[k]) need.push_back(scenes[k]);
        if (need.empty()) break;
        FGCHK(ensure_vmax(ctx, L, need));
        std::vector<fg_scene*> act;
        std::vector<fg_stats*> ast;
        std::vector<size_t> idx;
        std::vector<float> dts;
        for (size_t k = 0; k < n; k++) {
            if (finished[k]) continue;
            fg_scene* s = scenes[k];
            // FlipSolver::maxParticleVelocity + the substep split, flipsolver2d.cpp:476-490
            const float frame_dt = 1.f / s->p.fps;
            float v2;
            std::memcpy(&v2, &s->hsc->max_v2_bits, 4);
            const float vel = std::sqrt(v2);
            if (!std::isfinite(vel)) {   // a NaN/Inf particle velocity: no dt can be formed
                quarantine(s, stats[k]);
                finished[k] = 1;
                continue;
            }
            float mss = s->p.cfl_number / (vel + 1e-15f);
            if (t[k] + mss >= frame_dt || count[k] == (s->p.max_substeps - 1)) {
                mss = frame_dt - t[k];
                finished[k] = 1;
            } else if (t[k] + 2.f * mss >= frame_dt) {
                mss = 0.5f * (frame_dt - t[k]);
            }
            if (stats[k]) stats[k]->max_velocity = vel;
            act.push_back(s);
            ast.push_back(stats[k]);
            idx.push_back(k);
            dts.push_back(mss);
        }
        if (act.empty()) continue;
        FGCHK(substep_group(ctx, L, act, dts, ast));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (size_t a = 0; a < act.size(); a++) {
            const size_t k = idx[a];
            t[k] += dts[a];
            count[k]++;
            if (count[k] > 50) finished[k] = 1;
            if (act[a]->status != FG_OK) {   // quarantined by this substep
                finished[k] = 1;
                continue;
            }
            if (finished[k]) {
                act[a]->frame++;
                if (stats[k]) stats[k]->frame_ms = ms;
            }
        }
    }
    return FG_OK;
}

int ensure_lanes(fg_ctx* ctx, int n)
{
    while ((int)ctx->lanes.size() < n) {
        Lane* L = nullptr;
        FGCHK(lane_create(nullptr, &L));
        ctx->lanes.push_back(L);
    }
    return FG_OK;
}

// Split a batch into size classes, one lane each: scenes sorted by cell
// count, cut into contiguous runs of about equal work (cells^1.5 per scene,
// the PCG cost of a 2-D Poisson solve with this preconditioner).  The
// largest scenes get a lane of their own, so the smaller ones no longer wait
// on their solves' latency-bound tails.
std::vector<std::vector<size_t>> lane_plan(const std::vector<fg_scene*>& scenes, int lanes)
{
    const size_t n = scenes.size();
    std::vector<size_t> order(n);
    for (size_t k = 0; k < n; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return scenes[a]->C > scenes[b]->C; });
    double total = 0;
    std::vector<double> cost(n);
    for (size_t k = 0; k < n; k++) total += (cost[k] = std::pow((double)scenes[k]->C, 1.5));
    std::vector<std::vector<size_t>> plan(1);
    double acc = 0;
    for (size_t r = 0; r < n; r++) {
        const size_t k = order[r];
        const int lane = (int)plan.size() - 1;
        if (!plan.back().empty() && lane + 1 < lanes && acc >= total * (lane + 1) / lanes &&
            n - r >= (size_t)(lanes - 1 - lane))
            plan.emplace_back();
        plan.back().push_back(k);
        acc += cost[k];
    }
    return plan;
}

int quarantine_result(const std::vector<fg_scene*>& scenes);
int step_frames_split(fg_ctx* ctx, const std::vector<fg_scene*>& scenes, const std::vector<fg_stats*>& stats,
                      std::chrono::steady_clock::time_point t0);

// Every scene one frame: on lane 0 alone, or split over the context's lanes
// (one host thread each).  Results do not depend on the split.
int step_frames(fg_ctx* ctx, const std::vector<fg_scene*>& scenes, const std::vector<fg_stats*>& stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (fg_stats* st : stats) stats_reset(st);
    FGCHK(step_frames_split(ctx, scenes, stats, t0));
    return quarantine_result(scenes);
}

// FG_E_NUMERIC (with a message) when a scene of the call is quarantined.
int quarantine_result(const std::vector<fg_scene*>& scenes)
{
    int bad = 0;
    for (const fg_scene* s : scenes) bad += s->status != FG_OK;
    if (bad) return fail(FG_E_NUMERIC, std::to_string(bad) + " scene(s) quarantined (NaN/Inf); the others completed");
    return FG_OK;
}

int step_frames_split(fg_ctx* ctx, const std::vector<fg_scene*>& scenes, const std::vector<fg_stats*>& stats,
                      std::chrono::steady_clock::time_point t0)
{
    const int lanes = std::min<int>(ctx->nlanes, (int)scenes.size());
    if (lanes <= 1) return step_frames_lane(ctx, ctx->lanes[0], scenes, stats, t0);
    FGCHK(ensure_lanes(ctx, lanes));
    const auto plan = lane_plan(scenes, lanes);
    std::vector<int> rc(plan.size(), FG_OK);
    std::vector<std::string> msg(plan.size());
    auto run = [&](size_t li) {
        hipSetDevice(ctx->device);
        std::vector<fg_scene*> sc;
        std::vector<fg_stats*> st;
        for (size_t k : plan[li]) {
            sc.push_back(scenes[k]);
            st.push_back(stats[k]);
        }
        rc[li] = step_frames_lane(ctx, ctx->lanes[li], sc, st, t0);
        if (rc[li] != FG_OK) msg[li] = g_err;
    };
    std::vector<std::thread> workers;
    for (size_t li = 1; li < plan.size(); li++) workers.emplace_back(run, li);
    run(0);
    for (std::thread& w : workers) w.join();
    for (size_t li = 0; li < plan.size(); li++)
        if (rc[li] != FG_OK) return fail(rc[li], msg[li]);
    return FG_OK;
}


// The solver-only scene of an isolated solve (LinearSolver::solve replacement).
fg_scene_params solve_params(int32_t size_i, int32_t size_j, double rel_tol, int32_t max_iters)
{
    fg_scene_params p{};
    p.size_i = size_i;
    p.size_j = size_j;
    p.dx = 1.0;
    p.fluid_density = 1.0;
    p.fps = 1;
    p.solver_tolerance = rel_tol > 0 ? rel_tol : 1e-6;
    p.pcg_iter_limit = max_iters > 0 ? max_iters : 100000;
    return p;
}

// Upload the system into a solver-only scene, solve, download x.  The
// matrix comes as materials (assembled on the device, k_stencil) or as its
// bytes (one per cell, the stencil format of fg_scene_state).
int solve_on_scene(fg_scene* s, const int8_t* material, const uint8_t* stencil, double scale, const double* b,
                   double* x, fg_solve_stats* stats)
{
    fg_ctx* ctx = s->ctx;
    hipStream_t q = ctx->stream;
    const int64_t C = s->C;
    s->h.sA = scale;
    int rc = upload_desc(s);
    const auto t0 = std::chrono::steady_clock::now();
    if (rc == FG_OK) {
        hipError_t e = material ? hipMemcpyAsync(s->h.mat, material, C, hipMemcpyHostToDevice, q)
                                : hipMemcpyAsync(s->h.stencil, stencil, C, hipMemcpyHostToDevice, q);
        if (e == hipSuccess) e = hipMemcpyAsync(s->h.b, b, sizeof(double) * C, hipMemcpyHostToDevice, q);
        if (e != hipSuccess) rc = fail(FG_E_DEVICE, hipGetErrorString(e));
    }
    int32_t it = 0, sw = 0;
    if (rc == FG_OK) {
        Group G;
        rc = make_group(ctx, {s}, {}, G);
        if (rc == FG_OK) {
            if (material) launch_stencil(G.g.cells, q, G.d);
            rc = run_solve(G, false);
            it = s->hsc->total_iters;
            sw = s->hsc->sweeps;
            if (rc == FG_OK && s->hsc->numeric_bad) rc = fail(FG_E_NUMERIC, "NaN/Inf in the solve");
        }
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == FG_OK || rc == FG_E_NUMERIC) {
        hipError_t e = hipMemcpyAsync(x, s->h.x64, sizeof(double) * C, hipMemcpyDeviceToHost, q);
        if (e == hipSuccess) e = hipStreamSynchronize(q);
        if (e != hipSuccess && rc == FG_OK) rc = fail(FG_E_DEVICE, hipGetErrorString(e));
    }
    if (stats) {
        stats->iterations = it;
        stats->sweeps = sw;
        stats->rel_residual = s->hsc->bb > 0 ? std::sqrt(s->hsc->rr64 / s->hsc->bb) : 0.0;
        stats->converged = stats->rel_residual <= s->p.solver_tolerance;
        stats->solve_ms = ms;
    }
    return rc;
}

// Arguments of the batched entry points: the scenes (distinct, of this
// context, full scenes) and per-scene stats pointers (NULL when stats is).
int batch_args(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, fg_stats* stats, std::vector<fg_scene*>& v,
               std::vector<fg_stats*>& st)
{
    if (!ctx || n < 0 || (n > 0 && !scenes)) return fail(FG_E_INVALID, "NULL argument");
    v.assign(scenes, scenes + n);
    st.assign(n, nullptr);
    for (int32_t k = 0; k < n; k++) {
        fg_scene* s = scenes[k];
        if (!s) return fail(FG_E_INVALID, "scene is NULL");
        if (s->ctx != ctx) return fail(FG_E_INVALID, "scene belongs to another context");
        if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
        for (int32_t q = 0; q < k; q++)
            if (scenes[q] == s) return fail(FG_E_INVALID, "a scene appears twice in the batch");
        if (stats) st[k] = &stats[k];
    }
    return FG_OK;
}

// Polygons of a set-up packed for k_init_cells: vertex offsets, then the
// vertices as float2, in the set order sinks, sources, solids, fluids.
int upload_geometry(fg_scene* s, const fg_polygons* sets[4], const float* const visc[4], InitGeo* geo)
{
    std::vector<int32_t> voff(1, 0);
    std::vector<float2> verts;
    std::vector<float> pvisc;
    int first = 0;
    for (int q = 0; q < 4; q++) {
        const fg_polygons& P = *sets[q];
        if (P.n < 0 || (P.n > 0 && (!P.nverts || !P.verts))) return fail(FG_E_INVALID, "bad polygon set");
        geo->set[q] = PolySet{first, P.n};
        const float* v = P.verts;
        for (int32_t k = 0; k < P.n; k++) {
            pvisc.push_back(visc[q] ? visc[q][k] : 0.f);
            if (P.nverts[k] < 3) return fail(FG_E_INVALID, "a polygon needs 3 vertices");
            for (int32_t e = 0; e < P.nverts[k]; e++) verts.push_back(make_float2(v[2 * e], v[2 * e + 1]));
            v += 2 * P.nverts[k];
            voff.push_back((int32_t)verts.size());
        }
        first += P.n;
    }
    const size_t off_bytes = (sizeof(int32_t) * voff.size() + 255) & ~size_t(255);
    const size_t visc_bytes = (sizeof(float) * std::max<size_t>(pvisc.size(), 1) + 255) & ~size_t(255);
    const size_t bytes = off_bytes + visc_bytes + sizeof(float2) * std::max<size_t>(verts.size(), 1);
    Lane* L = s->ctx->lanes[0];
    if (bytes > L->bgeo_cap) {
        HIPCHK(hipStreamSynchronize(L->stream));
        if (L->bgeo) hipFree(L->bgeo);
        L->bgeo = nullptr;
        L->bgeo_cap = 0;
        HIPCHK(hipMalloc(&L->bgeo, bytes));
        L->bgeo_cap = bytes;
    }
    char* base = static_cast<char*>(L->bgeo);
    HIPCHK(hipMemcpyAsync(base, voff.data(), sizeof(int32_t) * voff.size(), hipMemcpyHostToDevice, L->stream));
    if (!pvisc.empty())
        HIPCHK(hipMemcpyAsync(base + off_bytes, pvisc.data(), sizeof(float) * pvisc.size(), hipMemcpyHostToDevice,
                              L->stream));
    if (!verts.empty())
        HIPCHK(hipMemcpyAsync(base + off_bytes + visc_bytes, verts.data(), sizeof(float2) * verts.size(),
                              hipMemcpyHostToDevice, L->stream));
    geo->voff = reinterpret_cast<const int32_t*>(base);
    geo->visc = reinterpret_cast<const float*>(base + off_bytes);
    geo->verts = reinterpret_cast<const float2*>(base + off_bytes + visc_bytes);
    geo->dx = (float)s->p.dx;
    // the host vectors are read by the copies above: wait for them
    HIPCHK(hipStreamSynchronize(L->stream));
    return FG_OK;
}

// The emitter table (fg_scene_set_emitters; with the sources' viscosity
// from fg_scene_init_objects).
int set_emitters(fg_scene* s, int32_t n, const float* velocity, const uint8_t* transfer, const float* viscosity,
                 uint32_t seed)
{
    if (!s || (n > 0 && !velocity)) return fail(FG_E_INVALID, "NULL argument");
    if (n < 0 || n > kMaxEmitters) return fail(FG_E_INVALID, "at most 256 emitters");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    // per emitter the face values afterTransfer writes: velocity / dx in
    // double, rounded to float (flipsolver2d.cpp:402-403)
    std::vector<float4> tab(kMaxEmitters, make_float4(0.f, 0.f, 0.f, 0.f));
    for (int32_t k = 0; k < n; k++)
        tab[k] = make_float4((float)((double)velocity[2 * k] / s->p.dx), (float)((double)velocity[2 * k + 1] / s->p.dx),
                             transfer && transfer[k] ? 1.f : 0.f, viscosity ? viscosity[k] : 0.f);
    HIPCHK(hipMemcpyAsync(s->h.emitters, tab.data(), sizeof(float4) * kMaxEmitters, hipMemcpyHostToDevice,
                          s->ctx->stream));
    s->h.n_emitters = n;
    s->h.rseed = seed;
    FGCHK(upload_desc(s));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return FG_OK;
}

}  // namespace

fg::LaunchEvents& fg::next_launch_events() { return g_next_events; }
int fg::report_error(int code, const std::string& msg) { return fail(code, msg); }

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

const char* fg_last_error(void) { return g_err.c_str(); }
int32_t fg_abi_version(void) { return FG_ABI_VERSION; }

int fg_ctx_create(int32_t device, uint64_t hbm_budget_bytes, fg_ctx** out)
{
    if (!out) return fail(FG_E_INVALID, "out is NULL");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(FG_E_INVALID, "no such device");
    HIPCHK(hipSetDevice(device));
    FGCHK(noalloc_resolve(device));
    auto* c = new fg_ctx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    Lane* L = nullptr;
    const int rc = lane_create(nullptr, &L);
    if (rc != FG_OK) {
        delete c;
        return rc;
    }
    c->lanes.push_back(L);
    c->stream = L->stream;
    *out = c;
    return FG_OK;
}

int fg_ctx_destroy(fg_ctx* ctx)
{
    if (!ctx) return FG_OK;
    for (Lane* L : ctx->lanes) lane_destroy(L);
    if (ctx->prof_ref) hipEventDestroy(ctx->prof_ref);
    delete ctx;
    return FG_OK;
}

int fg_ctx_bytes_in_use(fg_ctx* ctx, uint64_t* out)
{
    if (!ctx || !out) return fail(FG_E_INVALID, "NULL argument");
    *out = ctx->used;
    return FG_OK;
}

int fg_ctx_set_timing(fg_ctx* ctx, int32_t enabled)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    ctx->timing = enabled != 0;
    return FG_OK;
}

int fg_ctx_profile_kernel(fg_ctx* ctx, int32_t kernel)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    if (kernel < -1 || kernel >= FG_K_NUM) return fail(FG_E_INVALID, "bad kernel id");
    for (Lane* L : ctx->lanes) {
        HIPCHK(hipStreamSynchronize(L->stream));
        L->ev_used = 0;
        L->prof_units = 0;
    }
    if (!ctx->prof_ref) HIPCHK(hipEventCreate(&ctx->prof_ref));
    HIPCHK(hipEventRecord(ctx->prof_ref, ctx->lanes[0]->stream));
    HIPCHK(hipStreamSynchronize(ctx->lanes[0]->stream));
    ctx->prof_kernel = kernel;
    return FG_OK;
}

int fg_ctx_work_read(fg_ctx* ctx, int64_t* units, int32_t reset)
{
    if (!ctx || !units) return fail(FG_E_INVALID, "NULL argument");
    for (int k = 0; k < FG_K_NUM; k++) units[k] = 0;
    for (Lane* L : ctx->lanes) {
        for (int k = 0; k < FG_K_NUM; k++) {
            units[k] += L->work[k];
            if (reset) L->work[k] = 0;
        }
    }
    return FG_OK;
}

int fg_ctx_set_batch_lanes(fg_ctx* ctx, int32_t lanes)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    if (lanes < 1 || lanes > FG_MAX_LANES) return fail(FG_E_INVALID, "lanes must be in [1, FG_MAX_LANES]");
    FGCHK(ensure_lanes(ctx, lanes));
    ctx->nlanes = lanes;
    return FG_OK;
}

int fg_ctx_profile_read(fg_ctx* ctx, double* total_ms, double* busy_ms, int64_t* launches, int64_t* work_units)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    double tot = 0;
    int64_t n = 0, units = 0;
    std::vector<std::pair<double, double>> iv;   // launch intervals after prof_ref
    for (Lane* L : ctx->lanes) {
        HIPCHK(hipStreamSynchronize(L->stream));
        for (size_t k = 0; k + 1 < L->ev_used; k += 2) {
            float ms = 0, a = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&ms, L->ev_pool[k], L->ev_pool[k + 1]));
            tot += ms;
            if (ctx->prof_ref) {
                HIPCHK(hipEventElapsedTime(&a, ctx->prof_ref, L->ev_pool[k]));
                HIPCHK(hipEventElapsedTime(&b, ctx->prof_ref, L->ev_pool[k + 1]));
                iv.emplace_back(a, b);
            }
        }
        n += (int64_t)(L->ev_used / 2);
        units += L->prof_units;
        L->ev_used = 0;
        L->prof_units = 0;
    }
    // time during which at least one timed launch ran (launches of several
    // lanes overlap)
    std::sort(iv.begin(), iv.end());
    double busy = 0, cs = 0, ce = -1;
    for (const auto& p : iv) {
        if (p.first > ce) {
            if (ce > cs) busy += ce - cs;
            cs = p.first;
            ce = p.second;
        } else {
            ce = std::max(ce, p.second);
        }
    }
    if (ce > cs) busy += ce - cs;
    if (total_ms) *total_ms = tot;
    if (busy_ms) *busy_ms = busy;
    if (launches) *launches = n;
    if (work_units) *work_units = units;
    return FG_OK;
}

__global__ void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        b[k] = a[k];
}

int fg_ctx_bandwidth_probe(fg_ctx* ctx, uint64_t bytes, int32_t reps, double* gbps, double* ms_per_copy)
{
    if (!ctx || bytes < 16) return fail(FG_E_INVALID, "bad argument");
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) {
        hipFree(a);
        return fail(FG_E_DEVICE, hipGetErrorString(e));
    }
    const int64_t n = (int64_t)(bytes / 16);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipMemsetAsync(a, 0, bytes, ctx->stream);
    float best = 1e30f;
    for (int r = 0; r < std::max(1, reps); r++) {
        hipEventRecord(e0, ctx->stream);
        hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, ctx->stream, (const float4*)a, (float4*)b, n);
        hipEventRecord(e1, ctx->stream);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        best = std::min(best, ms);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (gbps) *gbps = 2.0 * (double)(n * 16) / (best * 1e-3) / 1e9;
    if (ms_per_copy) *ms_per_copy = best;
    return FG_OK;
}

int fg_scene_bytes(const fg_scene_params* p, uint64_t* out)
{
    FGCHK(validate_params(p));
    if (!out) return fail(FG_E_INVALID, "out is NULL");
    fg_scene tmp;
    tmp.p = *p;
    init_geometry_fields(&tmp);
    *out = arena_bytes(&tmp);
    return FG_OK;
}

int fg_scene_create(fg_ctx* ctx, const fg_scene_params* p, fg_scene** out)
{
    if (!ctx || !out) return fail(FG_E_INVALID, "NULL argument");
    HIPCHK(hipSetDevice(ctx->device));
    return create_scene(ctx, p, false, out);
}

int fg_scene_destroy(fg_scene* s)
{
    if (!s) return FG_OK;
    hipStreamSynchronize(s->ctx->stream);
    hipFree(s->arena);
    hipFree(s->d);
    hipHostFree(s->hpin);
    hipHostFree(s->hsc);
    s->ctx->used -= s->bytes;
    delete s;
    return FG_OK;
}

int fg_scene_set_friction(fg_scene* s, int32_t n, const float* friction)
{
    if (!s || (n > 0 && !friction)) return fail(FG_E_INVALID, "NULL argument");
    if (n < 0 || n > 256) return fail(FG_E_INVALID, "at most 256 obstacles");
    if (n > 0)
        HIPCHK(hipMemcpyAsync(s->h.friction, friction, sizeof(float) * n, hipMemcpyHostToDevice, s->ctx->stream));
    s->h.n_friction = n;
    s->h.friction_any = 0;
    for (int k = 0; k < n; k++)
        if (friction[k] != 0.f) s->h.friction_any = 1;
    FGCHK(upload_desc(s));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    return FG_OK;
}

int fg_scene_set_state(fg_scene* s, const fg_scene_state* st)
{
    if (!s || !st) return fail(FG_E_INVALID, "NULL argument");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    const int64_t n = st->n_particles;
    if (n < 0 || n > s->cap) return fail(FG_E_CAPACITY, "n_particles exceeds the scene's max_particles");
    s->last_stage = -1;
    if (n > 0 && (!st->px || !st->py || !st->vx || !st->vy))
        return fail(FG_E_INVALID, "particle arrays are required");
    if (!st->u || !st->v || !st->material || !st->solid_sdf)
        return fail(FG_E_INVALID, "u, v, material and solid_sdf are required");
    if (s->h.nb && !st->fluid_sdf)
        return fail(FG_E_INVALID, "an NB-FLIP scene needs fluid_sdf (its next substep advects the level set)");
    hipStream_t q = s->ctx->stream;
    const SceneDev& h = s->h;
    std::vector<float2> pos(n), vel(n);
    std::vector<uint32_t> pid(n);
    for (int64_t k = 0; k < n; k++) {
        pos[k] = make_float2(st->px[k], st->py[k]);
        vel[k] = make_float2(st->vx[k], st->vy[k]);
        pid[k] = st->pid ? st->pid[k] : (uint32_t)k;
    }
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(h.pos, pos.data(), sizeof(float2) * n, hipMemcpyHostToDevice, q));
        HIPCHK(hipMemcpyAsync(h.vel, vel.data(), sizeof(float2) * n, hipMemcpyHostToDevice, q));
        HIPCHK(hipMemcpyAsync(h.pid, pid.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, q));
        HIPCHK(hipMemsetAsync(h.key, 0, sizeof(uint32_t) * n, q));
        if (h.visc_on) {
            if (st->particle_viscosity)
                HIPCHK(hipMemcpyAsync(h.pvisc, st->particle_viscosity, sizeof(float) * n, hipMemcpyHostToDevice, q));
            else HIPCHK(hipMemsetAsync(h.pvisc, 0, sizeof(float) * n, q));
        }
    }
    HIPCHK(hipMemcpyAsync(h.uf, st->u, sizeof(float) * s->FU, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(h.vf, st->v, sizeof(float) * s->FV, hipMemcpyHostToDevice, q));
    if (st->u_valid) HIPCHK(hipMemcpyAsync(h.uvalF, st->u_valid, s->FU, hipMemcpyHostToDevice, q));
    else HIPCHK(hipMemsetAsync(h.uvalF, 0, s->FU, q));
    if (st->v_valid) HIPCHK(hipMemcpyAsync(h.vvalF, st->v_valid, s->FV, hipMemcpyHostToDevice, q));
    else HIPCHK(hipMemsetAsync(h.vvalF, 0, s->FV, q));
    HIPCHK(hipMemcpyAsync(h.mat, st->material, s->C, hipMemcpyHostToDevice, q));
    HIPCHK(hipMemcpyAsync(h.ssdf, st->solid_sdf, sizeof(float) * s->C, hipMemcpyHostToDevice, q));
    if (st->fluid_sdf) HIPCHK(hipMemcpyAsync(h.fsdf, st->fluid_sdf, sizeof(float) * s->C, hipMemcpyHostToDevice, q));
    if (st->particle_counts)
        HIPCHK(hipMemcpyAsync(h.pcount, st->particle_counts, sizeof(int32_t) * s->C, hipMemcpyHostToDevice, q));
    else HIPCHK(hipMemsetAsync(h.pcount, 0, sizeof(int32_t) * s->C, q));
    if (h.visc_on) {   // the viscosity grid (an NB-FLIP substep advects it)
        if (st->viscosity) HIPCHK(hipMemcpyAsync(h.cvisc, st->viscosity, sizeof(float) * s->C, hipMemcpyHostToDevice, q));
        else HIPCHK(hipMemsetAsync(h.cvisc, 0, sizeof(float) * s->C, q));
    }
    if (st->solid_id) HIPCHK(hipMemcpyAsync(h.sid, st->solid_id, sizeof(int32_t) * s->C, hipMemcpyHostToDevice, q));
    else HIPCHK(hipMemsetAsync(h.sid, 0xff, sizeof(int32_t) * s->C, q));
    if (st->emitter_id)
        HIPCHK(hipMemcpyAsync(h.eid, st->emitter_id, sizeof(int32_t) * s->C, hipMemcpyHostToDevice, q));
    else HIPCHK(hipMemsetAsync(h.eid, 0xff, sizeof(int32_t) * s->C, q));
    // SOURCE cells never change class, so whether the scene has any is fixed here
    int has_sources = 0;
    for (int64_t c = 0; c < s->C && !has_sources; c++) has_sources = st->material[c] == (int8_t)FG_SOURCE;
    uint32_t next_pid = st->next_pid;
    if (next_pid == 0)
        for (int64_t k = 0; k < n; k++) next_pid = std::max(next_pid, pid[k] + 1u);
    // the pressure iterate (the next solve's warm start) belongs to the old state
    HIPCHK(hipMemsetAsync(h.x64, 0, sizeof(double) * s->C, q));
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, q, s->d, (int)n);
    HIPCHK(hipStreamSynchronize(q));
    s->n_bound = n;
    s->n_live = n;
    s->vmax_ready = false;
    s->next_stage = 0;
    s->status = FG_OK;
    s->epoch = st->rng_epoch;
    s->next_pid = next_pid;
    s->h.has_sources = has_sources;
    FGCHK(upload_desc(s));
    HIPCHK(hipStreamSynchronize(q));
    return FG_OK;
}

int fg_scene_get_state(fg_scene* s, fg_scene_state* st)
{
    if (!s || !st) return fail(FG_E_INVALID, "NULL argument");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    hipStream_t q = s->ctx->stream;
    const SceneDev& h = s->h;
    FGCHK(read_scalars(s));
    const int64_t n = s->hsc->n_particles;
    std::vector<float2> pos(n), vel(n);
    std::vector<uint32_t> pid(n), key(n);
    std::vector<float> pv(h.visc_on ? n : 0);
    if (n > 0) {
        if (h.visc_on) HIPCHK(hipMemcpyAsync(pv.data(), h.pvisc, sizeof(float) * n, hipMemcpyDeviceToHost, q));
        HIPCHK(hipMemcpyAsync(pos.data(), h.pos, sizeof(float2) * n, hipMemcpyDeviceToHost, q));
        HIPCHK(hipMemcpyAsync(vel.data(), h.vel, sizeof(float2) * n, hipMemcpyDeviceToHost, q));
        HIPCHK(hipMemcpyAsync(pid.data(), h.pid, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, q));
        HIPCHK(hipMemcpyAsync(key.data(), h.key, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, q));
    }
    auto dl = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, q));
        return FG_OK;
    };
    FGCHK(dl(st->u, h.uf, sizeof(float) * s->FU));
    FGCHK(dl(st->v, h.vf, sizeof(float) * s->FV));
    FGCHK(dl(st->u_valid, h.uvalF, s->FU));
    FGCHK(dl(st->v_valid, h.vvalF, s->FV));
    FGCHK(dl(st->material, h.mat, s->C));
    FGCHK(dl(st->solid_sdf, h.ssdf, sizeof(float) * s->C));
    FGCHK(dl(st->fluid_sdf, h.fsdf, sizeof(float) * s->C));
    FGCHK(dl(st->solid_id, h.sid, sizeof(int32_t) * s->C));
    FGCHK(dl(st->particle_counts, h.pcount, sizeof(int32_t) * s->C));
    FGCHK(dl(st->saved_u, h.su, sizeof(float) * s->FU));
    FGCHK(dl(st->saved_v, h.sv, sizeof(float) * s->FV));
    FGCHK(dl(st->pressure, h.x64, sizeof(double) * s->C));
    FGCHK(dl(st->rhs, h.b, sizeof(double) * s->C));
    FGCHK(dl(st->stencil, h.stencil, s->C));
    FGCHK(dl(st->emitter_id, h.eid, sizeof(int32_t) * s->C));
    if (h.visc_on) FGCHK(dl(st->viscosity, h.cvisc, sizeof(float) * s->C));
    else if (st->viscosity) std::fill(st->viscosity, st->viscosity + s->C, 0.f);
    {   // the grid fluidVelocityGrid holds after the last stage hook
        const float *wu = h.u, *wv = h.v;   // P2G .. SDF, FORCES, PROJECT (+ SOLIDS)
        if (s->last_stage == FG_STAGE_EXTRAPOLATE) {   // extrapolated into the saved grid
            wu = h.su;
            wv = h.sv;
        } else if (s->last_stage < FG_STAGE_P2G || s->last_stage >= FG_STAGE_G2P) {   // the substep's grid
            wu = h.uf;
            wv = h.vf;
        }
        FGCHK(dl(st->work_u, wu, sizeof(float) * s->FU));
        FGCHK(dl(st->work_v, wv, sizeof(float) * s->FV));
    }
    st->rng_epoch = s->epoch;
    st->next_pid = s->next_pid;
    HIPCHK(hipStreamSynchronize(q));
    int64_t w = 0;
    for (int64_t k = 0; k < n; k++) {
        if (key[k] == kDead) continue;
        if (st->px) st->px[w] = pos[k].x;
        if (st->py) st->py[w] = pos[k].y;
        if (st->vx) st->vx[w] = vel[k].x;
        if (st->vy) st->vy[w] = vel[k].y;
        if (st->pid) st->pid[w] = pid[k];
        if (st->particle_viscosity) st->particle_viscosity[w] = h.visc_on ? pv[k] : 0.f;
        w++;
    }
    st->n_particles = w;
    return FG_OK;
}

int fg_scene_max_velocity(fg_scene* s, float* out)
{
    if (!s || !out) return fail(FG_E_INVALID, "NULL argument");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    FGCHK(ensure_vmax(s->ctx, s->ctx->lanes[0], {s}));
    float v2;
    std::memcpy(&v2, &s->hsc->max_v2_bits, 4);
    *out = std::sqrt(v2);
    return FG_OK;
}

int fg_scene_particle_count(fg_scene* s, int64_t* out)
{
    if (!s || !out) return fail(FG_E_INVALID, "NULL argument");
    *out = s->n_live;
    return FG_OK;
}

int fg_scene_set_emitters(fg_scene* s, int32_t n, const float* velocity, const uint8_t* transfer, uint32_t seed)
{
    return set_emitters(s, n, velocity, transfer, nullptr, seed);
}

int fg_scene_init_objects(fg_scene* s, const fg_scene_objects* o)
{
    if (!s || !o) return fail(FG_E_INVALID, "NULL argument");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    if (o->solids.n > 256) return fail(FG_E_INVALID, "at most 256 obstacles");
    if (o->sources.n > kMaxEmitters) return fail(FG_E_INVALID, "at most 256 sources");
    s->last_stage = -1;
    const fg_polygons* sets[4] = {&o->sinks, &o->sources, &o->solids, &o->fluids};
    const float* const visc[4] = {nullptr, o->source_viscosity, nullptr, o->fluid_viscosity};
    InitGeo geo{};
    FGCHK(upload_geometry(s, sets, visc, &geo));
    hipStream_t q = s->ctx->stream;
    SceneDev& h = s->h;
    // grids of a scene at rest; the cell pass writes materials, ids, SDFs
    HIPCHK(hipMemsetAsync(h.uf, 0, sizeof(float) * s->FU, q));
    HIPCHK(hipMemsetAsync(h.vf, 0, sizeof(float) * s->FV, q));
    HIPCHK(hipMemsetAsync(h.uvalF, 0, s->FU, q));
    HIPCHK(hipMemsetAsync(h.vvalF, 0, s->FV, q));
    const dim3 cells((unsigned)cdiv(s->C, kBlock));
    const dim3 scan((unsigned)cdiv(s->C, kScanTileCells));
    launch_init_cells(cells, q, s->d, geo);
    launch_scan_local(scan, q, s->d);
    launch_scan_top(dim3(1), q, s->d);
    launch_scan_add(scan, q, s->d);
    int32_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, h.cellStart + s->C, sizeof(int32_t), hipMemcpyDeviceToHost, q));
    HIPCHK(hipStreamSynchronize(q));
    const bool fits = total <= s->cap;
    if (fits) launch_seed(cells, q, s->d, o->seed);
    else launch_clear_counts(cells, q, s->d);
    const int32_t n = fits ? total : 0;
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, q, s->d, (int)n);
    HIPCHK(hipGetLastError());
    s->n_bound = n;
    s->n_live = n;
    s->vmax_ready = false;
    s->next_stage = 0;
    s->status = FG_OK;
    s->epoch = 0;
    s->next_pid = (uint32_t)n;
    s->h.has_sources = o->sources.n > 0;   // SOURCE cells can only come from sources
    if (o->friction && o->solids.n > 0) FGCHK(fg_scene_set_friction(s, o->solids.n, o->friction));
    else FGCHK(fg_scene_set_friction(s, 0, nullptr));
    std::vector<float> vel(2 * std::max(o->sources.n, 0), 0.f);
    if (o->source_velocity) std::copy(o->source_velocity, o->source_velocity + vel.size(), vel.begin());
    FGCHK(set_emitters(s, o->sources.n, vel.data(), o->source_transfer, o->source_viscosity, o->seed));
    if (!fits) return fail(FG_E_CAPACITY, "seeded particles exceed max_particles");
    return FG_OK;
}

int fg_scene_init_geometry(fg_scene* s, int32_t n_solids, const int32_t* snv, const float* sv,
                           const float* friction, int32_t n_fluids, const int32_t* fnv, const float* fv,
                           uint32_t seed)
{
    if (!s || (n_solids > 0 && (!snv || !sv)) || (n_fluids > 0 && (!fnv || !fv)))
        return fail(FG_E_INVALID, "NULL argument");
    fg_scene_objects o{};
    o.solids = fg_polygons{n_solids, snv, sv};
    o.friction = friction;
    o.fluids = fg_polygons{n_fluids, fnv, fv};
    o.seed = seed;
    return fg_scene_init_objects(s, &o);
}

int fg_scene_substep(fg_scene* s, float dt, fg_stats* stats)
{
    if (!s) return fail(FG_E_INVALID, "scene is NULL");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FG_E_INVALID, "dt must be positive and finite");
    stats_reset(stats);
    if (s->status != FG_OK) {
        if (stats) stats->status = s->status;
        return fail(s->status, "scene is quarantined (set_state gives it a new state)");
    }
    FGCHK(substep_group(s->ctx, s->ctx->lanes[0], {s}, {dt}, {stats}));
    return quarantine_result({s});
}

int fg_scene_step_frame(fg_scene* s, fg_stats* stats)
{
    if (!s) return fail(FG_E_INVALID, "scene is NULL");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    return step_frames(s->ctx, {s}, {stats});
}

int fg_scene_stage(fg_scene* s, int32_t stage, float dt)
{
    if (!s) return fail(FG_E_INVALID, "scene is NULL");
    if (s->solver_only) return fail(FG_E_STATE, "solver-only scene");
    if (stage < 0 || stage >= FG_STAGE_NUM) return fail(FG_E_INVALID, "bad stage");
    if (s->h.nb) return fail(FG_E_INVALID, "stage hooks follow FlipSolver::step; an NB-FLIP scene steps whole");
    if (s->status != FG_OK) return fail(s->status, "scene is quarantined (set_state gives it a new state)");
    if (stage == FG_STAGE_ADVECT) {
        if (!(dt > 0.f)) return fail(FG_E_INVALID, "dt must be positive");
        set_substep_scalars(s, dt);
    }
    fg_stats tmp;
    stats_reset(&tmp);
    Group G;
    FGCHK(make_group(s->ctx, {s}, {&tmp}, G));
    if (stage == FG_STAGE_ADVECT) hipLaunchKernelGGL(k_substep_begin, G.g.one, dim3(64), 0, s->ctx->stream, G.d);
    FGCHK(run_stage(G, stage));
    HIPCHK(hipGetLastError());
    FGCHK(read_scalars(s));
    if (stage == FG_STAGE_PROJECT && s->hsc->numeric_bad) {
        quarantine(s, nullptr);
        return fail(FG_E_NUMERIC, "NaN/Inf in the pressure solve: scene quarantined");
    }
    if (stage == FG_STAGE_G2P || stage == FG_STAGE_COUNT) {
        s->n_live = s->hsc->n_particles - s->hsc->erased;
        s->n_bound = s->hsc->n_particles;
        s->vmax_ready = true;
    }
    if (stage == FG_STAGE_COUNT) {   // the substep is complete
        s->epoch++;
        s->next_pid += (uint32_t)s->hsc->reseeded;
    }
    s->last_stage = stage == FG_STAGE_COUNT ? -1 : stage;
    return FG_OK;
}

int fg_batch_step_frame(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, fg_stats* stats)
{
    std::vector<fg_scene*> v;
    std::vector<fg_stats*> st;
    FGCHK(batch_args(ctx, scenes, n, stats, v, st));
    return step_frames(ctx, v, st);
}

int fg_batch_substep(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, const float* dts, fg_stats* stats)
{
    std::vector<fg_scene*> v;
    std::vector<fg_stats*> st;
    FGCHK(batch_args(ctx, scenes, n, stats, v, st));
    if (n > 0 && !dts) return fail(FG_E_INVALID, "dts is NULL");
    for (int32_t k = 0; k < n; k++)
        if (!(dts[k] > 0.f) || !std::isfinite(dts[k])) return fail(FG_E_INVALID, "dt must be positive and finite");
    std::vector<fg_scene*> act;
    std::vector<fg_stats*> ast;
    std::vector<float> adt;
    for (int32_t k = 0; k < n; k++) {
        stats_reset(st[k]);
        if (v[k]->status != FG_OK) {   // quarantined earlier: skipped
            if (st[k]) {
                st[k]->status = v[k]->status;
                st[k]->converged = 0;
            }
            continue;
        }
        act.push_back(v[k]);
        ast.push_back(st[k]);
        adt.push_back(dts[k]);
    }
    if (!act.empty()) FGCHK(substep_group(ctx, ctx->lanes[0], act, adt, ast));
    return quarantine_result(v);
}

int fg_ctx_set_solver_tiling(fg_ctx* ctx, int32_t target_blocks, int32_t min_chunk)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    if (target_blocks < 1 || target_blocks > kPcgTilingMaxBlocks) return fail(FG_E_INVALID, "bad target_blocks");
    if (min_chunk < kPcgTilingMinChunk || min_chunk % kBlock) return fail(FG_E_INVALID, "bad min_chunk");
    ctx->pcg_target_blocks = target_blocks;
    ctx->pcg_min_chunk = min_chunk;
    return FG_OK;
}

int fg_ctx_set_preconditioner(fg_ctx* ctx, int32_t kind)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    if (kind != FG_PRECOND_MG && kind != FG_PRECOND_IPP) return fail(FG_E_INVALID, "bad preconditioner");
    ctx->precond = kind == FG_PRECOND_MG ? kPrecondMG : kPrecondIPP;
    return FG_OK;
}

int fg_ctx_set_vcycle_fused(fg_ctx* ctx, int32_t enabled)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    ctx->vcycle_fused = enabled != 0;
    return FG_OK;
}

int fg_ctx_set_solver_graphs(fg_ctx* ctx, int32_t enabled)
{
    if (!ctx) return fail(FG_E_INVALID, "ctx is NULL");
    if (enabled) return fail(FG_E_INVALID, "graph replay of PCG iterations is switched off");
    return FG_OK;
}

int fg_ctx_preconditioner(fg_ctx* ctx, int32_t* kind)
{
    if (!ctx || !kind) return fail(FG_E_INVALID, "NULL argument");
    *kind = ctx->precond == kPrecondMG ? FG_PRECOND_MG : FG_PRECOND_IPP;
    return FG_OK;
}

int fg_ctx_batch_lanes(fg_ctx* ctx, int32_t* lanes)
{
    if (!ctx || !lanes) return fail(FG_E_INVALID, "NULL argument");
    *lanes = ctx->nlanes;
    return FG_OK;
}

int fg_batch_autotune(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, int32_t rounds, int32_t* lanes_out,
                      double* rates_out)
{
    if (!ctx || !scenes || n < 1 || rounds < 1) return fail(FG_E_INVALID, "bad argument");
    std::vector<double> ms(FG_MAX_LANES + 1, 0.0), work(FG_MAX_LANES + 1, 0.0);
    std::vector<fg_stats> st(n);
    const int saved = ctx->nlanes;
    for (int r = 0; r < rounds; r++)
        for (int q = 0; q < FG_MAX_LANES; q++) {
            const int lanes = (r & 1) ? FG_MAX_LANES - q : q + 1;   // alternate the order
            int rc = fg_ctx_set_batch_lanes(ctx, lanes);
            const auto t0 = std::chrono::steady_clock::now();
            if (rc == FG_OK) rc = fg_batch_step_frame(ctx, scenes, n, st.data());
            if (rc != FG_OK) {
                ctx->nlanes = saved;
                return rc;
            }
            ms[lanes] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            for (const fg_stats& s : st) work[lanes] += (double)s.particle_substeps;
        }
    // each count is measured on other frames of the advancing simulation,
    // so rates within kTie of the fastest count as equal, and the most lanes
    // among them win (more lanes keep more independent launches in flight)
    constexpr double kTie = 0.97;
    std::vector<double> rate(FG_MAX_LANES + 1, 0.0);
    double top = 0.0;
    for (int l = 1; l <= FG_MAX_LANES; l++) {
        rate[l] = ms[l] > 0 ? work[l] / (ms[l] * 1e-3) : 0.0;
        if (rates_out) rates_out[l - 1] = rate[l];
        top = std::max(top, rate[l]);
    }
    int best = 1;
    for (int l = 1; l <= FG_MAX_LANES; l++)
        if (rate[l] >= kTie * top) best = l;
    ctx->nlanes = best;
    if (lanes_out) *lanes_out = best;
    return FG_OK;
}

int fg_batch_autotune_tiling(fg_ctx* ctx, fg_scene* const* scenes, int32_t n, int32_t rounds,
                             const int32_t* candidates, int32_t n_cand, int32_t* best_out, double* rates_out)
{
    static const int32_t kDefault[2 * FG_TILING_CANDIDATES] = {1024, 256, 1024, 512, 1024, 1024, 1024, 2048,
                                                               1024, 4096};
    if (!candidates) {
        candidates = kDefault;
        n_cand = FG_TILING_CANDIDATES;
    }
    if (!ctx || !scenes || n < 1 || rounds < 1 || n_cand < 1) return fail(FG_E_INVALID, "bad argument");
    for (int k = 0; k < n_cand; k++) {
        const int tb = candidates[2 * k], mc = candidates[2 * k + 1];
        if (tb < 1 || tb > kPcgTilingMaxBlocks || mc < kPcgTilingMinChunk || mc % kBlock)
            return fail(FG_E_INVALID, "bad tiling candidate");
    }
    const int saved_tb = ctx->pcg_target_blocks, saved_mc = ctx->pcg_min_chunk;
    std::vector<double> ms(n_cand, 0.0), work(n_cand, 0.0);
    std::vector<fg_stats> st(n);
    for (int r = 0; r < rounds; r++)
        for (int q = 0; q < n_cand; q++) {
            const int k = (r & 1) ? n_cand - 1 - q : q;   // alternate the order
            ctx->pcg_target_blocks = candidates[2 * k];
            ctx->pcg_min_chunk = candidates[2 * k + 1];
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = fg_batch_step_frame(ctx, scenes, n, st.data());
            if (rc != FG_OK) {
                ctx->pcg_target_blocks = saved_tb;
                ctx->pcg_min_chunk = saved_mc;
                return rc;
            }
            ms[k] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            for (const fg_stats& s : st) work[k] += (double)s.particle_substeps;
        }
    int best = 0;
    for (int k = 0; k < n_cand; k++) {
        const double rate = ms[k] > 0 ? work[k] / (ms[k] * 1e-3) : 0.0;
        if (rates_out) rates_out[k] = rate;
        const double rb = ms[best] > 0 ? work[best] / (ms[best] * 1e-3) : 0.0;
        if (rate > rb) best = k;
    }
    ctx->pcg_target_blocks = candidates[2 * best];
    ctx->pcg_min_chunk = candidates[2 * best + 1];
    if (best_out) *best_out = best;
    return FG_OK;
}

int fg_poisson_solve(fg_ctx* ctx, int32_t size_i, int32_t size_j, const int8_t* material, double scale,
                     const double* b, double* x, double rel_tol, int32_t max_iters, fg_solve_stats* stats)
{
    if (!ctx || !material || !b || !x) return fail(FG_E_INVALID, "NULL argument");
    if (!(scale > 0)) return fail(FG_E_INVALID, "scale must be > 0");
    const fg_scene_params p = solve_params(size_i, size_j, rel_tol, max_iters);
    fg_scene* s = nullptr;
    FGCHK(create_scene(ctx, &p, true, &s));
    const int rc = solve_on_scene(s, material, nullptr, scale, b, x, stats);
    fg_scene_destroy(s);
    return rc;
}

int fg_poisson_solve_matrix(fg_ctx* ctx, int32_t size_i, int32_t size_j, const uint8_t* matrix, double scale,
                            const double* b, double* x, double rel_tol, int32_t max_iters, fg_solve_stats* stats)
{
    if (!ctx || !matrix || !b || !x) return fail(FG_E_INVALID, "NULL argument");
    if (!(scale > 0)) return fail(FG_E_INVALID, "scale must be > 0");
    const fg_scene_params p = solve_params(size_i, size_j, rel_tol, max_iters);
    fg_scene* s = nullptr;
    FGCHK(create_scene(ctx, &p, true, &s));
    const int rc = solve_on_scene(s, nullptr, matrix, scale, b, x, stats);
    fg_scene_destroy(s);
    return rc;
}

int fg_poisson_solve_noalloc(fg_ctx* ctx, int32_t size_i, int32_t size_j, const int8_t* material, double scale,
                             const double* b, double* x, double rel_tol, int32_t max_iters, fg_solve_stats* stats)
{
    if (!ctx || !material || !b || !x) return fail(FG_E_INVALID, "NULL argument");
    if (!(scale > 0)) return fail(FG_E_INVALID, "scale must be > 0");
    const fg_scene_params p = solve_params(size_i, size_j, rel_tol, max_iters);
    FGCHK(validate_params(&p));
    if (ctx->device < 0 || ctx->device >= kNoAllocDevices) return fail(FG_E_INVALID, "device index out of range");
    NoAllocSlot& slot = g_noalloc[ctx->device];
    std::lock_guard<std::mutex> lock(slot.mu);
    if (!slot.dev) return fail(FG_E_STATE, "scratch address not resolved for this device");
    HIPCHK(hipSetDevice(ctx->device));
    fg_scene& sc = slot.scene;
    sc = fg_scene{};
    sc.ctx = ctx;
    sc.p = p;
    sc.solver_only = true;
    init_geometry_fields(&sc);
    sc.bytes = arena_bytes(&sc);
    const uint64_t desc_off = (sc.bytes + 255) & ~uint64_t(255);
    if (desc_off + sizeof(SceneDev) > kNoAllocScratchBytes)
        return fail(FG_E_CAPACITY, "the grid needs " + std::to_string(desc_off + sizeof(SceneDev)) +
                                       " bytes of device scratch; the allocation-free solve has " +
                                       std::to_string(kNoAllocScratchBytes) + " (use fg_poisson_solve)");
    sc.arena = slot.dev;
    sc.d = reinterpret_cast<SceneDev*>(static_cast<char*>(slot.dev) + desc_off);
    sc.hpin = &slot.desc_stage;
    sc.hsc = &slot.scalars;
    FGCHK(setup_scene_memory(&sc));
    return solve_on_scene(&sc, material, nullptr, scale, b, x, stats);
}

}  // extern "C"
