// fg_particles.hip -- particle-space kernels of the substep.
//
// Particles live in one SoA arena per scene (float2 position, float2
// velocity, u32 stable id), kept sorted by (cell, id) after every substep.
// The reference's 3x3-cell bins and rebin sets
// (FlipSolver2dLib/markerparticlesystem.cpp:47-58, 260-269, 317-342) become a
// counting sort on the cell key with a per-cell id fixup: deterministic (no
// float atomics anywhere) and O(particles + cells).
//
// Counters (kills, erasures, max |v|^2) are per-block partials reduced by a
// one-block kernel: no two blocks ever update the same address.  The only
// atomics are the sort's per-cell counters, aggregated over runs of equal
// keys within a wave (particles arrive sorted by their old cell, so most
// runs are long).
#include "fg_kernels.h"

namespace fg {

// ---------------------------------------------------------------------------
// wave64 runs of equal keys: head lane of this lane's run and the run length
// ---------------------------------------------------------------------------
struct Run {
    int head, len;
};
__device__ __forceinline__ Run key_run(uint32_t key)
{
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool is_head = lane == 0 || prev != key;
    const unsigned long long heads = __ballot(is_head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);   // bits <= lane
    const int head = 63 - __clzll(heads & upto);
    const unsigned long long above = heads & ~upto;
    const int end = above ? (__ffsll((long long)above) - 1) : 64;
    return Run{head, end - head};
}

// ---------------------------------------------------------------------------
// max |v|^2 over live particles (FlipSolver::maxParticleVelocity,
// flipsolver2d.cpp:1560-1574): per-block maxima, reduced by k_finalize.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_max_v2(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || (int)blockIdx.x >= s.pblocks) return;
    __shared__ float red[kBlock / 64];
    const int n = s.sc->n_particles;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    float m = 0.f;
    if (k < n && s.key[k] != kDead) {
        const float2 v = s.vel[k];
        m = v.x * v.x + v.y * v.y;
    }
    m = block_maxf(m, red);
    if (threadIdx.x == 0) {
        s.fpart[blockIdx.x] = m;
        s.ipart[s.max_pblocks + blockIdx.x] = 0;
    }
}

// ---------------------------------------------------------------------------
// Advection: FlipSolver::advectThread (flipsolver2d.cpp:305-338) with
// rk4Integrate (1195-1203, fg_device.h) and SdfGrid::closestSurfacePoint
// (sdfgrid.cpp:15-46).  Fused with the cell histogram of the counting sort.
// ---------------------------------------------------------------------------
// simmath::gradCenteredGrid with truncating (ssize_t) arguments
__device__ __forceinline__ float2 grad_trunc(const float* g, int I, int J, float x, float y)
{
    const int i = (int)x, j = (int)y;
    const float ip1 = get_ext(g, I, J, i + 1, j), im1 = get_ext(g, I, J, i - 1, j);
    const float jp1 = get_ext(g, I, J, i, j + 1), jm1 = get_ext(g, I, J, i, j - 1);
    return make_float2((ip1 - im1) / 2.f, (jp1 - jm1) / 2.f);
}

// The 3x3 block of clamped SDF values (get_ext) around a centre cell.  The
// search below moves by fractions of a cell per step, so most of its lookups
// hit the block and need no memory access; a lookup outside reloads the
// block around its own cell (nine independent loads).  Values and
// arithmetic are those of lerp_grid / grad_trunc, so results are unchanged.
struct SdfBlock {
    int ci, cj;
    float v[3][3];
};

__device__ __forceinline__ void sdf_block_load(SdfBlock& B, const float* g, int I, int J, int ci, int cj)
{
    B.ci = ci;
    B.cj = cj;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) B.v[a][b] = get_ext(g, I, J, ci - 1 + a, cj - 1 + b);
}

// value at logical cell (i, j), |i - ci| <= 1 and |j - cj| <= 1
__device__ __forceinline__ float sdf_block_at(const SdfBlock& B, int i, int j)
{
    const int a = i - B.ci, b = j - B.cj;
    float r[3];
#pragma unroll
    for (int q = 0; q < 3; q++) r[q] = b < 0 ? B.v[q][0] : (b == 0 ? B.v[q][1] : B.v[q][2]);
    return a < 0 ? r[0] : (a == 0 ? r[1] : r[2]);
}

__device__ __forceinline__ bool sdf_block_has(const SdfBlock& B, int i, int j)
{
    return i >= B.ci - 1 && i <= B.ci + 1 && j >= B.cj - 1 && j <= B.cj + 1;
}

// lerp_grid (grid2d.h:186-216) of the solid SDF through the block
__device__ __forceinline__ float sdf_lerp_cached(SdfBlock& B, const float* g, int I, int J, float i, float j)
{
    i = fclamp(i, 0.f, (float)(I - 1));
    j = fclamp(j, 0.f, (float)(J - 1));
    const int ci = integr(i), cj = integr(j);
    const float fi = frac_(i), fj = frac_(j);
    const int nj = fj >= 0.5f ? cj + 1 : cj - 1;
    const int ni = fi >= 0.5f ? ci + 1 : ci - 1;
    const float iF = fi < 0.5f ? 0.5f - fi : fi - 0.5f;
    const float jF = fj < 0.5f ? 0.5f - fj : fj - 0.5f;
    if (!sdf_block_has(B, ci, cj) || !sdf_block_has(B, ni, nj)) sdf_block_load(B, g, I, J, ci, cj);
    const float v1 = lerp1(sdf_block_at(B, ci, cj), sdf_block_at(B, ni, cj), iF);
    const float v2 = lerp1(sdf_block_at(B, ci, nj), sdf_block_at(B, ni, nj), iF);
    return lerp1(v1, v2, jF);
}

// grad_trunc through the block
__device__ __forceinline__ float2 sdf_grad_cached(SdfBlock& B, const float* g, int I, int J, float x, float y)
{
    const int i = (int)x, j = (int)y;
    if (i != B.ci || j != B.cj) sdf_block_load(B, g, I, J, i, j);
    const float ip1 = sdf_block_at(B, i + 1, j), im1 = sdf_block_at(B, i - 1, j);
    const float jp1 = sdf_block_at(B, i, j + 1), jm1 = sdf_block_at(B, i, j - 1);
    return make_float2((ip1 - im1) / 2.f, (jp1 - jm1) / 2.f);
}

__device__ float2 closest_surface_point(const float* sdf, int I, int J, float2 pos)
{
    SdfBlock B;
    sdf_block_load(B, sdf, I, J, (int)pos.x, (int)pos.y);
    float2 cp = pos;
    float value = sdf_lerp_cached(B, sdf, I, J, pos.x, pos.y);
    float2 grad = sdf_grad_cached(B, sdf, I, J, pos.x, pos.y);
    for (int it = 0; it < 100; it++) {
        float alpha = 1;
        bool moved = false;
        for (int k = 0; k < 10; k++) {
            const float t = alpha * value;
            const float2 q = make_float2(cp.x - grad.x * t, cp.y - grad.y * t);
            const float vq = sdf_lerp_cached(B, sdf, I, J, q.x, q.y);
            if (fabsf(vq) < fabsf(value)) {
                cp = q;
                value = vq;
                grad = sdf_grad_cached(B, sdf, I, J, q.x, q.y);
                moved = true;
                if (fabsf(value) < 1e-5f) return cp;
            } else {
                alpha *= 0.7f;
            }
        }
        // an outer iteration that accepted no step leaves (cp, value, grad)
        // as they were and alpha restarts at 1, so every later one repeats
        // it exactly: the reference's remaining iterations change nothing
        if (!moved) break;
    }
    return cp;
}

// NBFlipSolver::pruneNarrowBand (nbflipsolver.cpp:228-252) for a particle
// advect left alive in cell c (the position's truncation equals its floor
// inside the grid): below the narrow band it dies unless its cell is a
// SOURCE; in a SOURCE cell or below the resampling band it dies when its cell
// held more than 2*ppc particles at the last count.  The level set, the
// materials and the counts are the previous substep's.
__device__ __forceinline__ bool nb_pruned(const SceneDev& s, int c, float2 p)
{
    const float sdf = lerp_grid(s.fsdf, s.I, s.J, 0.f, 0.f, p.x, p.y);
    const bool src = s.mat[c] == M_SOURCE;
    if (!src && sdf < kNbNarrowBand) return true;
    return (src || sdf < kNbResamplingBand) && s.pcount[c] > 2 * s.ppc;
}

__global__ void __launch_bounds__(kBlock) k_advect(const SceneDev* scenes, int alive_from_key)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || (int)blockIdx.x >= s.pblocks) return;
    __shared__ int red[kBlock / 64];
    const int n = s.sc->n_particles;
    const int I = s.I, J = s.J;
    int killed = 0;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    uint32_t key = kDead;
    if (k < n) {
        const bool was_alive = !alive_from_key || s.key[k] != kDead;
        float2 p = s.pos[k];
        const float2 vel = s.vel[k];
        if (was_alive) {
            p = rk4(s.uf, s.vf, I, J, p, s.dt);
            if (lerp_grid(s.ssdf, I, J, 0.f, 0.f, p.x, p.y) < 0.f) p = closest_surface_point(s.ssdf, I, J, p);
            // a non-finite position (from a NaN velocity field) has no cell:
            // it is killed before the conversion to a cell index
            const bool fin = isfinite(p.x) && isfinite(p.y);
            const int pi = fin ? integr(p.x) : -1, pj = fin ? integr(p.y) : -1;
            const bool inb = pi >= 0 && pi < I && pj >= 0 && pj < J;
            if (!inb || is_sink(s.mat[(size_t)pi * J + pj])) {
                killed = 1;
            } else if (s.nb && nb_pruned(s, pi * J + pj, p)) {
                killed = 1;
            } else {
                key = (uint32_t)(pi * J + pj);
            }
        }
        s.posA[k] = p;
        s.velA[k] = vel;
        if (s.visc_on) s.pviscA[k] = s.pvisc[k];
        s.key[k] = key;
    }
    // histogram: one atomic per run of equal keys in the wave
    const Run run = key_run(key);
    if ((int)(threadIdx.x & 63) == run.head && key != kDead) atomicAdd(&s.cellCount[key], run.len);
    killed = block_sumi(killed, red);
    if (threadIdx.x == 0) s.ipart[blockIdx.x] = killed;
}

// ---------------------------------------------------------------------------
// Counting sort on cell key.
//   scan: exclusive prefix sum of cellCount -> cellStart (3 passes)
//   scatter: slots from cellStart[key] and a per-run atomic on cellCount
//            (leaves cellCount at 0 for the next substep)
//   fixup: per cell, order the (id, source) pairs by id
//   gather: move the payload into slot order (coalesced writes)
// ---------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;   // 4096 cells per block

__device__ __forceinline__ int block_excl_scan(int v, int* lds, int* total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    int woff = 0, tot = 0;
    for (int k = 0; k < kBlock / 64; k++) {
        if (k < w) woff += lds[k];
        tot += lds[k];
    }
    __syncthreads();
    *total = tot;
    return woff + inc - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_local(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int C = s.I * s.J;
    const int base = blockIdx.x * kScanTile;
    if (base >= C) return;
    __shared__ int lds[kBlock / 64];
    int vals[kScanItems];
    int sum = 0;
    const int t0 = base + threadIdx.x * kScanItems;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        const int c = t0 + q;
        vals[q] = c < C ? s.cellCount[c] : 0;
        sum += vals[q];
    }
    int total;
    int run = block_excl_scan(sum, lds, &total);
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        const int c = t0 + q;
        if (c < C) s.cellStart[c] = run;
        run += vals[q];
    }
    if (threadIdx.x == 0) s.scanSums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_top(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int C = s.I * s.J;
    const int nb = (C + kScanTile - 1) / kScanTile;
    __shared__ int lds[kBlock / 64];
    int* sums = s.scanSums;
    int carry = 0;
    for (int base = 0; base < nb; base += kBlock) {
        const int b = base + threadIdx.x;
        const int v = b < nb ? sums[b] : 0;
        int total;
        const int ex = block_excl_scan(v, lds, &total);
        if (b < nb) sums[b] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) s.cellStart[C] = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_add(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int C = s.I * s.J;
    const int base = blockIdx.x * kScanTile;
    if (base >= C) return;
    const int off = s.scanSums[blockIdx.x];
    if (off == 0) return;
    for (int c = base + threadIdx.x; c < base + kScanTile && c < C; c += kBlock) s.cellStart[c] += off;
}

// Runs over the advected (pre-sort) arrays: sc->n_particles is still the
// pre-sort count until k_finalize commits the new one.
__global__ void __launch_bounds__(kBlock) k_scatter(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const int n = s.sc->n_particles;
    const uint32_t key = k < n ? s.key[k] : kDead;
    const Run run = key_run(key);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == run.head && key != kDead) base = atomicSub(&s.cellCount[key], run.len) - run.len;
    base = __shfl(base, run.head, 64);
    if (key == kDead) return;
    const int slot = s.cellStart[key] + base + (lane - run.head);
    s.tmpPid[slot] = s.pid[k];
    s.tmpSrc[slot] = (uint32_t)k;
}

// Per cell: order the slots by (stable id, source slot).  The source slot
// is unique, so the order is total even if a caller repeats an id, and the
// padding pair (0xffffffff, 0xffffffff) sorts after every real entry (a
// source slot is below the particle capacity).
__device__ __forceinline__ bool pair_gt(uint32_t p0, uint32_t s0, uint32_t p1, uint32_t s1)
{
    return p0 > p1 || (p0 == p1 && s0 > s1);
}

// A cell of 2..N entries, sorted in registers by a bitonic network.
template <int N>
__device__ __forceinline__ void sort_cell_regs(uint32_t* tp, uint32_t* ts, int a, int n)
{
    uint32_t kp[N], ks[N];
#pragma unroll
    for (int q = 0; q < N; q++) {
        kp[q] = q < n ? tp[a + q] : 0xffffffffu;
        ks[q] = q < n ? ts[a + q] : 0xffffffffu;
    }
    bool sorted = true;
#pragma unroll
    for (int q = 1; q < N; q++)
        if (q < n && pair_gt(kp[q - 1], ks[q - 1], kp[q], ks[q])) sorted = false;
    if (sorted) return;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int l = i ^ j;
                if (l <= i) continue;
                const bool gt = pair_gt(kp[i], ks[i], kp[l], ks[l]);
                const bool sw = (i & k) == 0 ? gt : !gt;
                const uint32_t p0 = sw ? kp[l] : kp[i], p1 = sw ? kp[i] : kp[l];
                const uint32_t s0 = sw ? ks[l] : ks[i], s1 = sw ? ks[i] : ks[l];
                kp[i] = p0;
                kp[l] = p1;
                ks[i] = s0;
                ks[l] = s1;
            }
#pragma unroll
    for (int q = 0; q < N; q++)
        if (q < n) {
            tp[a + q] = kp[q];
            ts[a + q] = ks[q];
        }
}

// Cells of up to 32 entries: one thread each, in registers.  Longer cells
// (pile-ups): the whole wave ranks the cell's entries through LDS, one cell
// at a time; beyond kFixupLongCap entries, insertion sort by the owning lane.
__global__ void __launch_bounds__(kBlock) k_fixup(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    __shared__ uint32_t lp[kBlock / 64][kFixupLongCap], ls[kBlock / 64][kFixupLongCap];
    const int C = s.I * s.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    int a = 0, n = 0;
    if (c < C) {
        a = s.cellStart[c];
        n = s.cellStart[c + 1] - a;
    }
    uint32_t* tp = s.tmpPid;
    uint32_t* ts = s.tmpSrc;
    if (n > 1 && n <= 16) sort_cell_regs<16>(tp, ts, a, n);
    else if (n > 16 && n <= 32) sort_cell_regs<32>(tp, ts, a, n);
    unsigned long long longs = __ballot(n > 32);   // every lane of the wave is here
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    while (longs) {
        const int owner = __ffsll((long long)longs) - 1;
        longs &= longs - 1;
        const int la = __shfl(a, owner, 64), ln = __shfl(n, owner, 64);
        if (ln > kFixupLongCap) {
            if (lane == owner) {
                for (int x = la + 1; x < la + ln; x++) {
                    const uint32_t kp = tp[x], ks = ts[x];
                    int y = x - 1;
                    while (y >= la && pair_gt(tp[y], ts[y], kp, ks)) {
                        tp[y + 1] = tp[y];
                        ts[y + 1] = ts[y];
                        y--;
                    }
                    tp[y + 1] = kp;
                    ts[y + 1] = ks;
                }
            }
            continue;
        }
        constexpr int Q = kFixupLongCap / 64;
        uint32_t kp[Q], ks[Q];
        int rank[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int j = lane + 64 * q;
            kp[q] = j < ln ? tp[la + j] : 0u;
            ks[q] = j < ln ? ts[la + j] : 0u;
            rank[q] = 0;
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int j = lane + 64 * q;
            if (j < ln) {
                lp[w][j] = kp[q];
                ls[w][j] = ks[q];
            }
        }
        // one wave writes and reads its own LDS rows: its LDS operations
        // complete in issue order, the barrier keeps the compiler's order
        __builtin_amdgcn_wave_barrier();
        for (int m = 0; m < ln; m++) {
            const uint32_t pm = lp[w][m], sm = ls[w][m];
#pragma unroll
            for (int q = 0; q < Q; q++) rank[q] += pair_gt(kp[q], ks[q], pm, sm) ? 1 : 0;
        }
        __builtin_amdgcn_wave_barrier();
        // every entry of the cell is in registers (the LDS stores consumed
        // all loads), so the in-place scatter reads nothing it overwrites
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int j = lane + 64 * q;
            if (j < ln) {
                tp[la + rank[q]] = kp[q];
                ts[la + rank[q]] = ks[q];
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_gather(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int total = s.cellStart[s.I * s.J];
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= total) return;
    const uint32_t src = s.tmpSrc[k];
    s.pos[k] = s.posA[src];
    s.vel[k] = s.velA[src];
    if (s.visc_on) s.pvisc[k] = s.pviscA[src];
    s.pid[k] = s.tmpPid[k];
    s.key[k] = 0;   // live
}

// ---------------------------------------------------------------------------
// One-block reduction of the s.pblocks per-block particle partials.
//   what & 1: commit the sorted count, sum the advect kills
//   what & 2: max |v|^2 (flipsolver2d.cpp:1562 starts from FLT_MIN) and
//             the countParticles erasures
//   what & 4: commit the sorted count only (re-sort after density correction)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_finalize(const SceneDev* scenes, int what)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int nb = s.pblocks;
    __shared__ int redi[kBlock / 64];
    __shared__ float redf[kBlock / 64];
    int kills = 0, erased = 0;
    float m = 0.f;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        if (what & 1) kills += s.ipart[b];
        if (what & 2) {
            erased += s.ipart[s.max_pblocks + b];
            m = fmaxf(m, s.fpart[b]);
        }
    }
    kills = block_sumi(kills, redi);
    erased = block_sumi(erased, redi);
    m = block_maxf(m, redf);
    if (threadIdx.x == 0) {
        SceneScalars* sc = s.sc;
        if (what & 1) {
            sc->n_particles = s.cellStart[s.I * s.J];
            sc->killed = kills;
        }
        if (what & 4) sc->n_particles = s.cellStart[s.I * s.J];
        if (what & 2) {
            const float fmin = 1.17549435e-38f;
            sc->max_v2_bits = __float_as_uint(m > fmin ? m : fmin);
            sc->erased = erased;
        }
    }
}

// ---------------------------------------------------------------------------
// G2P (FlipSolver::particleUpdate, flipsolver2d.cpp:361-388) fused with
// countParticles (1021-1051: per cell, particles beyond 2*ppc are erased;
// survivors are the smallest ids) and the next substep's max |v|^2.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_g2p_count(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || (int)blockIdx.x >= s.pblocks) return;
    __shared__ float redf[kBlock / 64];
    __shared__ int redi[kBlock / 64];
    const int n = s.sc->n_particles;
    const int I = s.I, J = s.J;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    float m = 0.f;
    int erased = 0;
    if (k < n) {
        const float2 p = s.pos[k];
        float2 v = s.vel[k];
        const float ou = lerp_grid(s.su, I + 1, J, 0.5f, 0.f, p.x, p.y);
        const float ov = lerp_grid(s.sv, I, J + 1, 0.f, 0.5f, p.x, p.y);
        const float2 nv = velocity_at(s.uf, s.vf, I, J, p.x, p.y);
        const float pic = s.pic;
        const float ax = nv.x * pic, ay = nv.y * pic;
        const float bx = ((v.x + nv.x) - ou) * (1.f - pic);
        const float by = ((v.y + nv.y) - ov) * (1.f - pic);
        v.x = ax + bx;
        v.y = ay + by;
        s.vel[k] = v;
        const int cell = integr(p.x) * J + integr(p.y);
        const int rank = k - s.cellStart[cell];
        if (rank >= 2 * s.ppc) {
            s.key[k] = kDead;
            erased = 1;
        } else {
            m = v.x * v.x + v.y * v.y;
        }
    }
    m = block_maxf(m, redf);
    erased = block_sumi(erased, redi);
    if (threadIdx.x == 0) {
        s.fpart[blockIdx.x] = m;
        s.ipart[s.max_pblocks + blockIdx.x] = erased;
    }
}

}  // namespace fg

namespace fg {
#define FG_LAUNCH0(name) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc) { fg_launch(k_##name, g, dim3(kBlock), st, sc); }
#define FG_LAUNCH1(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, T a) { fg_launch(k_##name, g, dim3(kBlock), st, sc, a); }
FG_LAUNCH0(max_v2)
FG_LAUNCH1(advect, int, alive_from_key)
FG_LAUNCH0(scan_local)
FG_LAUNCH0(scan_top)
FG_LAUNCH0(scan_add)
FG_LAUNCH0(scatter)
FG_LAUNCH0(fixup)
FG_LAUNCH0(gather)
FG_LAUNCH0(g2p_count)
void launch_finalize(dim3 g, hipStream_t st, const SceneDev* sc, int what)
{
    hipLaunchKernelGGL(k_finalize, g, dim3(kBlock), 0, st, sc, what);
}
}  // namespace fg
