// fg_particles.hip -- particle-space kernels of the substep.
//
// Particles live in one SoA arena per scene (float2 position, float2
// velocity, u32 stable id), kept sorted by (cell, id) after every substep.
// The reference's 3x3-cell bins and rebin sets
// (FlipSolver2dLib/markerparticlesystem.cpp:47-58, 260-269, 317-342) become a
// counting sort on the cell key with a per-cell id fixup, which is
// deterministic (no float atomics anywhere) and O(particles + cells).
#include "fg_kernels.h"

namespace fg {

// ---------------------------------------------------------------------------
// max |v|^2 over live particles (FlipSolver::maxParticleVelocity,
// flipsolver2d.cpp:1560-1574).  Max of non-negative floats is exact and
// order-free, so an integer atomicMax on the float bits is deterministic.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_max_v2(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    __shared__ float red[kBlock / 64];
    const int n = s.sc->n_particles;
    float m = 0.f;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        if (s.key[k] == kDead) continue;
        const float2 v = s.vel[k];
        m = fmaxf(m, v.x * v.x + v.y * v.y);
    }
    m = block_maxf(m, red);
    if (threadIdx.x == 0 && m > 0.f) atomicMax(&s.sc->max_v2_bits, __float_as_uint(m));
}

// ---------------------------------------------------------------------------
// Advection: FlipSolver::advectThread (flipsolver2d.cpp:305-338) with
// rk4Integrate (1195-1203) and SdfGrid::closestSurfacePoint
// (sdfgrid.cpp:15-46).  Fused with the cell histogram of the counting sort.
// ---------------------------------------------------------------------------
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

// simmath::gradCenteredGrid with truncating (ssize_t) arguments
__device__ __forceinline__ float2 grad_trunc(const float* g, int I, int J, float x, float y)
{
    const int i = (int)x, j = (int)y;
    const float ip1 = get_ext(g, I, J, i + 1, j), im1 = get_ext(g, I, J, i - 1, j);
    const float jp1 = get_ext(g, I, J, i, j + 1), jm1 = get_ext(g, I, J, i, j - 1);
    return make_float2((ip1 - im1) / 2.f, (jp1 - jm1) / 2.f);
}

__device__ float2 closest_surface_point(const float* sdf, int I, int J, float2 pos)
{
    float2 cp = pos;
    float value = lerp_grid(sdf, I, J, 0.f, 0.f, pos.x, pos.y);
    float2 grad = grad_trunc(sdf, I, J, pos.x, pos.y);
    for (int it = 0; it < 100; it++) {
        float alpha = 1;
        for (int k = 0; k < 10; k++) {
            const float t = alpha * value;
            const float2 q = make_float2(cp.x - grad.x * t, cp.y - grad.y * t);
            const float vq = lerp_grid(sdf, I, J, 0.f, 0.f, q.x, q.y);
            if (fabsf(vq) < fabsf(value)) {
                cp = q;
                value = vq;
                grad = grad_trunc(sdf, I, J, q.x, q.y);
                if (fabsf(value) < 1e-5f) return cp;
            } else {
                alpha *= 0.7f;
            }
        }
    }
    return cp;
}

__global__ void __launch_bounds__(kBlock) k_advect(const SceneDev* scenes, int alive_from_key)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    __shared__ int red[kBlock / 64];
    const int n = s.sc->n_particles;
    const int I = s.I, J = s.J;
    int killed = 0;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        const bool was_alive = !alive_from_key || s.key[k] != kDead;
        float2 p = s.pos[k];
        const float2 vel = s.vel[k];
        uint32_t key = kDead;
        if (was_alive) {
            p = rk4(s.uf, s.vf, I, J, p, s.dt);
            if (lerp_grid(s.ssdf, I, J, 0.f, 0.f, p.x, p.y) < 0.f) p = closest_surface_point(s.ssdf, I, J, p);
            const int pi = integr(p.x), pj = integr(p.y);
            const bool inb = pi >= 0 && pi < I && pj >= 0 && pj < J;
            if (!inb || is_sink(s.mat[(size_t)pi * J + pj])) {
                killed = 1;
            } else {
                key = (uint32_t)(pi * J + pj);
                atomicAdd(&s.cellCount[key], 1);
            }
        }
        s.posA[k] = p;
        s.velA[k] = vel;
        s.key[k] = key;
    }
    killed = block_sumi(killed, red);
    if (threadIdx.x == 0 && killed) atomicAdd(&s.sc->killed, killed);
}

// ---------------------------------------------------------------------------
// Counting sort on cell key.
//   scan: exclusive prefix sum of cellCount -> cellStart (3 passes)
//   scatter: slot = cellStart[key] + (--cellCount[key]); leaves cellCount 0
//   fixup: per cell, order the slots by stable id, then gather the payload
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
// pre-sort count until k_commit_count.
__global__ void __launch_bounds__(kBlock) k_scatter(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const int n = s.sc->n_particles;
    if (k >= n) return;
    const uint32_t key = s.key[k];
    if (key == kDead) return;
    const int slot = s.cellStart[key] + atomicSub(&s.cellCount[key], 1) - 1;
    s.tmpPid[slot] = s.pid[k];
    s.tmpSrc[slot] = (uint32_t)k;
}

__global__ void __launch_bounds__(kBlock) k_fixup_gather(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int C = s.I * s.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const int a = s.cellStart[c], e = s.cellStart[c + 1];
    // insertion sort of (pid, src) by pid -- segments hold ~ppc entries
    for (int x = a + 1; x < e; x++) {
        const uint32_t kp = s.tmpPid[x], ks = s.tmpSrc[x];
        int y = x - 1;
        while (y >= a && s.tmpPid[y] > kp) {
            s.tmpPid[y + 1] = s.tmpPid[y];
            s.tmpSrc[y + 1] = s.tmpSrc[y];
            y--;
        }
        s.tmpPid[y + 1] = kp;
        s.tmpSrc[y + 1] = ks;
    }
    for (int x = a; x < e; x++) {
        const uint32_t src = s.tmpSrc[x];
        s.pos[x] = s.posA[src];
        s.vel[x] = s.velA[src];
        s.pid[x] = s.tmpPid[x];
        s.key[x] = (uint32_t)c;
    }
}

__global__ void k_commit_count(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || threadIdx.x != 0) return;
    s.sc->n_particles = s.cellStart[s.I * s.J];
}

// ---------------------------------------------------------------------------
// G2P (FlipSolver::particleUpdate, flipsolver2d.cpp:361-388) fused with
// countParticles (1021-1051: per cell, particles beyond 2*ppc are erased;
// survivors are the smallest ids) and the next substep's max |v|^2.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_g2p_count(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
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
    if (threadIdx.x == 0 && m > 0.f) atomicMax(&s.sc->max_v2_bits, __float_as_uint(m));
    erased = block_sumi(erased, redi);
    if (threadIdx.x == 0 && erased) atomicAdd(&s.sc->erased, erased);
}

}  // namespace fg

namespace fg {
#define FG_LAUNCH0(name) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc) { hipLaunchKernelGGL(k_##name, g, dim3(kBlock), 0, st, sc); }
#define FG_LAUNCH1(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, T a) { hipLaunchKernelGGL(k_##name, g, dim3(kBlock), 0, st, sc, a); }
FG_LAUNCH0(max_v2)
FG_LAUNCH1(advect, int, alive_from_key)
FG_LAUNCH0(scan_local)
FG_LAUNCH0(scan_top)
FG_LAUNCH0(scan_add)
FG_LAUNCH0(scatter)
FG_LAUNCH0(fixup_gather)
void launch_commit_count(dim3 g, hipStream_t st, const SceneDev* sc) { hipLaunchKernelGGL(k_commit_count, g, dim3(64), 0, st, sc); }
FG_LAUNCH0(g2p_count)
}  // namespace fg
