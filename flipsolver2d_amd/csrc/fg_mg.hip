// fg_mg.hip -- multigrid V-cycle preconditioner of the pressure PCG.
//
// The reference preconditions its PCG with the inverse-Poisson
// preconditioner (inversepoissonpreconditioner.h:45-79), whose iteration
// count grows linearly with the grid side (about 140 / 280 / 565 iterations
// to 1e-6 on dam-break systems of 128^2 / 256^2 / 512^2 cells).  This V-cycle
// is a fixed symmetric positive definite linear operator, so PCG converges
// to the same solution of the same system; it takes 11 / 13 / 17.
//
// Hierarchy (rebuilt for every solve, since the fluid region changes):
//   level 0      the pressure matrix A' itself: 5-point, integer
//                coefficients, from SceneDev::stencil;
//   level l+1    cells of 2x2 level-l cells, active where any of the four
//                is; operator A_{l+1} = R A_l P (Galerkin), P the bilinear
//                cell-centred prolongation restricted to active cells
//                (weights 9/16, 3/16, 3/16, 1/16 from the four nearest
//                coarse cells), R = P^T / 4.  The coarse operators have 5x5
//                stencils; their entries are summed in FP64 and mirrored so
//                that each stored operator is exactly symmetric.
//   bottom       the first level with at most kMgBottomMax cells per side
//                (at least level 1).
// Smoother: one L1-Jacobi sweep before and after the coarse correction,
// x += (r - A x) / sum_k |a_k|, which converges for every symmetric
// positive definite A (a damped plain Jacobi sweep does not on all of these
// Galerkin operators); the bottom level gets kMgBottomSweeps sweeps from
// zero.  Every step is the same before and after, so the cycle is symmetric.
//
// Non-fluid rows of A' are identity rows: z = r there, as in the IPP.
// Level 0 as the V-cycle sees it is A' without the row-wrapping couplings
// the reference's linear offsets create at j = 0 and j = J-1 (a fluid cell
// on the grid edge, see k_stencil); the smoother's L1 diagonal still counts
// them, which only enlarges it.  The cycle stays symmetric; PCG itself
// multiplies by the full A'.
#include <type_traits>
#include <utility>

#include "fg_kernels.h"

namespace fg {

namespace {

__device__ __forceinline__ bool mg_iterating(const SceneDev& s)
{
    return s.active && s.precond == kPrecondMG && !s.sc->outer_done && !s.sc->inner_done;
}

// L1 diagonal of a level-0 row: n + (number of fluid neighbours)
__device__ __forceinline__ float l1_diag0(uint8_t code)
{
    return (float)(((code >> 4) & 7) + __popc(code & 15u));
}

// 1-D restriction weights of the fine cells 2c-1 .. 2c+2 (times 4)
__device__ __forceinline__ float w1(int a) { return (a == 1 || a == 2) ? 0.75f : 0.25f; }

// Coarse correction at fine cell (fi, fj) of a level whose coarser level is
// C: sum over the (up to) four nearest active coarse cells of weight * y.
__device__ __forceinline__ float prolong_at(const MgLevel& Cl, int fi, int fj)
{
    const int p0 = fi >> 1, q0 = fj >> 1;
    const int p1 = (fi & 1) ? p0 + 1 : p0 - 1, q1 = (fj & 1) ? q0 + 1 : q0 - 1;
    const int pi[2] = {p0, p1}, qj[2] = {q0, q1};
    const float wi[2] = {0.75f, 0.25f};
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int i = pi[a], j = qj[b];
            if (i < 0 || i >= Cl.I || j < 0 || j >= Cl.J) continue;
            acc += (wi[a] * wi[b]) * Cl.y[i * Cl.J + j];   // inactive coarse cells hold y = 0
        }
    return acc;
}

}  // namespace

// ---------------------------------------------------------------------------
// set-up, per level l >= 1 in increasing order
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_mg_mark(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.precond != kPrecondMG) return;
    const MgLevels& M = *s.mg;
    if (l >= M.n) return;
    const MgLevel& L = M.lv[l];
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= L.I * L.J) return;
    const int ci = c / L.J, cj = c - (c / L.J) * L.J;
    const int FI = l == 1 ? s.I : M.lv[l - 1].I, FJ = l == 1 ? s.J : M.lv[l - 1].J;
    uint8_t act = 0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int fi = 2 * ci + a, fj = 2 * cj + b;
            if (fi >= FI || fj >= FJ) continue;
            const int f = fi * FJ + fj;
            act |= l == 1 ? (uint8_t)(s.stencil[f] >> 7) : M.lv[l - 1].act[f];
        }
    L.act[c] = act;
}

// floor(t / 2) for small t of either sign
__device__ __forceinline__ constexpr int half_floor(int t) { return t >= 0 ? t / 2 : -((1 - t) / 2); }
__device__ __forceinline__ constexpr double w1c(int a) { return (a == 1 || a == 2) ? 0.75 : 0.25; }

// compile-time loop: f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>)
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int DI, int DJ>
__device__ __forceinline__ void galerkin_add(double (&acc)[kMgStencil], uint32_t am, double w)
{
    if constexpr (DI >= -2 && DI <= 2 && DJ >= -2 && DJ <= 2) {
        constexpr int k = (DI + 2) * 5 + (DJ + 2);
        if ((am >> k) & 1u) acc[k] += w;
    }
}

// One entry (value v) of fine row f = (2ci-1+A, 2cj-1+B) coupling it to
// g = f + (EI, EJ): spread R_f v P_g over g's coarse parents.
template <int A, int B, int EI, int EJ>
__device__ __forceinline__ void galerkin_spread(double (&acc)[kMgStencil], uint32_t am, double v)
{
    constexpr int ti = -1 + A + EI, tj = -1 + B + EJ;   // g - 2C
    constexpr int pi0 = half_floor(ti), pj0 = half_floor(tj);
    constexpr int pi1 = (ti & 1) ? pi0 + 1 : pi0 - 1, pj1 = (tj & 1) ? pj0 + 1 : pj0 - 1;
    constexpr double wc = w1c(A) * w1c(B);
    galerkin_add<pi0, pj0>(acc, am, (wc * 0.5625) * v);
    galerkin_add<pi0, pj1>(acc, am, (wc * 0.1875) * v);
    galerkin_add<pi1, pj0>(acc, am, (wc * 0.1875) * v);
    galerkin_add<pi1, pj1>(acc, am, (wc * 0.0625) * v);
}

// A_l = R A_{l-1} P, one coarse row C = (ci, cj) per thread: the fine cells
// f, their stencil entries g and g's coarse parents sit at compile-time
// offsets from C, so the 25 FP64 accumulators stay in registers.
template <bool FINE0>
__device__ __forceinline__ void galerkin_row(const SceneDev& s, const MgLevels& M, int l, int c, int ci, int cj)
{
    const MgLevel& L = M.lv[l];
    const int CL = L.I * L.J;
    // activity of the 5x5 coarse neighbourhood, bit (di+2)*5 + (dj+2)
    uint32_t am = 0;
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
        if (ni >= 0 && ni < L.I && nj >= 0 && nj < L.J && L.act[ni * L.J + nj]) am |= 1u << k;
    }
    double acc[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) acc[k] = 0.0;
    const int FI = FINE0 ? s.I : M.lv[l - 1].I, FJ = FINE0 ? s.J : M.lv[l - 1].J;
    const int CF = FI * FJ;
    static_for<16>([&](auto ab) {
        constexpr int a = decltype(ab)::value / 4, b = decltype(ab)::value % 4;
        const int fi = 2 * ci - 1 + a, fj = 2 * cj - 1 + b;
        if (fi < 0 || fi >= FI || fj < 0 || fj >= FJ) return;
        const int f = fi * FJ + fj;
        if constexpr (FINE0) {
            const uint8_t code = s.stencil[f];
            if (!(code & 0x80)) return;
            galerkin_spread<a, b, 0, 0>(acc, am, (double)((code >> 4) & 7));
            galerkin_spread<a, b, -1, 0>(acc, am, (code & 1) ? -1.0 : 0.0);
            galerkin_spread<a, b, 1, 0>(acc, am, (code & 2) ? -1.0 : 0.0);
            galerkin_spread<a, b, 0, -1>(acc, am, ((code & 4) && fj > 0) ? -1.0 : 0.0);   // not the row-wrapping coupling
            galerkin_spread<a, b, 0, 1>(acc, am, ((code & 8) && fj < FJ - 1) ? -1.0 : 0.0);
        } else {
            if (!M.lv[l - 1].act[f]) return;
            const float* fc = M.lv[l - 1].coef;
            static_for<kMgStencil>([&](auto ee) {
                constexpr int e = decltype(ee)::value;
                galerkin_spread<a, b, e / 5 - 2, e % 5 - 2>(acc, am, (double)fc[(size_t)e * CF + f]);
            });
        }
    });
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) L.coef[(size_t)k * CL + c] = (float)(0.25 * acc[k]);
}

__global__ void __launch_bounds__(kBlock) k_mg_galerkin(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.precond != kPrecondMG) return;
    const MgLevels& M = *s.mg;
    if (l >= M.n) return;
    const MgLevel& L = M.lv[l];
    const int CL = L.I * L.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= CL) return;
    if (!L.act[c]) {
#pragma unroll
        for (int k = 0; k < kMgStencil; k++) L.coef[(size_t)k * CL + c] = 0.f;
        return;
    }
    const int ci = c / L.J, cj = c - (c / L.J) * L.J;
    if (l == 1) galerkin_row<true>(s, M, l, c, ci, cj);
    else galerkin_row<false>(s, M, l, c, ci, cj);
}

// Mirror the upper entries into the lower ones (exact symmetry) and form the
// L1 diagonal.  Reads only entries k > 12, writes only k < 12 of its own row.
__global__ void __launch_bounds__(kBlock) k_mg_sym(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.precond != kPrecondMG) return;
    const MgLevels& M = *s.mg;
    if (l >= M.n) return;
    const MgLevel& L = M.lv[l];
    const int CL = L.I * L.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= CL) return;
    const int ci = c / L.J, cj = c - (c / L.J) * L.J;
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        float v;
        if (k < 12) {
            const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
            v = (ni >= 0 && ni < L.I && nj >= 0 && nj < L.J) ? L.coef[(size_t)(24 - k) * CL + ni * L.J + nj] : 0.f;
            L.coef[(size_t)k * CL + c] = v;
        } else {
            v = L.coef[(size_t)k * CL + c];
        }
        sum += fabsf(v);
    }
    L.dinv[c] = (L.act[c] && sum > 0.f) ? 1.0f / sum : 0.f;
}

// ---------------------------------------------------------------------------
// the V-cycle: z = M r, r the current PCG residual (which: 0 = cr, 1 = cr2)
// ---------------------------------------------------------------------------

// Level 0 -> 1: x0 = r / D1 on the fine level, res0 = r - A' x0, and
// r1 = R res0, one level-1 cell per thread over its 4x4 fine support (a 6x6
// fine patch with the stencil's reach).
__global__ void __launch_bounds__(kBlock) k_mg_down0(const SceneDev* scenes, int which)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!mg_iterating(s)) return;
    const MgLevel& L1 = s.mg->lv[1];
    const int c = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= L1.I * L1.J) return;
    if (!L1.act[c]) {
        L1.r[c] = 0.f;
        return;
    }
    const int ci = c / L1.J, cj = c - (c / L1.J) * L1.J;
    const int I = s.I, J = s.J;
    const float* __restrict__ r = which ? s.cr2 : s.cr;
    uint8_t cd[6][6];
    float rv[6][6];
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const int fi = 2 * ci - 2 + a, fj = 2 * cj - 2 + b;
            const bool in = fi >= 0 && fi < I && fj >= 0 && fj < J;
            cd[a][b] = in ? s.stencil[fi * J + fj] : 0;
            rv[a][b] = in ? r[fi * J + fj] : 0.f;
        }
    float x0[6][6];
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const float d = l1_diag0(cd[a][b]);
            x0[a][b] = ((cd[a][b] & 0x80) && d > 0.f) ? rv[a][b] / d : 0.f;
        }
    float acc = 0.f;
#pragma unroll
    for (int a = 1; a < 5; a++)
#pragma unroll
        for (int b = 1; b < 5; b++) {
            const uint8_t k = cd[a][b];
            if (!(k & 0x80)) continue;
            float ax = (float)((k >> 4) & 7) * x0[a][b];
            if (k & 1) ax -= x0[a - 1][b];
            if (k & 2) ax -= x0[a + 1][b];
            if (k & 4) ax -= x0[a][b - 1];
            if (k & 8) ax -= x0[a][b + 1];
            acc += (w1(a - 1) * w1(b - 1)) * (rv[a][b] - ax);
        }
    L1.r[c] = 0.25f * acc;
}

// Level l (1 <= l <= n-2): x = r * dinv, y = res = r - A_l x.
__global__ void __launch_bounds__(kBlock) k_mg_resid(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!mg_iterating(s)) return;
    const MgLevels& M = *s.mg;
    if (l > M.n - 2) return;
    const MgLevel& L = M.lv[l];
    const int CL = L.I * L.J;
    const int c = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= CL) return;
    if (!L.act[c]) {
        L.y[c] = 0.f;
        return;
    }
    const int ci = c / L.J, cj = c - (c / L.J) * L.J;
    float ax = 0.f;
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
        if (ni < 0 || ni >= L.I || nj < 0 || nj >= L.J) continue;
        const int n = ni * L.J + nj;
        ax += L.coef[(size_t)k * CL + c] * (L.r[n] * L.dinv[n]);
    }
    L.y[c] = L.r[c] - ax;
}

// Level l-1 -> l (2 <= l <= n-1): r_l = R y_{l-1}.
__global__ void __launch_bounds__(kBlock) k_mg_restrict(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!mg_iterating(s)) return;
    const MgLevels& M = *s.mg;
    if (l > M.n - 1) return;
    const MgLevel& L = M.lv[l];
    const MgLevel& F = M.lv[l - 1];
    const int c = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= L.I * L.J) return;
    if (!L.act[c]) {
        L.r[c] = 0.f;
        return;
    }
    const int ci = c / L.J, cj = c - (c / L.J) * L.J;
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int fi = 2 * ci - 1 + a;
        if (fi < 0 || fi >= F.I) continue;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int fj = 2 * cj - 1 + b;
            if (fj < 0 || fj >= F.J) continue;
            acc += (w1(a) * w1(b)) * F.y[fi * F.J + fj];   // inactive fine cells hold y = 0
        }
    }
    L.r[c] = 0.25f * acc;
}

// Bottom level: kMgBottomSweeps L1-Jacobi sweeps from zero, one block per
// scene, one cell per thread, the iterate in LDS with a two-cell zero halo
// (so the stencil needs no bounds tests); result in y.
__global__ void __launch_bounds__(kBlock) k_mg_bottom(const SceneDev* scenes)
{
    static_assert(kMgBottomMax * kMgBottomMax <= kBlock, "one bottom cell per thread");
    constexpr int W = kMgBottomMax + 4;
    __shared__ float xs[2][W * W];
    const SceneDev& s = scenes[blockIdx.y];
    if (!mg_iterating(s)) return;
    const MgLevels& M = *s.mg;
    const MgLevel& L = M.lv[M.n - 1];
    const int CL = L.I * L.J;   // <= kBlock (the hierarchy stops there)
    const int c = threadIdx.x;
    const bool in = c < CL;
    const int ci = in ? c / L.J : 0, cj = in ? c - ci * L.J : 0;
    float cf[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) cf[k] = in ? L.coef[(size_t)k * CL + c] : 0.f;
    const float dv = in ? L.dinv[c] : 0.f, rv = in ? L.r[c] : 0.f;
    for (int t = threadIdx.x; t < 2 * W * W; t += kBlock) (&xs[0][0])[t] = 0.f;
    __syncthreads();
    const int home = (ci + 2) * W + (cj + 2);
    int cur = 0;
    for (int sw = 0; sw < kMgBottomSweeps; sw++) {
        if (in) {
            float ax = 0.f;
#pragma unroll
            for (int k = 0; k < kMgStencil; k++) ax += cf[k] * xs[cur][home + (k / 5 - 2) * W + (k % 5 - 2)];
            xs[cur ^ 1][home] = xs[cur][home] + dv * (rv - ax);
        }
        __syncthreads();
        cur ^= 1;
    }
    if (in) L.y[c] = xs[cur][home];
}

// Level l (1 <= l <= n-2), up: x = r * dinv + P y_{l+1}.
__global__ void __launch_bounds__(kBlock) k_mg_prolong(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!mg_iterating(s)) return;
    const MgLevels& M = *s.mg;
    if (l > M.n - 2) return;
    const MgLevel& L = M.lv[l];
    const int c = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= L.I * L.J) return;
    if (!L.act[c]) {
        L.x[c] = 0.f;
        return;
    }
    const int ci = c / L.J, cj = c - (c / L.J) * L.J;
    L.x[c] = L.r[c] * L.dinv[c] + prolong_at(M.lv[l + 1], ci, cj);
}

// Level l (1 <= l <= n-2), up: y = x + dinv (r - A_l x).
__global__ void __launch_bounds__(kBlock) k_mg_post(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!mg_iterating(s)) return;
    const MgLevels& M = *s.mg;
    if (l > M.n - 2) return;
    const MgLevel& L = M.lv[l];
    const int CL = L.I * L.J;
    const int c = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (c >= CL) return;
    if (!L.act[c]) {
        L.y[c] = 0.f;
        return;
    }
    const int ci = c / L.J, cj = c - (c / L.J) * L.J;
    float ax = 0.f;
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
        if (ni < 0 || ni >= L.I || nj < 0 || nj >= L.J) continue;
        ax += L.coef[(size_t)k * CL + c] * L.x[ni * L.J + nj];
    }
    L.y[c] = L.x[c] + L.dinv[c] * (L.r[c] - ax);
}

// Level 0, up: z0 = r / D1 + P y1 at the cell and its fluid neighbours, then
// z = z0 + (r - A' z0) / D1; z = r on non-fluid rows.  Writes z and the
// partials of r.z and r.r the next PCG iteration reduces (solver chunks).
__global__ void __launch_bounds__(kBlock) k_mg_up0(const SceneDev* scenes, int which)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!mg_iterating(s)) return;
    if ((int)blockIdx.x >= s.pcg_blocks) return;
    const int J = s.J;
    const MgLevel& L1 = s.mg->lv[1];
    const float* __restrict__ r = which ? s.cr2 : s.cr;
    const uint8_t* __restrict__ code = s.stencil;
    const Chunk ch = my_chunk(s);
    double rz = 0.0, rr = 0.0;
    for (int c = ch.begin + threadIdx.x; c < ch.end; c += kBlock) {
        const uint8_t k = code[c];
        const float rc = r[c];
        float z = rc;
        const float dc = l1_diag0(k);
        if ((k & 0x80) && dc > 0.f) {
            const int ci = c / J, cj = c - (c / J) * J;
            const int offs[5] = {0, -J, J, -1, 1};
            const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
            float z0[5];
#pragma unroll
            for (int e = 0; e < 5; e++) {
                z0[e] = 0.f;
                if (e > 0 && !(k & (1 << (e - 1)))) continue;   // only fluid neighbours enter A'
                if ((e == 3 && cj == 0) || (e == 4 && cj == J - 1)) continue;   // row-wrapping coupling
                const int g = c + offs[e];
                const uint8_t kg = e == 0 ? k : code[g];
                const float dg = l1_diag0(kg);
                const float rg = e == 0 ? rc : r[g];
                z0[e] = (dg > 0.f ? rg / dg : 0.f) + prolong_at(L1, ci + di[e], cj + dj[e]);
            }
            float ax = (float)((k >> 4) & 7) * z0[0];
#pragma unroll
            for (int e = 1; e < 5; e++) ax -= z0[e];   // zero where no coupling
            z = z0[0] + (rc - ax) / dc;
        }
        s.cu[c] = z;
        rz += (double)rc * (double)z;
        rr += (double)rc * (double)rc;
    }
    double v[2] = {rz, rr};
    put_partials<2>(s, kSlotCG, v);
}

}  // namespace fg

namespace fg {
#define FG_MG_LAUNCH1(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, T a) { fg_launch(k_##name, g, dim3(kBlock), st, sc, a); }
FG_MG_LAUNCH1(mg_mark, int, l)
FG_MG_LAUNCH1(mg_galerkin, int, l)
FG_MG_LAUNCH1(mg_sym, int, l)
FG_MG_LAUNCH1(mg_down0, int, which)
FG_MG_LAUNCH1(mg_resid, int, l)
FG_MG_LAUNCH1(mg_restrict, int, l)
FG_MG_LAUNCH1(mg_prolong, int, l)
FG_MG_LAUNCH1(mg_post, int, l)
FG_MG_LAUNCH1(mg_up0, int, which)
void launch_mg_bottom(dim3 g, hipStream_t st, const SceneDev* sc) { fg_launch(k_mg_bottom, g, dim3(kBlock), st, sc); }
}  // namespace fg
