// fg_mg.hip -- multigrid V-cycle preconditioner of the pressure PCG.
//
// The reference preconditions its PCG with the inverse-Poisson
// preconditioner (inversepoissonpreconditioner.h:45-79), whose iteration
// count grows linearly with the grid side (about 140 / 280 / 565 iterations
// to 1e-6 on dam-break systems of 128^2 / 256^2 / 512^2 cells).  This V-cycle
// is a fixed symmetric positive definite linear operator, so PCG converges
// to the same solution of the same system; it takes 9 / 11 / 14.
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
// x += w (r - A x) / sum_k |a_k| with w = kMgOmega = 1.5.  The L1 diagonal
// dominates A, so the eigenvalues of A / sum_k |a_k| lie in (0, 1] and the
// sweep converges for every symmetric positive definite A and every w < 2
// (a damped plain Jacobi sweep does not on all of these Galerkin
// operators).  Over-relaxed weights take fewer PCG iterations: from zero
// on dam-break systems of 128^2 / 256^2 / 512^2 cells 11 / 13 / 17 with
// w = 1, 9 / 11 / 14 with 1.5 or 1.6, more again from 1.8; on the
// warm-started solves of 256 mixed scenes 6.2 a solve on average with 1.5
// (1.3: 6.6, 1.4: 6.4, 1.6: 6.5, 1.7: 7.7).  The bottom
// level gets kMgBottomSweeps sweeps from zero.  Every step is the same
// before and after, so the cycle is symmetric.
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

// Both stop flags are read unconditionally (scalar loads issued together
// with the caller's other uniform loads, no branch between them).
__device__ __forceinline__ bool mg_iterating(const SceneDev& s)
{
    cptr<SceneScalars> sc = cst(s.sc);
    const int od = sc->outer_done, id = sc->inner_done;
    return (s.active != 0) & (s.precond == kPrecondMG) & (od == 0) & (id == 0);
}

// L1 diagonal of a level-0 row: n + (number of fluid neighbours)
__device__ __forceinline__ float l1_diag0(uint8_t code)
{
    return (float)(((code >> 4) & 7) + __popc(code & 15u));
}

// 1-D restriction weights of the fine cells 2c-1 .. 2c+2 (times 4)
__device__ __forceinline__ float w1(int a) { return (a == 1 || a == 2) ? 0.75f : 0.25f; }

// Coarse correction at fine cell (fi, fj) from the coarser level's y (I x J):
// the (up to) four nearest coarse cells, weights 9/16, 3/16, 3/16, 1/16.
// The four loads are unconditional (clamped index, weight 0 outside), so
// they issue together.  Inactive coarse cells hold y = 0.
template <typename P>
__device__ __forceinline__ float prolong_at(P y, int I, int J, int fi, int fj)
{
    const int p0 = fi >> 1, q0 = fj >> 1;
    const int p1 = (fi & 1) ? p0 + 1 : p0 - 1, q1 = (fj & 1) ? q0 + 1 : q0 - 1;
    const int pi[2] = {p0, p1}, qj[2] = {q0, q1};
    const float wi[2] = {0.75f, 0.25f};
    float v[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int i = min(max(pi[a], 0), I - 1), j = min(max(qj[b], 0), J - 1);
            v[a][b] = y[i * J + j];
        }
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const bool ok = pi[a] >= 0 && pi[a] < I && qj[b] >= 0 && qj[b] < J;
            acc += ((wi[a] * wi[b]) * v[a][b]) * mask01(ok);
        }
    return acc;
}

// The 25 stored coefficients of row c: one 128-byte line (the row's
// kMgRowStride floats), seven 16-byte loads issued together.
typedef float coef4 __attribute__((ext_vector_type(4)));   // a builtin vector: loads from any address space
__device__ __forceinline__ void load_coef_row(gptr<const float> coef, int c, float (&a)[kMgStencil])
{
    gptr<const coef4> r4 = (gptr<const coef4>)(coef + (size_t)c * kMgRowStride);
    coef4 q[7];
#pragma unroll
    for (int u = 0; u < 7; u++) q[u] = r4[u];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) a[k] = q[k >> 2][k & 3];
}

// sum_k a_k v_k over an active level-l row (dinv != 0) with the level's 25
// stencil entries: the canonical stencil on canonical rows (dinv < 0), else
// the stored coefficients, loaded only by the lanes whose rows need them.
template <typename MT>
__device__ __forceinline__ float row_dot(const MT& M, int l, gptr<const float> coef, int CL, int c,
                                         float dinv_c, const float (&v)[kMgStencil])
{
    (void)CL;
    const auto& cn = M.canon[l];
    float a[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) a[k] = cn[k];
    if (dinv_c > 0.f) load_coef_row(coef, c, a);
    float ax = 0.f;
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) ax += a[k] * v[k];
    return ax;
}

// r at coarse cell (ci, cj) from the finer level's y (FI x FJ): 1-D weights
// 1/4, 3/4, 3/4, 1/4 over fine cells 2c-1 .. 2c+2, times 1/4.  Inactive fine
// cells hold y = 0; the loads are unconditional (clamped index, weight 0
// outside the grid).
template <typename P>
__device__ __forceinline__ float restrict_at(P fy, int FI, int FJ, int ci, int cj)
{
    float v[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int fi = min(max(2 * ci - 1 + a, 0), FI - 1), fj = min(max(2 * cj - 1 + b, 0), FJ - 1);
            v[a][b] = fy[fi * FJ + fj];
        }
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int fi = 2 * ci - 1 + a, fj = 2 * cj - 1 + b;
            const bool ok = fi >= 0 && fi < FI && fj >= 0 && fj < FJ;
            acc += ((w1(a) * w1(b)) * v[a][b]) * mask01(ok);
        }
    return 0.25f * acc;
}

// The 5x5 neighbourhood of (ci, cj) in x (LI x LJ), 0 outside the grid
// (clamped loads times a 0/1 factor, so every load issues).
template <typename P>
__device__ __forceinline__ void gather25(P x, int LI, int LJ, int ci, int cj, float (&xv)[kMgStencil])
{
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
        const bool ok = ni >= 0 && ni < LI && nj >= 0 && nj < LJ;
        const float v = x[min(max(ni, 0), LI - 1) * LJ + min(max(nj, 0), LJ - 1)];
        xv[k] = v * mask01(ok);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// set-up, per level l >= 1 in increasing order (one launch per level for all
// scenes of the group; a thread per level-l cell)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mg_mark_block(const SceneDev& s, int l, int bx)
{
    if (!s.active || s.precond != kPrecondMG) return;
    const auto& M = *cst(s.mg);
    if (l >= M.n) return;
    const auto& L = M.lv[l];
    const int c = bx * kBlock + threadIdx.x;
    if (c >= L.I * L.J) return;
    const int ci = c / L.J, cj = c - ci * L.J;
    const int FI = l == 1 ? s.I : M.lv[l - 1].I, FJ = l == 1 ? s.J : M.lv[l - 1].J;
    gptr<const uint8_t> fact = gbl<const uint8_t>(l == 1 ? s.stencil : M.lv[l - 1].act);
    uint8_t act = 0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int fi = 2 * ci + a, fj = 2 * cj + b;
            const bool ok = fi < FI && fj < FJ;
            const uint8_t v = fact[ok ? fi * FJ + fj : 0];
            if (ok) act |= l == 1 ? (uint8_t)(v >> 7) : v;
        }
    gbl(L.act)[c] = act;
}

__global__ void __launch_bounds__(kBlock) k_mg_mark(const SceneDev* scenes, int l)
{
    mg_mark_block(scenes[blockIdx.y], l, blockIdx.x);
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
        acc[k] += ((am >> k) & 1u) ? w : 0.0;   // a select, not a branch
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
template <bool FINE0, typename MT>
__device__ __forceinline__ void galerkin_row(const SceneDev& s, const MT& M, int l, int c, int ci, int cj)
{
    const auto& L = M.lv[l];
    const int CL = L.I * L.J;
    // activity of the 5x5 coarse neighbourhood, bit (di+2)*5 + (dj+2)
    uint32_t am = 0;
    gptr<const uint8_t> act = gbl<const uint8_t>(L.act);
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {   // clamped loads, all issued together
        const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
        const bool in = ni >= 0 && ni < L.I && nj >= 0 && nj < L.J;
        const uint8_t v = act[min(max(ni, 0), L.I - 1) * L.J + min(max(nj, 0), L.J - 1)];
        am |= (in && v) ? 1u << k : 0u;
    }
    double acc[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) acc[k] = 0.0;
    const int FI = FINE0 ? s.I : M.lv[l - 1].I, FJ = FINE0 ? s.J : M.lv[l - 1].J;
    const int CF = FI * FJ;
    // Every load is unconditional (clamped fine index; a fine cell outside
    // the grid contributes zero), so the loads of all 16 fine rows can be in
    // flight together instead of one branch-guarded batch after another.
    // Inactive fine rows hold all-zero coefficients (level >= 1) or a
    // non-fluid code (level 0): they add nothing.
    if constexpr (FINE0) {
        gptr<const uint8_t> code = gbl<const uint8_t>(s.stencil);
        uint8_t cd[16];
        static_for<16>([&](auto ab) {
            constexpr int a = decltype(ab)::value / 4, b = decltype(ab)::value % 4;
            const int fi = 2 * ci - 1 + a, fj = 2 * cj - 1 + b;
            const bool in = fi >= 0 && fi < FI && fj >= 0 && fj < FJ;
            const uint8_t k = code[min(max(fi, 0), FI - 1) * FJ + min(max(fj, 0), FJ - 1)];
            cd[decltype(ab)::value] = in ? k : 0;
        });
        static_for<16>([&](auto ab) {
            constexpr int a = decltype(ab)::value / 4, b = decltype(ab)::value % 4;
            const int fj = 2 * cj - 1 + b;
            const uint8_t k = cd[decltype(ab)::value];
            const double on = (k & 0x80) ? 1.0 : 0.0;
            galerkin_spread<a, b, 0, 0>(acc, am, on * (double)((k >> 4) & 7));
            galerkin_spread<a, b, -1, 0>(acc, am, (k & 1) ? -on : 0.0);
            galerkin_spread<a, b, 1, 0>(acc, am, (k & 2) ? -on : 0.0);
            galerkin_spread<a, b, 0, -1>(acc, am, ((k & 4) && fj > 0) ? -on : 0.0);   // not the row-wrapping coupling
            galerkin_spread<a, b, 0, 1>(acc, am, ((k & 8) && fj < FJ - 1) ? -on : 0.0);
        });
    } else {
        gptr<const float> fc = gbl<const float>(M.lv[l - 1].coef);
        gptr<const char> fb = (gptr<const char>)fc;
        static_for<8>([&](auto hh) {   // two fine cells at a time: 50 loads in flight
            constexpr int a = decltype(hh)::value / 2, b0 = 2 * (decltype(hh)::value % 2);
            float v[2][kMgStencil];
            bool in[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int fi = 2 * ci - 1 + a, fj = 2 * cj - 1 + b0 + h;
                in[h] = fi >= 0 && fi < FI && fj >= 0 && fj < FJ;
                const uint32_t f = (uint32_t)(min(max(fi, 0), FI - 1) * FJ + min(max(fj, 0), FJ - 1));
#pragma unroll
                for (int e = 0; e < kMgStencil; e++)
                    v[h][e] = *(gptr<const float>)(fb + (f * (uint32_t)kMgRowStride + (uint32_t)e) * 4u);
            }
            static_for<2>([&](auto h2) {
                constexpr int h = decltype(h2)::value, b = b0 + h;
                static_for<kMgStencil>([&](auto ee) {
                    constexpr int e = decltype(ee)::value;
                    galerkin_spread<a, b, e / 5 - 2, e % 5 - 2>(acc, am, in[h] ? (double)v[h][e] : 0.0);
                });
            });
        });
    }
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) L.coef[(size_t)c * kMgRowStride + k] = (float)(0.25 * acc[k]);
}

// An inactive row must hold zeros; its dinv, not yet rewritten for this
// hierarchy, says whether it held any: +0 (bit pattern 0) only on a row that
// was inactive before (or never set: the arena starts zeroed), whose
// coefficients are zero already -- most of the level, and most of its writes.
__global__ void __launch_bounds__(kBlock) k_mg_galerkin(const SceneDev* scenes, int l)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.precond != kPrecondMG) return;
    const auto& M = *cst(s.mg);
    if (l >= M.n) return;
    const auto& L = M.lv[l];
    const int CL = L.I * L.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= CL) return;
    if (!L.act[c]) {
        if (__float_as_uint(gbl<const float>(L.dinv)[c]) != 0u) {
#pragma unroll
            for (int k = 0; k < kMgStencil; k++) L.coef[(size_t)c * kMgRowStride + k] = 0.f;
        }
        return;
    }
    const int ci = c / L.J, cj = c - ci * L.J;
    if (l == 1) galerkin_row<true>(s, M, l, c, ci, cj);
    else galerkin_row<false>(s, M, l, c, ci, cj);
}

// Mirror the upper entries into the lower ones (exact symmetry) and form the
// L1 diagonal.  Reads only entries k > 12, writes only k < 12 of its own row.
// Rows equal to the level's canonical interior stencil (every entry within
// 1e-5 of the canonical diagonal; other rows differ by more than 1 %) are
// flagged by a negative dinv and read the canonical stencil later.
// A block takes a 16 x 16 tile of rows: the upper entries (13..24) of the
// tile's rows and of its two-row halo are staged in LDS with row-contiguous
// loads (a row's entries share one line), so the mirrored reads, which
// touch twelve neighbouring rows each, come from LDS.
constexpr int kSymTile = 16, kSymR = kSymTile + 4, kSymUp = 13;   // 12 entries, padded to an odd stride
__device__ __forceinline__ int sym_tiles(int LI, int LJ)
{
    return ((LI + kSymTile - 1) / kSymTile) * ((LJ + kSymTile - 1) / kSymTile);
}

__device__ __forceinline__ void mg_sym_block(const SceneDev& s, int l, int bx)
{
    static_assert(kBlock == kSymTile * kSymTile, "one tile row per thread");
    if (!s.active || s.precond != kPrecondMG) return;
    const auto& M = *cst(s.mg);
    if (l >= M.n) return;
    const auto& L = M.lv[l];
    const int LI = L.I, LJ = L.J;
    if (bx >= sym_tiles(LI, LJ)) return;   // the whole block is past the level
    __shared__ float up[kSymR * kSymR][kSymUp];   // entries 13..24 of the region's rows
    const int tilesJ = (LJ + kSymTile - 1) / kSymTile;
    const int ti = bx / tilesJ, tj = bx - ti * tilesJ;
    const int ri0 = ti * kSymTile - 2, rj0 = tj * kSymTile - 2;
    gptr<float> coef = gbl(L.coef);
    typedef float v4 __attribute__((ext_vector_type(4)));
    // stage: region row q's floats 12..27 as four 16-byte loads, entries 13..24 kept
    for (int w = threadIdx.x; w < kSymR * kSymR * 4; w += kBlock) {
        const int q = w >> 2, part = w & 3;
        const int gi = ri0 + q / kSymR, gj = rj0 + q % kSymR;
        const bool in = gi >= 0 && gi < LI && gj >= 0 && gj < LJ;
        const int n = min(max(gi, 0), LI - 1) * LJ + min(max(gj, 0), LJ - 1);
        const v4 x = *(gptr<const v4>)(coef + (size_t)n * kMgRowStride + 12 + 4 * part);
        const float m = mask01(in);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int k = 12 + 4 * part + e;
            if (k >= 13 && k <= 24) up[q][k - 13] = x[e] * m;
        }
    }
    __syncthreads();
    const int li = threadIdx.x / kSymTile, lj = threadIdx.x % kSymTile;
    const int ci0 = ti * kSymTile + li, cj0 = tj * kSymTile + lj;
    const bool valid = ci0 < LI && cj0 < LJ;
    const int ci = min(ci0, LI - 1), cj = min(cj0, LJ - 1), c = ci * LJ + cj;
    float v[kMgStencil];
    {   // own row: entries 12..24 (and the next three floats) in four loads
        v4 o[4];
#pragma unroll
        for (int u = 0; u < 4; u++) o[u] = *(gptr<const v4>)(coef + (size_t)c * kMgRowStride + 12 + 4 * u);
#pragma unroll
        for (int k = 12; k < kMgStencil; k++) v[k] = o[(k - 12) >> 2][(k - 12) & 3];
    }
    const uint8_t av = gbl<const uint8_t>(L.act)[c];
    // lower entries: the mirrored upper entry 24 - k of neighbour k (outside
    // the level: 0, as staged)
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const int q = (li + 2 + k / 5 - 2) * kSymR + (lj + 2 + k % 5 - 2);
        v[k] = up[q][(24 - k) - 13];
    }
    float cn[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) cn[k] = M.canon[l][k];
    const float tol = 1e-5f * cn[12];
    float sum = 0.f;
    bool canon_row = true;
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        sum += fabsf(v[k]);
        canon_row = canon_row && fabsf(v[k] - cn[k]) <= tol;
    }
    const bool act = av != 0;
    // an inactive row is all zero already (k_mg_galerkin), mirrored entries
    // included; the lower entries go out as three 16-byte stores
    if (valid && act) {
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const v4 x = {v[4 * u], v[4 * u + 1], v[4 * u + 2], v[4 * u + 3]};
            *(gptr<v4>)(coef + (size_t)c * kMgRowStride + 4 * u) = x;
        }
    }
    const float di = (act && sum > 0.f) ? kMgOmega / sum : 0.f;
    if (valid) gbl(L.dinv)[c] = (canon_row && di > 0.f) ? -di : (act && di == 0.f ? -0.f : di);
    // active rows that will read their stored coefficients (for the kernels'
    // byte counts): one atomic per block, once per solve
    const int stored = __syncthreads_count(valid && di > 0.f && !canon_row);
    if (threadIdx.x == 0 && stored) atomicAdd(&s.sc->mg_stored[l], stored);
}

__global__ void __launch_bounds__(kBlock) k_mg_sym(const SceneDev* scenes, int l)
{
    mg_sym_block(scenes[blockIdx.y], l, blockIdx.x);
}

// k_mg_sym of level l (blocks [0, nsym)) and k_mg_mark of level l+1 (the
// rest) in one launch: they touch different arrays, and both must be done
// before the Galerkin product of level l+1.
__global__ void __launch_bounds__(kBlock) k_mg_symmark(const SceneDev* scenes, int l, int nsym)
{
    const SceneDev& s = scenes[blockIdx.y];
    if ((int)blockIdx.x < nsym) mg_sym_block(s, l, blockIdx.x);
    else mg_mark_block(s, l + 1, blockIdx.x - nsym);
}

// ---------------------------------------------------------------------------
// the V-cycle: z = M r, r the current PCG residual (which: 0 = cr, 1 = cr2).
// Each kernel takes its scenes and blocks from a block map (flat launch over
// the packed iterating scenes) or from a (blocks, scenes) grid.  All loads
// of a thread are unconditional (clamped indices, masked afterwards) so they
// issue together; the kernels are latency-bound.
// ---------------------------------------------------------------------------

// Level 0 -> 1: x0 = w r / D1 on the fine level, res0 = r - A' x0, and
// r1 = R res0 over each level-1 cell's 4x4 fine support.  A block owns a
// kD0Tile x kD0Tile tile of level-1 cells (mg_d0_tiles blocks a scene) and
// stages its fine patch (the supports with the stencil's reach, kD0R cells
// a side) in LDS: code and r, then x0 once per fine cell, then res0 once
// per fine cell of the supports, then each thread sums its cell's 16
// weighted res0.  Each value is formed by the same expression as it would
// be per coarse cell, so r1 does not depend on the tiling.
constexpr int kD0R = 2 * kD0Tile + 4;
__global__ void __launch_bounds__(kBlock) k_mg_down0(const SceneDev* scenes, const int2* map, int which)
{
    static_assert(kD0Tile * kD0Tile == kBlock, "one level-1 cell per thread");
    __shared__ float xs[kD0R][kD0R + 1];   // x0
    __shared__ float rs[kD0R][kD0R + 1];   // r, then res0 (own cell only)
    __shared__ uint8_t cs[kD0R][kD0R + 4];
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& L1 = cst(s.mg)->lv[1];
    const int I1 = L1.I, J1 = L1.J;
    const int tjn = (J1 + kD0Tile - 1) / kD0Tile;
    // one test after the uniform loads, so that they issue together; the
    // whole block takes it
    if (!mg_iterating(s) | (sb.block >= mg_d0_tiles(I1, J1))) return;
    const int ti = sb.block / tjn, tj = sb.block - ti * tjn;
    const int I = s.I, J = s.J;
    const int fi0 = 2 * kD0Tile * ti - 2, fj0 = 2 * kD0Tile * tj - 2;   // the patch's first fine cell
    gptr<const float> r = gbl<const float>(which ? s.cr2 : s.cr);
    gptr<const uint8_t> code = gbl<const uint8_t>(s.stencil);
    {   // code, r and x0 of the patch (outside the grid: code 0, r 0)
        constexpr int Q = (kD0R * kD0R + kBlock - 1) / kBlock;
        uint8_t kq[Q];
        float vq[Q];
#pragma unroll
        for (int u = 0; u < Q; u++) {
            const int q = min((int)threadIdx.x + u * kBlock, kD0R * kD0R - 1);
            const int fi = fi0 + q / kD0R, fj = fj0 + q % kD0R;
            const int f = min(max(fi, 0), I - 1) * J + min(max(fj, 0), J - 1);   // clamped, not selected
            kq[u] = code[f];
            vq[u] = r[f];
        }
#pragma unroll
        for (int u = 0; u < Q; u++) {
            const int q = threadIdx.x + u * kBlock;
            if (q >= kD0R * kD0R) break;
            const int a = q / kD0R, b = q % kD0R;
            const int fi = fi0 + a, fj = fj0 + b;
            const bool in = fi >= 0 && fi < I && fj >= 0 && fj < J;
            const uint8_t k = in ? kq[u] : 0;
            const float rv = vq[u] * mask01(in);
            const float d = l1_diag0(k);
            xs[a][b] = ((k & 0x80) && d > 0.f) ? kMgOmega * rv / d : 0.f;
            rs[a][b] = rv;
            cs[a][b] = k;
        }
    }
    __syncthreads();
    // res0 on the supports (patch cells 1 .. kD0R-2 each way); 0 off fluid rows
    for (int q = threadIdx.x; q < (kD0R - 2) * (kD0R - 2); q += kBlock) {
        const int a = 1 + q / (kD0R - 2), b = 1 + q % (kD0R - 2);
        const uint8_t k = cs[a][b];
        float ax = (float)((k >> 4) & 7) * xs[a][b];
        ax -= (k & 1) ? xs[a - 1][b] : 0.f;
        ax -= (k & 2) ? xs[a + 1][b] : 0.f;
        ax -= (k & 4) ? xs[a][b - 1] : 0.f;
        ax -= (k & 8) ? xs[a][b + 1] : 0.f;
        rs[a][b] = (k & 0x80) ? rs[a][b] - ax : 0.f;
    }
    __syncthreads();
    const int li = threadIdx.x / kD0Tile, lj = threadIdx.x % kD0Tile;
    const int ci = ti * kD0Tile + li, cj = tj * kD0Tile + lj;
    if ((ci >= I1) | (cj >= J1)) return;
    const int c = ci * J1 + cj;
    if (!gbl<const uint8_t>(L1.act)[c]) {   // inactive coarse cell: nothing to restrict
        gbl(L1.r)[c] = 0.f;
        return;
    }
    float acc = 0.f;
#pragma unroll
    for (int a = 1; a < 5; a++)
#pragma unroll
        for (int b = 1; b < 5; b++) {
            const uint8_t k = cs[2 * li + a][2 * lj + b];
            acc += (k & 0x80) ? (w1(a - 1) * w1(b - 1)) * rs[2 * li + a][2 * lj + b] : 0.f;
        }
    gbl(L1.r)[c] = 0.25f * acc;
}

// Level l (1 <= l < t): x = r * dinv, y = res = r - A_l x.
__global__ void __launch_bounds__(kBlock) k_mg_resid(const SceneDev* scenes, const int2* map, int l)
{
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& M = *cst(s.mg);
    const auto& L = M.lv[l];
    const int LI = L.I, LJ = L.J, CL = LI * LJ;
    const int c = sb.block * kBlock + threadIdx.x;
    if (!mg_iterating(s) | (l >= M.t) | (c >= CL)) return;   // l >= t: in k_mg_coarse
    gptr<const float> r = gbl<const float>(L.r), dinv = gbl<const float>(L.dinv);
    const float dc = dinv[c];
    if (dc == 0.f) {   // inactive row: r = 0 and no coefficients
        gbl(L.y)[c] = 0.f;
        return;
    }
    const int ci = c / LJ, cj = c - ci * LJ;
    float xv[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
        const bool ok = ni >= 0 && ni < LI && nj >= 0 && nj < LJ;
        const int n = min(max(ni, 0), LI - 1) * LJ + min(max(nj, 0), LJ - 1);   // clamped, not selected
        const float rn = r[n], dn = dinv[n];
        xv[k] = (rn * fabsf(dn)) * mask01(ok);   // a 0/1 factor, not a select: the load stays unconditional
    }
    const float rc = r[c];
    const float ax = row_dot(M, l, gbl<const float>(L.coef), CL, c, dc, xv);
    gbl(L.y)[c] = rc - ax;
}

// Level l-1 -> l (2 <= l < t): r_l = R y_{l-1}.
__global__ void __launch_bounds__(kBlock) k_mg_restrict(const SceneDev* scenes, const int2* map, int l)
{
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& M = *cst(s.mg);
    const auto& L = M.lv[l];
    const auto& F = M.lv[l - 1];
    const int LJ = L.J, FI = F.I, FJ = F.J;
    const int c = sb.block * kBlock + threadIdx.x;
    if (!mg_iterating(s) | (l >= M.t) | (c >= L.I * LJ)) return;   // l >= t: in k_mg_coarse
    gptr<const float> fy = gbl<const float>(F.y);
    if (!gbl<const uint8_t>(L.act)[c]) {
        gbl(L.r)[c] = 0.f;
        return;
    }
    const int ci = c / LJ, cj = c - ci * LJ;
    gbl(L.r)[c] = restrict_at(fy, FI, FJ, ci, cj);
}

// Levels t .. n-1 of one scene in one block (t = MgLevels::t, chosen so
// that level t has at most kMgCoarseTop = kMgCoarseThreads cells and levels
// t .. n-1 at most kMgCoarseCells together): r_t = R y_{t-1} (r_1 comes
// from k_mg_down0 when t = 1), the down steps of k_mg_resid and
// k_mg_restrict to the bottom, kMgBottomSweeps L1-Jacobi sweeps from zero
// there, then the up steps of k_mg_prolong and k_mg_post back to level t,
// whose y it writes for the level above.  The arithmetic and its order are
// those of the per-level kernels, so where t falls does not change z.
// Thread c owns cell c of every level.  The block is latency-bound, so its
// global loads come in two rounds: every level's dinv with level t's input
// (its 4x4 patch of y_{t-1}), then the stored coefficient rows of the
// thread's level-t and bottom cells, held in registers for the way down
// and the way up.  LDS: every level's r and dinv (a cell is active exactly
// where its dinv's bit pattern is nonzero) and one work array for the level
// at hand: its residual y on the way down (dead once the coarser r is
// made), its x on the way up.  On the way up each level's y overwrites its
// r in place (a thread reads only its own cell's r there), where the finer
// level reads it.  On the way down each level's x = r dinv is made once
// into its own array before the residual reads 25 of them.  The bottom's
// iterate has its own array with a two-cell zero halo, so its stencil needs
// no bounds tests.
__device__ __forceinline__ bool mg_active(float dinv) { return __float_as_uint(dinv) != 0u; }

// The 25 stored coefficients of row c, loaded together.
__device__ __forceinline__ void coarse_row_load(gptr<const float> coef, int CL, int c, float (&a)[kMgStencil])
{
    (void)CL;
    load_coef_row(coef, c, a);
}

// row_dot with v_k = V(k) made as it is summed (same terms, same order):
// the canonical stencil (LDS) on canonical rows (dinv < 0), else the
// coefficients in a.
template <typename F>
__device__ __forceinline__ float coarse_row_dot(const float* cn, const float (&a)[kMgStencil], float dinv_c, F V)
{
    const bool stored = dinv_c > 0.f;
    float ax = 0.f;
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) ax += (stored ? a[k] : cn[k]) * V(k);
    return ax;
}

__global__ void __launch_bounds__(kMgCoarseThreads) k_mg_coarse(const SceneDev* scenes)
{
    constexpr int NT = kMgCoarseThreads;
    constexpr int W = kMgBottomMax + 4;
    static_assert(kMgBottomMax * kMgBottomMax <= NT, "one bottom cell per thread");
    static_assert(kMgCoarseTop <= NT, "one level-t cell per thread");
    __shared__ float rs[kMgCoarseCells], ds[kMgCoarseCells];
    __shared__ float tx[kMgCoarseTop], xw[kMgCoarseTop];
    __shared__ float xb[2][W * W];
    __shared__ float cns[kMgMaxLevels][kMgStencil];
    constexpr int kBS = kMgBottomMax * kMgBottomMax + 1;   // the bottom's entry stride in LDS (max)
    __shared__ float cbs[kMgStencil * kBS];   // bottom coefficients, entry-major
    const SceneDev& s = scenes[blockIdx.y];
    const auto& M = *cst(s.mg);
    const int t = M.t, nb = M.n - 1;
    if (!mg_iterating(s)) return;
    const int tid = threadIdx.x;
    if (tid < kMgMaxLevels * kMgStencil) (&cns[0][0])[tid] = (&M.canon[0][0])[tid];
    const auto& T = M.lv[t];
    const int TI = T.I, TJ = T.J, CT = TI * TJ;   // roff[t] = 0
    const bool tin = tid < CT;
    const int tc = tin ? tid : CT - 1;
    const int tci = tc / TJ, tcj = tc - tci * TJ;
    // round 1: dinv of levels t .. n-1 (contiguous) and level t's input
    {
        gptr<const float> dg = gbl<const float>(T.dinv);
        const int TC = M.tcells;
        constexpr int Q = (kMgCoarseCells + NT - 1) / NT;
        float dv[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) dv[q] = dg[min(tid + q * NT, TC - 1)];
        float rt;
        if (t > 1) {   // r_t = R y_{t-1} from the launch before
            const auto& F = M.lv[t - 1];
            rt = restrict_at(gbl<const float>(F.y), F.I, F.J, tci, tcj);
        } else {
            rt = gbl<const float>(M.lv[1].r)[tc];
        }
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (tid + q * NT < TC) ds[tid + q * NT] = dv[q];
        if (tin) rs[tc] = (t == 1 || mg_active(dv[0])) ? rt : 0.f;   // dv[0] = level t's dinv of cell tc
    }
    __syncthreads();
    // round 2: the coefficients of this thread's level-t row where it reads
    // stored ones (level t above the bottom), and the bottom's, whose sweeps
    // read the stored coefficients of every row
    const auto& B = M.lv[nb];
    const int CB = B.I * B.J, ob = M.roff[nb];   // CB <= kMgBottomMax^2 (the hierarchy stops there)
    const bool bin = tid < CB;
    const int bc = bin ? tid : 0;
    const float dct = tin ? ds[tc] : 0.f;
    float ct[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) ct[k] = 0.f;
    if (tin && nb != t && dct > 0.f) coarse_row_load(gbl<const float>(T.coef), CT, tc, ct);
    // the bottom's entry stride in LDS: a multiple of the 32 banks plus one,
    // so that the copy below (consecutive threads: consecutive entries of a
    // row) and a sweep (consecutive cells: one entry) are both conflict-free
    const int SB = ((CB + 31) & ~31) + 1;
    {   // the bottom's coefficient rows (contiguous) to LDS by the whole block,
        // transposed to entry-major there
        gptr<const float> bcoef = gbl<const float>(B.coef);
        const int NBC = kMgRowStride * CB;
        constexpr int QB = (kMgRowStride * kMgBottomMax * kMgBottomMax + NT - 1) / NT;
        float v[QB];
#pragma unroll
        for (int q = 0; q < QB; q++) v[q] = bcoef[min(tid + q * NT, NBC - 1)];
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int i = tid + q * NT, c = i / kMgRowStride, k = i % kMgRowStride;
            if (i < NBC && k < kMgStencil) cbs[k * SB + c] = v[q];
        }
    }
    for (int l = t; l < nb; l++) {   // down: x = r dinv, y = r - A x, r_{l+1} = R y
        const auto& L = M.lv[l];
        const int LI = L.I, LJ = L.J, CL = LI * LJ, o = M.roff[l];
        if (tid < CL) xw[tid] = rs[o + tid] * fabsf(ds[o + tid]);   // x = r dinv
        __syncthreads();
        if (tid < CL) {
            const int c = tid;
            const float dc = ds[o + c];
            float y = 0.f;   // inactive row: r = 0 and no coefficients
            if (dc != 0.f) {
                const int ci = c / LJ, cj = c - ci * LJ;
                auto V = [&](int k) {
                    const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
                    const bool ok = ni >= 0 && ni < LI && nj >= 0 && nj < LJ;
                    return xw[min(max(ni, 0), LI - 1) * LJ + min(max(nj, 0), LJ - 1)] * mask01(ok);
                };
                float ax;
                if (l == t) {
                    ax = coarse_row_dot(cns[l], ct, dc, V);
                } else {
                    float a[kMgStencil];
#pragma unroll
                    for (int k = 0; k < kMgStencil; k++) a[k] = 0.f;
                    if (dc > 0.f) coarse_row_load(gbl<const float>(L.coef), CL, c, a);
                    ax = coarse_row_dot(cns[l], a, dc, V);
                }
                y = rs[o + c] - ax;
            }
            tx[c] = y;
        }
        __syncthreads();
        const auto& N = M.lv[l + 1];
        const int NJ = N.J, NC = N.I * NJ, on = M.roff[l + 1];
        if (tid < NC) {
            const int c = tid, ci = c / NJ, cj = c - ci * NJ;
            rs[on + c] = mg_active(ds[on + c]) ? restrict_at(static_cast<const float*>(tx), LI, LJ, ci, cj) : 0.f;
        }
        __syncthreads();
    }
    {   // bottom: y = kMgBottomSweeps sweeps of x += dinv (r - A x) from 0
        const int ci = bc / B.J, cj = bc - ci * B.J;
        const float dv = fabsf(ds[ob + bc]), rv = rs[ob + bc];
        for (int k = tid; k < 2 * W * W; k += NT) (&xb[0][0])[k] = 0.f;
        __syncthreads();
        const int home = (ci + 2) * W + (cj + 2);
        int cur = 0;
        for (int sw = 0; sw < kMgBottomSweeps; sw++) {
            if (bin) {
                float ax = 0.f;
#pragma unroll
                for (int k = 0; k < kMgStencil; k++)
                    ax += cbs[k * SB + bc] * xb[cur][home + (k / 5 - 2) * W + (k % 5 - 2)];
                xb[cur ^ 1][home] = xb[cur][home] + dv * (rv - ax);
            }
            __syncthreads();
            cur ^= 1;
        }
        if (bin) {
            const float y = xb[cur][home];
            rs[ob + bc] = y;
            if (nb == t) gbl(B.y)[bc] = y;
        }
        __syncthreads();
    }
    for (int l = nb - 1; l >= t; l--) {   // up: x = r dinv + P y_{l+1}, y = x + dinv (r - A x)
        const auto& L = M.lv[l];
        const auto& N = M.lv[l + 1];
        const int LI = L.I, LJ = L.J, CL = LI * LJ, o = M.roff[l];
        const float* yn = rs + M.roff[l + 1];   // y_{l+1}, in place of r_{l+1}
        if (tid < CL) {
            const int c = tid;
            const float dc = ds[o + c];
            float x = 0.f;
            if (mg_active(dc)) {
                const int ci = c / LJ, cj = c - ci * LJ;
                x = rs[o + c] * fabsf(dc) + prolong_at(yn, N.I, N.J, ci, cj);
            }
            tx[c] = x;
        }
        __syncthreads();
        if (tid < CL) {
            const int c = tid;
            const float dc = ds[o + c];
            float y = 0.f;   // inactive row: x = 0
            if (dc != 0.f) {
                const int ci = c / LJ, cj = c - ci * LJ;
                auto V = [&](int k) {
                    const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
                    const bool ok = ni >= 0 && ni < LI && nj >= 0 && nj < LJ;
                    return tx[min(max(ni, 0), LI - 1) * LJ + min(max(nj, 0), LJ - 1)] * mask01(ok);
                };
                float ax;
                if (l == t) {
                    ax = coarse_row_dot(cns[l], ct, dc, V);
                } else {
                    float a[kMgStencil];
#pragma unroll
                    for (int k = 0; k < kMgStencil; k++) a[k] = 0.f;
                    if (dc > 0.f) coarse_row_load(gbl<const float>(L.coef), CL, c, a);
                    ax = coarse_row_dot(cns[l], a, dc, V);
                }
                y = tx[c] + fabsf(dc) * (rs[o + c] - ax);   // tx[c] = the centre term x_12
            }
            rs[o + c] = y;
            if (l == t) gbl(L.y)[c] = y;
        }
        __syncthreads();
    }
}

// The fused V-cycle levels (the default; the per-level kernels above and
// below stay for comparison, fg_ctx_set_vcycle_fused).  A block owns a
// kTile x kTile tile of cells and stages what its cells' stencils reach in
// LDS, so a level's two steps take one launch instead of two.  Every value
// is formed by the expression the per-level kernels use, at the same clamped
// coordinates and with the same 0/1 masks, so z keeps its bits.
static_assert(kVcTile * kVcTile == kBlock, "one cell of the tile per thread");
constexpr int kRRY = 2 * kRRTile + 2;   // fine cells a side whose residual the restriction reads
constexpr int kRRX = kRRY + 4;          // and the x their stencils read
static_assert(4 * kRRTile * kRRTile <= kBlock, "one fine cell of the tile per thread");
constexpr int kPPX = kVcTile + 4;       // level-l cells a side whose x the post-smoothing reads

// Level l (1 <= l < t), down: x = r dinv and y = r - A_l x over the fine
// patch of a kRRTile x kRRTile tile of level-(l+1) cells, then r_{l+1} = R y
// for the tile's cells (l + 1 < t) or y written out for k_mg_coarse to
// restrict (l + 1 = t).  The tile is smaller than k_mg_pp's: each fine
// residual is a 25-term row with its own coefficient loads, and the small
// tile keeps those to one or two rounds a thread.
// Positions outside level l hold the values of their clamped cells (the
// per-level kernels load those and mask them).
__global__ void __launch_bounds__(kBlock) k_mg_rr(const SceneDev* scenes, const int2* map, int l)
{
    __shared__ float xs[kRRX][kRRX + 1];
    __shared__ float ys[kRRY][kRRY + 1];
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& M = *cst(s.mg);
    const auto& L = M.lv[l];
    const auto& N = M.lv[l + 1];
    const int LI = L.I, LJ = L.J, CL = LI * LJ, NI = N.I, NJ = N.J;
    const int tjn = (NJ + kRRTile - 1) / kRRTile;
    if (!mg_iterating(s) | (l + 1 > M.t) | (sb.block >= ((NI + kRRTile - 1) / kRRTile) * tjn)) return;
    const int ti = sb.block / tjn, tj = sb.block - ti * tjn;
    const int yi0 = 2 * kRRTile * ti - 1, yj0 = 2 * kRRTile * tj - 1;   // the residual patch's first fine cell
    const int xi0 = yi0 - 2, xj0 = yj0 - 2;
    gptr<const float> r = gbl<const float>(L.r), dinv = gbl<const float>(L.dinv);
    {   // x at the clamped cell of every patch position
        constexpr int Q = (kRRX * kRRX + kBlock - 1) / kBlock;
        float rq[Q], dq[Q];
#pragma unroll
        for (int u = 0; u < Q; u++) {
            const int q = min((int)threadIdx.x + u * kBlock, kRRX * kRRX - 1);
            const int fi = min(max(xi0 + q / kRRX, 0), LI - 1), fj = min(max(xj0 + q % kRRX, 0), LJ - 1);
            rq[u] = r[fi * LJ + fj];
            dq[u] = dinv[fi * LJ + fj];
        }
#pragma unroll
        for (int u = 0; u < Q; u++) {
            const int q = threadIdx.x + u * kBlock;
            if (q < kRRX * kRRX) xs[q / kRRX][q % kRRX] = rq[u] * fabsf(dq[u]);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kRRY * kRRY; q += kBlock) {   // y at the clamped cell of every position
        const int ci = min(max(yi0 + q / kRRY, 0), LI - 1), cj = min(max(yj0 + q % kRRY, 0), LJ - 1);
        const int c = ci * LJ + cj;
        const float dc = dinv[c];
        float y = 0.f;   // inactive row: r = 0 and no coefficients
        if (dc != 0.f) {
            float xv[kMgStencil];
#pragma unroll
            for (int k = 0; k < kMgStencil; k++) {
                const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
                const bool ok = ni >= 0 && ni < LI && nj >= 0 && nj < LJ;
                xv[k] = xs[ni - xi0][nj - xj0] * mask01(ok);
            }
            const float rc = r[c];
            const float ax = row_dot(M, l, gbl<const float>(L.coef), CL, c, dc, xv);
            y = rc - ax;
        }
        ys[q / kRRY][q % kRRY] = y;
    }
    __syncthreads();
    if (l + 1 < M.t) {   // r_{l+1} = R y (restrict_at on the patch)
        if ((int)threadIdx.x >= kRRTile * kRRTile) return;
        const int li = threadIdx.x / kRRTile, lj = threadIdx.x % kRRTile;
        const int ci = kRRTile * ti + li, cj = kRRTile * tj + lj;
        if ((ci >= NI) | (cj >= NJ)) return;
        const int c = ci * NJ + cj;
        if (!gbl<const uint8_t>(N.act)[c]) {
            gbl(N.r)[c] = 0.f;
            return;
        }
        float acc = 0.f;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int fi = 2 * ci - 1 + a, fj = 2 * cj - 1 + b;
                const bool ok = fi >= 0 && fi < LI && fj >= 0 && fj < LJ;
                acc += ((w1(a) * w1(b)) * ys[2 * li + a][2 * lj + b]) * mask01(ok);
            }
        gbl(N.r)[c] = 0.25f * acc;
    } else {   // level t's input: y of the tile's own fine cells
        const int q = threadIdx.x;
        if (q >= 4 * kRRTile * kRRTile) return;
        const int a = q / (2 * kRRTile), b = q % (2 * kRRTile);
        const int fi = 2 * kRRTile * ti + a, fj = 2 * kRRTile * tj + b;
        if ((fi < LI) & (fj < LJ)) gbl(L.y)[fi * LJ + fj] = ys[a + 1][b + 1];
    }
}

// Level l (1 <= l < t), up: x = r dinv + P y_{l+1} over a tile of level-l
// cells and its two-cell halo, then y = x + dinv (r - A_l x) for the tile.
__global__ void __launch_bounds__(kBlock) k_mg_pp(const SceneDev* scenes, const int2* map, int l)
{
    __shared__ float xs[kPPX][kPPX + 1];
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& M = *cst(s.mg);
    const auto& L = M.lv[l];
    const auto& Cl = M.lv[l + 1];
    const int LI = L.I, LJ = L.J, CL = LI * LJ;
    const int tjn = (LJ + kVcTile - 1) / kVcTile;
    if (!mg_iterating(s) | (l >= M.t) | (sb.block >= ((LI + kVcTile - 1) / kVcTile) * tjn)) return;
    const int ti = sb.block / tjn, tj = sb.block - ti * tjn;
    const int xi0 = kVcTile * ti - 2, xj0 = kVcTile * tj - 2;
    gptr<const float> r = gbl<const float>(L.r), dinv = gbl<const float>(L.dinv);
    gptr<const uint8_t> act = gbl<const uint8_t>(L.act);
    gptr<const float> yc = gbl<const float>(Cl.y);
    for (int q = threadIdx.x; q < kPPX * kPPX; q += kBlock) {   // x at the clamped cell of every position
        const int ci = min(max(xi0 + q / kPPX, 0), LI - 1), cj = min(max(xj0 + q % kPPX, 0), LJ - 1);
        const int c = ci * LJ + cj;
        const uint8_t a = act[c];
        const float rc = r[c], dc = dinv[c];
        const float pc = prolong_at(yc, Cl.I, Cl.J, ci, cj);
        xs[q / kPPX][q % kPPX] = a ? rc * fabsf(dc) + pc : 0.f;
    }
    __syncthreads();
    const int li = threadIdx.x / kVcTile, lj = threadIdx.x % kVcTile;
    const int ci = kVcTile * ti + li, cj = kVcTile * tj + lj;
    if ((ci >= LI) | (cj >= LJ)) return;
    const int c = ci * LJ + cj;
    const float dc = dinv[c];
    if (dc == 0.f) {   // inactive row: x = 0
        gbl(L.y)[c] = 0.f;
        return;
    }
    float xv[kMgStencil];
#pragma unroll
    for (int k = 0; k < kMgStencil; k++) {
        const int ni = ci + k / 5 - 2, nj = cj + k % 5 - 2;
        const bool ok = ni >= 0 && ni < LI && nj >= 0 && nj < LJ;
        xv[k] = xs[li + k / 5][lj + k % 5] * mask01(ok);
    }
    const float rc = r[c];
    const float ax = row_dot(M, l, gbl<const float>(L.coef), CL, c, dc, xv);
    gbl(L.y)[c] = xv[12] + fabsf(dc) * (rc - ax);
}

// Level l (1 <= l < t), up: x = r * dinv + P y_{l+1}.
__global__ void __launch_bounds__(kBlock) k_mg_prolong(const SceneDev* scenes, const int2* map, int l)
{
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& M = *cst(s.mg);
    const auto& L = M.lv[l];
    const auto& Cl = M.lv[l + 1];
    const int LJ = L.J;
    const int c = sb.block * kBlock + threadIdx.x;
    if (!mg_iterating(s) | (l >= M.t) | (c >= L.I * LJ)) return;   // l >= t: in k_mg_coarse
    if (!gbl<const uint8_t>(L.act)[c]) {
        gbl(L.x)[c] = 0.f;
        return;
    }
    const int ci = c / LJ, cj = c - ci * LJ;
    const float rc = gbl<const float>(L.r)[c], dc = gbl<const float>(L.dinv)[c];
    const float pc = prolong_at(gbl<const float>(Cl.y), Cl.I, Cl.J, ci, cj);
    gbl(L.x)[c] = rc * fabsf(dc) + pc;
}

// Level l (1 <= l < t), up: y = x + dinv (r - A_l x).
__global__ void __launch_bounds__(kBlock) k_mg_post(const SceneDev* scenes, const int2* map, int l)
{
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& M = *cst(s.mg);
    const auto& L = M.lv[l];
    const int LI = L.I, LJ = L.J, CL = LI * LJ;
    const int c = sb.block * kBlock + threadIdx.x;
    if (!mg_iterating(s) | (l >= M.t) | (c >= CL)) return;   // l >= t: in k_mg_coarse
    gptr<const float> x = gbl<const float>(L.x);
    const float dc = gbl<const float>(L.dinv)[c];
    if (dc == 0.f) {   // inactive row: x = 0
        gbl(L.y)[c] = 0.f;
        return;
    }
    const int ci = c / LJ, cj = c - ci * LJ;
    float xv[kMgStencil];
    gather25(x, LI, LJ, ci, cj, xv);
    const float rc = gbl<const float>(L.r)[c];
    const float ax = row_dot(M, l, gbl<const float>(L.coef), CL, c, dc, xv);
    gbl(L.y)[c] = xv[12] + fabsf(dc) * (rc - ax);
}

// Level 0, up: z0 = w r / D1 + P y1 at the cell and its fluid neighbours, then
// z = z0 + w (r - A' z0) / D1; z = r on non-fluid rows.  Writes z and the
// partials of r.z and r.r the next PCG iteration reduces (solver chunks).
// z0 of a fine cell depends on that cell alone, so a block forms it once
// per cell into an LDS ring over the flat cell index (slot f mod RB, RB =
// 2J + 2 kUpPass floats of dynamic LDS) instead of once per neighbour.
// Pass p covers cells p .. p + kUpPass - 1, thread t taking p + t + u kBlock
// (u = 0 .. kUpItems - 1: the cells and their order of a plain kBlock
// stride, so the partials keep their order), and reads the ring over
// [p - J, p + J + kUpPass); the prologue fills the first pass's window,
// each pass the cells the next one adds (their loads issued first; their
// slots are those of the cells just behind this pass's window).  The
// window costs a prologue of 2J + kUpPass cells per block and a barrier per
// pass: it pays where chunks span several rows of a grid too large for the
// L2 to keep the neighbouring rows (4096^2); on one-row chunks (1024^2) and
// on batches of small scenes k_mg_up0_wide is faster, and the host
// launches that there.
constexpr int kUpItems = 4;
constexpr int kUpPass = kUpItems * kBlock;
constexpr int kUpMaxJ = (16384 - 2 * kUpPass) / 2;   // the ring within 64 KB
__global__ void __launch_bounds__(kBlock) k_mg_up0(const SceneDev* scenes, const int2* map, int which)
{
    extern __shared__ float zr[];
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& L1 = cst(s.mg)->lv[1];
    if (!mg_iterating(s) | (sb.block >= s.pcg_blocks)) return;
    const int I = s.I, J = s.J, C = I * J;
    gptr<const float> y1 = gbl<const float>(L1.y);
    const int I1 = L1.I, J1 = L1.J;
    gptr<const float> r = gbl<const float>(which ? s.cr2 : s.cr);
    gptr<const uint8_t> code = gbl<const uint8_t>(s.stencil);
    gptr<float> z = gbl(s.cu);
    const Chunk ch = my_chunk(s, sb.block);
    const int RB = 2 * J + 2 * kUpPass;
    // z0 of cell f from its code, r and P y1 at it
    auto z0_of = [&](uint8_t kf, float rf, float pg) {
        const float dg = l1_diag0(kf);
        return (dg > 0.f ? kMgOmega * rf / dg : 0.f) + pg;
    };
    {   // prologue: the first pass's window
        const int lo = max(ch.begin - J, 0), hi = min(ch.begin + J + kUpPass, C);
        for (int f0 = lo; f0 < hi; f0 += kUpPass) {
            uint8_t kq[kUpItems];
            float rq[kUpItems], pq[kUpItems];
#pragma unroll
            for (int u = 0; u < kUpItems; u++) {
                const int f = min(f0 + (int)threadIdx.x + u * kBlock, hi - 1);   // clamped, not selected
                const int fi = f / J, fj = f - fi * J;
                kq[u] = code[f];
                rq[u] = r[f];
                pq[u] = prolong_at(y1, I1, J1, fi, fj);
            }
#pragma unroll
            for (int u = 0; u < kUpItems; u++) {
                const int f = f0 + threadIdx.x + u * kBlock;
                if (f < hi) zr[f % RB] = z0_of(kq[u], rq[u], pq[u]);
            }
        }
    }
    __syncthreads();
    const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
    double rz = 0.0, rr = 0.0;
    int sp = ch.begin % RB;   // the ring slot of cell p
    for (int p = ch.begin; p < ch.end; p += kUpPass) {
        // the cells this thread adds for the next pass, and its own cells'
        // code and r: loads first
        const bool more = p + kUpPass < ch.end;
        uint8_t kf[kUpItems], kc[kUpItems];
        float rf[kUpItems], pf[kUpItems], rcv[kUpItems];
#pragma unroll
        for (int u = 0; u < kUpItems; u++) {
            const int fc = min(p + J + kUpPass + (int)threadIdx.x + u * kBlock, C - 1);
            const int fi = fc / J, fj = fc - fi * J;
            kf[u] = code[fc];
            rf[u] = r[fc];
            pf[u] = prolong_at(y1, I1, J1, fi, fj);
            const int c = min(p + (int)threadIdx.x + u * kBlock, ch.end - 1);
            kc[u] = code[c];
            rcv[u] = r[c];
        }
#pragma unroll
        for (int u = 0; u < kUpItems; u++) {
            const int c = p + threadIdx.x + u * kBlock;
            if (c >= ch.end) break;
            const uint8_t k = kc[u];
            const float rc = rcv[u];
            float zc = rc;
            if (k & 0x80) {
                const int ci = c / J, cj = c - ci * J;
                // the fluid neighbours that enter A' (not the row-wrapping coupling)
                bool use[5];
                use[0] = true;
#pragma unroll
                for (int e = 1; e < 5; e++) use[e] = (k & (1 << (e - 1))) != 0;
                use[3] = use[3] && cj > 0;
                use[4] = use[4] && cj < J - 1;
                int sc0 = sp + (int)threadIdx.x + u * kBlock;
                sc0 -= sc0 >= RB ? RB : 0;
                float z0[5];
#pragma unroll
                for (int e = 0; e < 5; e++) {
                    const int gi = min(max(ci + di[e], 0), I - 1), gj = min(max(cj + dj[e], 0), J - 1);
                    int sg = sc0 + (gi - ci) * J + (gj - cj);
                    sg += sg < 0 ? RB : 0;
                    sg -= sg >= RB ? RB : 0;
                    z0[e] = use[e] ? zr[sg] : 0.f;
                }
                const float dc = l1_diag0(k);
                float ax = (float)((k >> 4) & 7) * z0[0];
#pragma unroll
                for (int e = 1; e < 5; e++) ax -= z0[e];
                zc = ((k & 0x80) && dc > 0.f) ? z0[0] + kMgOmega * (rc - ax) / dc : rc;
                rz += (double)rc * (double)zc;
            } else {   // identity row: z = r
                rz += (double)rc * (double)rc;
            }
            z[c] = zc;
            rr += (double)rc * (double)rc;
        }
#pragma unroll
        for (int u = 0; u < kUpItems; u++) {
            const int f = p + J + kUpPass + threadIdx.x + u * kBlock;
            if (more && f < C) zr[f % RB] = z0_of(kf[u], rf[u], pf[u]);
        }
        __syncthreads();
        sp += kUpPass;
        sp -= sp >= RB ? RB : 0;
    }
    double v[2] = {rz, rr};
    put_partials<2>(s, kSlotCG, v, sb.block);
}

// k_mg_up0 for grids wider than its LDS window allows: each cell forms the
// smoothing start z0 of its fluid neighbours itself (the same values).
__global__ void __launch_bounds__(kBlock) k_mg_up0_wide(const SceneDev* scenes, const int2* map, int which)
{
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    const auto& L1 = cst(s.mg)->lv[1];
    if (!mg_iterating(s) | (sb.block >= s.pcg_blocks)) return;
    const int J = s.J;
    gptr<const float> y1 = gbl<const float>(L1.y);
    const int I1 = L1.I, J1 = L1.J;
    gptr<const float> r = gbl<const float>(which ? s.cr2 : s.cr);
    gptr<const uint8_t> code = gbl<const uint8_t>(s.stencil);
    gptr<float> z = gbl(s.cu);
    const Chunk ch = my_chunk(s, sb.block);
    double rz = 0.0, rr = 0.0;
    for (int c = ch.begin + threadIdx.x; c < ch.end; c += kBlock) {
        const uint8_t k = code[c];
        const float rc = r[c];
        if (!(k & 0x80)) {   // identity row: z = r, no neighbourhood
            z[c] = rc;
            rz += (double)rc * (double)rc;
            rr += (double)rc * (double)rc;
            continue;
        }
        const int ci = c / J, cj = c - ci * J;
        const int di[5] = {0, -1, 1, 0, 0}, dj[5] = {0, 0, 0, -1, 1};
        // the fluid neighbours that enter A' (not the row-wrapping coupling)
        bool use[5];
        use[0] = true;
#pragma unroll
        for (int e = 1; e < 5; e++) use[e] = (k & (1 << (e - 1))) != 0;
        use[3] = use[3] && cj > 0;
        use[4] = use[4] && cj < J - 1;
        uint8_t kg[5];
        float rg[5], pg[5];
        const int I = s.I;
#pragma unroll
        for (int e = 0; e < 5; e++) {
            // clamped coordinates (not a select of c), so every load issues
            const int gi = min(max(ci + di[e], 0), I - 1), gj = min(max(cj + dj[e], 0), J - 1);
            kg[e] = code[gi * J + gj];
            rg[e] = r[gi * J + gj];
            pg[e] = prolong_at(y1, I1, J1, gi, gj);
        }
        float z0[5];
#pragma unroll
        for (int e = 0; e < 5; e++) {
            const float dg = l1_diag0(kg[e]);
            z0[e] = use[e] ? (dg > 0.f ? kMgOmega * rg[e] / dg : 0.f) + pg[e] : 0.f;
        }
        const float dc = l1_diag0(k);
        float ax = (float)((k >> 4) & 7) * z0[0];
#pragma unroll
        for (int e = 1; e < 5; e++) ax -= z0[e];
        const float zc = ((k & 0x80) && dc > 0.f) ? z0[0] + kMgOmega * (rc - ax) / dc : rc;
        z[c] = zc;
        rz += (double)rc * (double)zc;
        rr += (double)rc * (double)rc;
    }
    double v[2] = {rz, rr};
    put_partials<2>(s, kSlotCG, v, sb.block);
}

}  // namespace fg

namespace fg {
#define FG_MG_LAUNCH1(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, T a) { fg_launch(k_##name, g, dim3(kBlock), st, sc, a); }
#define FG_MG_LAUNCH_MAP(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, const int2* map, T a) \
    { \
        fg_launch(k_##name, g, dim3(kBlock), st, sc, map, a); \
    }
FG_MG_LAUNCH1(mg_mark, int, l)
FG_MG_LAUNCH1(mg_galerkin, int, l)
FG_MG_LAUNCH1(mg_sym, int, l)
void launch_mg_symmark(dim3 g, hipStream_t st, const SceneDev* sc, int l, int nsym)
{
    fg_launch(k_mg_symmark, g, dim3(kBlock), st, sc, l, nsym);
}
FG_MG_LAUNCH_MAP(mg_down0, int, which)
FG_MG_LAUNCH_MAP(mg_resid, int, l)
FG_MG_LAUNCH_MAP(mg_restrict, int, l)
FG_MG_LAUNCH_MAP(mg_prolong, int, l)
FG_MG_LAUNCH_MAP(mg_post, int, l)
FG_MG_LAUNCH_MAP(mg_rr, int, l)
FG_MG_LAUNCH_MAP(mg_pp, int, l)
void launch_mg_up0(dim3 g, hipStream_t st, const SceneDev* sc, const int2* map, int which, int max_j, bool ring)
{
    if (ring && max_j <= kUpMaxJ)
        fg_launch_shm(k_mg_up0, g, dim3(kBlock), sizeof(float) * (2 * max_j + 2 * kUpPass), st, sc, map, which);
    else
        fg_launch(k_mg_up0_wide, g, dim3(kBlock), st, sc, map, which);
}
void launch_mg_coarse(dim3 g, hipStream_t st, const SceneDev* sc)
{
    fg_launch(k_mg_coarse, g, dim3(kMgCoarseThreads), st, sc);
}
}  // namespace fg
