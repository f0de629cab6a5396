// fg_pcg.hip -- pressure projection solve.
//
// System: the reference's matrix-free 5-point operator
// (IndexedPressureParameterUnit::multiply, pressuredata.h:132-146):
//   fluid row c:  (A x)_c = s * (n_c x_c - sum_{fluid nbrs} x_k),  s = dt/(rho dx^2)
//   other rows:   (A x)_c = x_c
// solved for the FP64 relative residual ||b - A x||2 / ||b||2 <= tol by
// iterative refinement: each sweep forms r64 = b - A x in FP64, solves the
// correction A' d = r' with FP32 PCG (A' = A/s on fluid rows, identity
// elsewhere, so the stencil has integer coefficients), and adds d into x in
// FP64.  Dot products are FP64; each is reduced per block and the block
// partials are summed in block order by the last block to finish, so every
// result is deterministic.
//
// Preconditioner: the symmetric inverse-Poisson preconditioner the
// reference defines in inversepoissonpreconditioner.h:45-79,
// M = (I - L D^-1)(I - L D^-1)^T on A's 5-point pattern (no fill-in):
//   z_c = (1 + sum_lower (1/n_k)^2) r_c + sum_lower r_k/n_k + sum_upper r_u/n_c
// ("lower" = smaller linear index: i-1 and j-1).  The reference's own
// matrix-free IPP (PressureIPPCoeficients.h:23-40) is mis-scaled and not
// symmetric and does not converge; it is reproduced only in the oracle.
//
// Each PCG iteration is two kernels, each a single pass over the grid:
//   KA: p = z + beta p_old (recomputed at the 4 neighbours), q = A' p, p.q
//   KB: r = r_old - alpha q (recomputed at the neighbours), d += alpha p,
//       z = M r, r.z and r.r
// so no vector is re-read by a separate axpy kernel.
#include "fg_kernels.h"

namespace fg {

__device__ __forceinline__ float inv_n(uint8_t code)
{
    const int n = (code >> 4) & 7;
    return n == 1 ? 1.0f : n == 2 ? 0.5f : n == 3 ? (1.0f / 3.0f) : n == 4 ? 0.25f : 0.0f;
}

// x64 += d; d = 0 (start of a refinement sweep)
__global__ void __launch_bounds__(kBlock) k_refine_update(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.sc->outer_done) return;
    const int C = s.I * s.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    s.x64[c] = s.x64[c] + (double)s.d[c];
    s.d[c] = 0.f;
}

// r64 = b - A x64 in FP64; inner rhs r = r64/s on fluid rows; outer test.
__global__ void __launch_bounds__(kBlock) k_refine_residual(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.sc->outer_done) return;
    const int J = s.J, C = s.I * J;
    const int nb = (C + kBlock - 1) / kBlock;
    if ((int)blockIdx.x >= nb) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double rr = 0.0;
    if (c < C) {
        const uint8_t code = s.stencil[c];
        const double* x = s.x64;
        double ax;
        if (code & 0x80) {
            const double sc = s.sA;
            const int n = (code >> 4) & 7;
            ax = (sc * n) * x[c];
            if (code & 1) ax += (-sc) * x[c - J];
            if (code & 2) ax += (-sc) * x[c + J];
            if (code & 4) ax += (-sc) * x[c - 1];
            if (code & 8) ax += (-sc) * x[c + 1];
        } else {
            ax = x[c];
        }
        const double r64 = s.b[c] - ax;
        rr = r64 * r64;
        s.r[c] = (float)((code & 0x80) ? r64 / s.sA : r64);
    }
    double v[1] = {rr};
    reduce_last<1>(s, v, 4, nb, [&](const double* t) {
        SceneScalars* sc = s.sc;
        sc->rr64 = t[0];
        const double bn = sqrt(sc->bb), rn = sqrt(t[0]);
        if (!finite_d(t[0])) {
            sc->numeric_bad = 1;
            sc->outer_done = 1;
            return;
        }
        if (rn <= s.tol * bn || bn == 0.0 || sc->sweeps >= kMaxSweeps) {
            sc->outer_done = 1;
            return;
        }
        sc->sweeps += 1;
        sc->inner_done = 0;
        sc->inner_iters = 0;
        // ask the FP32 solve for the reduction still needed (with margin)
        double f = 0.3 * s.tol * bn / rn;
        sc->inner_eps2 = (f < 1e-6 ? 1e-6 : f) * (f < 1e-6 ? 1e-6 : f);
    });
}

// z = M r, p-direction reset (beta = 0), r.z and r.r
__global__ void __launch_bounds__(kBlock) k_inner_init(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.sc->outer_done) return;
    const int J = s.J, C = s.I * J;
    const int nb = (C + kBlock - 1) / kBlock;
    if ((int)blockIdx.x >= nb) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (c < C) {
        const uint8_t code = s.stencil[c];
        const float* r = s.r;
        const float rc = r[c];
        float z = rc;
        if (code & 0x80) {
            float diag = 1.0f, off = 0.0f;
            if (code & 1) {
                const float w = inv_n(s.stencil[c - J]);
                diag += w * w;
                off += w * r[c - J];
            }
            if (code & 4) {
                const float w = inv_n(s.stencil[c - 1]);
                diag += w * w;
                off += w * r[c - 1];
            }
            const float wc = inv_n(code);
            if (code & 2) off += wc * r[c + J];
            if (code & 8) off += wc * r[c + 1];
            z = diag * rc + off;
        }
        s.z[c] = z;
        rz = (double)rc * (double)z;
        rr = (double)rc * (double)rc;
    }
    double v[2] = {rz, rr};
    reduce_last<2>(s, v, 2, nb, [&](const double* t) {
        SceneScalars* sc = s.sc;
        sc->rz = t[0];
        sc->rr0 = t[1];
        sc->rr = t[1];
        sc->beta = 0.0;
        if (!finite_d(t[0]) || t[1] == 0.0) sc->inner_done = 1;
    });
}

// KA: p = z + beta p_old; q = A' p; p.q
__global__ void __launch_bounds__(kBlock) k_pcg_a(const SceneDev* scenes, int par)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    SceneScalars* sc = s.sc;
    if (sc->outer_done || sc->inner_done) return;
    const int J = s.J, C = s.I * J;
    const int nb = (C + kBlock - 1) / kBlock;
    if ((int)blockIdx.x >= nb) return;
    const float* pold = par ? s.p2 : s.p;
    float* pnew = par ? s.p : s.p2;
    const float beta = (float)sc->beta;
    const bool first = sc->inner_iters == 0;
    const float* z = s.z;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double pq = 0.0;
    if (c < C) {
        auto P = [&](int k) { return first ? z[k] : z[k] + beta * pold[k]; };
        const uint8_t code = s.stencil[c];
        const float pc = P(c);
        float q = pc;
        if (code & 0x80) {
            const int n = (code >> 4) & 7;
            q = (float)n * pc;
            if (code & 1) q -= P(c - J);
            if (code & 2) q -= P(c + J);
            if (code & 4) q -= P(c - 1);
            if (code & 8) q -= P(c + 1);
        }
        pnew[c] = pc;
        s.q[c] = q;
        pq = (double)pc * (double)q;
    }
    double v[1] = {pq};
    reduce_last<1>(s, v, 1, nb, [&](const double* t) {
        sc->pq = t[0];
        if (!(t[0] > 0.0) || !finite_d(t[0])) {
            sc->inner_done = 1;   // breakdown: stop this sweep
            sc->alpha = 0.0;
            return;
        }
        sc->alpha = sc->rz / t[0];
    });
}

// KB: r = r_old - alpha q; d += alpha p; z = M r; r.z, r.r
__global__ void __launch_bounds__(kBlock) k_pcg_b(const SceneDev* scenes, int par)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    SceneScalars* sc = s.sc;
    if (sc->outer_done || sc->inner_done) return;
    const int J = s.J, C = s.I * J;
    const int nb = (C + kBlock - 1) / kBlock;
    if ((int)blockIdx.x >= nb) return;
    const float* rold = par ? s.r2 : s.r;
    float* rnew = par ? s.r : s.r2;
    const float* pnew = par ? s.p : s.p2;
    const float alpha = (float)sc->alpha;
    const float* q = s.q;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (c < C) {
        auto R = [&](int k) { return rold[k] - alpha * q[k]; };
        const uint8_t code = s.stencil[c];
        const float rc = R(c);
        float z = rc;
        if (code & 0x80) {
            float diag = 1.0f, off = 0.0f;
            if (code & 1) {
                const float w = inv_n(s.stencil[c - J]);
                diag += w * w;
                off += w * R(c - J);
            }
            if (code & 4) {
                const float w = inv_n(s.stencil[c - 1]);
                diag += w * w;
                off += w * R(c - 1);
            }
            const float wc = inv_n(code);
            if (code & 2) off += wc * R(c + J);
            if (code & 8) off += wc * R(c + 1);
            z = diag * rc + off;
        }
        rnew[c] = rc;
        s.z[c] = z;
        s.d[c] = s.d[c] + alpha * pnew[c];
        rz = (double)rc * (double)z;
        rr = (double)rc * (double)rc;
    }
    double v[2] = {rz, rr};
    reduce_last<2>(s, v, 2, nb, [&](const double* t) {
        sc->inner_iters += 1;
        sc->total_iters += 1;
        sc->rr = t[1];
        if (!finite_d(t[0]) || !finite_d(t[1])) {
            sc->numeric_bad = 1;
            sc->inner_done = 1;
            sc->outer_done = 1;
            return;
        }
        sc->beta = t[0] / sc->rz;
        sc->rz = t[0];
        if (t[1] <= sc->inner_eps2 * sc->rr0 || sc->inner_iters >= s.max_inner) sc->inner_done = 1;
    });
}

// ||b||^2 for an externally supplied right-hand side; x = 0, d = 0.
__global__ void __launch_bounds__(kBlock) k_solve_begin(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int C = s.I * s.J;
    const int nb = (C + kBlock - 1) / kBlock;
    if ((int)blockIdx.x >= nb) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double bb = 0.0;
    if (c < C) {
        const double b = s.b[c];
        bb = b * b;
        s.x64[c] = 0.0;
        s.d[c] = 0.f;
    }
    double v[1] = {bb};
    reduce_last<1>(s, v, 0, nb, [&](const double* t) {
        s.sc->bb = t[0];
        s.sc->rr64 = t[0];
    });
}

// reset the per-solve scalars (one thread per scene)
__global__ void k_solve_reset(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || threadIdx.x != 0) return;
    SceneScalars* sc = s.sc;
    sc->outer_done = 0;
    sc->inner_done = 1;
    sc->inner_iters = 0;
    sc->total_iters = 0;
    sc->sweeps = 0;
    sc->numeric_bad = 0;
    sc->beta = 0.0;
    sc->alpha = 0.0;
}

}  // namespace fg

namespace fg {
#define FG_LAUNCH0(name) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc) { hipLaunchKernelGGL(k_##name, g, dim3(kBlock), 0, st, sc); }
#define FG_LAUNCH1(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, T a) { hipLaunchKernelGGL(k_##name, g, dim3(kBlock), 0, st, sc, a); }
FG_LAUNCH0(refine_update)
FG_LAUNCH0(refine_residual)
FG_LAUNCH0(inner_init)
FG_LAUNCH1(pcg_a, int, par)
FG_LAUNCH1(pcg_b, int, par)
FG_LAUNCH0(solve_begin)
void launch_solve_reset(dim3 g, hipStream_t st, const SceneDev* sc) { hipLaunchKernelGGL(k_solve_reset, g, dim3(64), 0, st, sc); }
}  // namespace fg
