// fg_pcg.hip -- pressure projection solve.
//
// System: the reference's matrix-free 5-point operator
// (IndexedPressureParameterUnit::multiply, pressuredata.h:132-146):
//   fluid row c:  (A x)_c = s * (n_c x_c - sum_{fluid nbrs} x_k),  s = dt/(rho dx^2)
//   other rows:   (A x)_c = x_c
// solved for the FP64 relative residual ||b - A x||2 / ||b||2 <= tol by
// iterative refinement: each sweep forms r64 = b - A x in FP64 and solves
// the correction A' d = r' with PCG in FP32 (A' = A/s on fluid rows,
// identity elsewhere, so the stencil has integer coefficients), except the
// iterate: each step's alpha p goes straight into x in FP64.  An FP32
// correction d would limit the true residual to about eps32 ||A'|| ||d||,
// above 1e-6 ||b|| for the smooth corrections of a developed flow, and cost
// a second sweep whatever the inner target; with x in FP64 only the FP32
// recurrence's drift (a few eps32 ||r'||) separates the inner residual from
// the true one.
//
// Preconditioner: the inverse-Poisson preconditioner the reference defines
// in inversepoissonpreconditioner.h:45-79, M = (I - L D^-1)(I - L D^-1)^T,
// restricted to A's 5-point pattern (its two fill-in diagonals dropped):
//   z_c = (1 + sum_lower (1/n_k)^2) r_c + sum_lower r_k/n_k + sum_upper r_u/n_c
// ("lower" = smaller linear index: i-1 and j-1).  With the fill-in the
// solves took 3-5 % fewer iterations but each iteration 10 % longer (four
// more loads per cell), so the restricted form is kept.  The reference's own
// matrix-free IPP (PressureIPPCoeficients.h:23-40) is mis-scaled and not
// symmetric and does not converge; it is reproduced only in the oracle.
//
// Each PCG iteration is two single-pass kernels:
//   KA(it): rho, ||r||^2 <- partials of KB(it-1); stop test; beta;
//           p = z + beta p_old (recomputed at the 4 neighbours), q = A' p, p.q
//   KB(it): p.q <- partials of KA(it); alpha; r = r_old - alpha q (recomputed
//           at the neighbours); x += alpha p (FP64); z = M r; r.z and r.r
// A thread issues the loads of its cells before it reduces the previous
// kernel's partials, so the two memory latencies overlap.
//
// Reductions without atomics: every kernel writes one FP64 partial per block
// (blocks own contiguous cell chunks), and every block of the *next* kernel
// sums all partials in the same fixed order, so all blocks derive the same
// scalars (alpha, beta, convergence) deterministically and no block waits on
// another.
#include "fg_kernels.h"

namespace fg {

__device__ __forceinline__ float inv_n(uint8_t code)
{
    const int n = (code >> 4) & 7;
    return n == 1 ? 1.0f : n == 2 ? 0.5f : n == 3 ? (1.0f / 3.0f) : n == 4 ? 0.25f : 0.0f;
}

// r64 = b - A x64 in FP64 (partials of ||r64||^2); inner rhs r = r64/s on
// fluid rows.  Also folds the ||b||^2 partials into sc->bb.  spec: only the
// scenes whose inner PCG has stopped (the FP64 test ahead of a poll, while
// other scenes may still iterate; a stopped sweep's r is dead).
__global__ void __launch_bounds__(kBlock) k_refine_residual(const SceneDev* scenes, int spec)
{
    const SceneDev& s = scenes[blockIdx.y];
    cptr<SceneScalars> scr = cst(s.sc);
    const int od = scr->outer_done, idn = scr->inner_done;
    const int nb = s.pcg_blocks, blk = solver_block();
    if (!s.active | (od != 0) | ((spec != 0) & (idn == 0)) | (blk >= nb)) return;
    double bb[1];
    sum_partials<1>(s, kSlotBB, nb, bb);
    if (blk == 0 && threadIdx.x == 0) s.sc->bb = bb[0];
    const int J = s.J, C = s.I * s.J;
    const Chunk ch = my_chunk(s, blk);
    gptr<const double> x = gbl<const double>(s.x64), bv = gbl<const double>(s.b);
    gptr<const uint8_t> codes = gbl<const uint8_t>(s.stencil);
    const double sc = s.sA;
    double rr = 0.0;
    // kRefItems cells a thread a pass (a plain kBlock stride's cells in its
    // order), every load issued before the stencil code is looked at
    // (clamped indices; the code selects which neighbours enter)
    constexpr int kRefItems = 4;
    for (int base = ch.begin; base < ch.end; base += kRefItems * kBlock) {
        uint8_t cd[kRefItems];
        double xc[kRefItems][5], bc[kRefItems];
#pragma unroll
        for (int u = 0; u < kRefItems; u++) {
            const int c = min(base + (int)threadIdx.x + u * kBlock, ch.end - 1);
            const int offs[5] = {0, -J, J, -1, 1};
            cd[u] = codes[c];
            bc[u] = bv[c];
#pragma unroll
            for (int e = 0; e < 5; e++) xc[u][e] = x[min(max(c + offs[e], 0), C - 1)];
        }
#pragma unroll
        for (int u = 0; u < kRefItems; u++) {
            const int c = base + threadIdx.x + u * kBlock;
            if (c >= ch.end) break;
            const uint8_t code = cd[u];
            double ax;
            if (code & 0x80) {
                const int n = (code >> 4) & 7;
                ax = (sc * n) * xc[u][0];
                if (code & 1) ax += (-sc) * xc[u][1];
                if (code & 2) ax += (-sc) * xc[u][2];
                if (code & 4) ax += (-sc) * xc[u][3];
                if (code & 8) ax += (-sc) * xc[u][4];
            } else {
                ax = xc[u][0];
            }
            const double r64 = bc[u] - ax;
            rr += r64 * r64;
            s.cr[c] = (float)((code & 0x80) ? r64 / sc : r64);
        }
    }
    double v[1] = {rr};
    put_partials<1>(s, kSlotRR64, v, blk);
}

constexpr int kItems = 2;                  // cells per thread per pass (4: 1.7 % slower on the batch, 8: 7 %)
constexpr int kPass = kItems * kBlock;     // cells per block per pass

// Outer test on ||r64||; if not converged: z = M r and the partials of r.z,
// r.r for the first iteration.  spec: the test alone, on the scenes whose
// inner PCG has stopped (after k_refine_residual's spec pass): a scene that
// passes is done, one that does not waits for the next sweep's k_cg_init.
__global__ void __launch_bounds__(kBlock) k_cg_init(const SceneDev* scenes, int sweep, int spec)
{
    const SceneDev& s = scenes[blockIdx.y];
    SceneScalars* sc = s.sc;
    cptr<SceneScalars> scr = cst(sc);
    const int od = scr->outer_done, idn = scr->inner_done;
    const int nb = s.pcg_blocks, blk = solver_block();
    if (!s.active | (od != 0) | ((spec != 0) & (idn == 0)) | (blk >= nb)) return;
    double rr64[1];
    sum_partials<1>(s, kSlotRR64, nb, rr64);
    const double bb = sc->bb;
    if (spec) {
        const double bn = sqrt(bb), rn = sqrt(rr64[0]);
        if (blk == 0 && threadIdx.x == 0 && (!finite_d(rr64[0]) || rn <= s.tol * bn || bn == 0.0)) {
            sc->rr64 = rr64[0];
            if (!finite_d(rr64[0])) sc->numeric_bad = 1;
            sc->outer_done = 1;
        }
        return;
    }
    if (sweep == 0 && s.precond == kPrecondMG && !(rr64[0] <= 25.0 * bb)) {
        // the warm start (see k_rhs) is far off (residual above 5 ||b||, or
        // not finite): start from zero, r = b (every block resets its
        // chunk; all see the same test)
        const Chunk ch = my_chunk(s, blk);
        for (int c = ch.begin + threadIdx.x; c < ch.end; c += kBlock) {
            const double b = s.b[c];
            s.x64[c] = 0.0;
            s.cr[c] = (float)((s.stencil[c] & 0x80) ? b / s.sA : b);
        }
        rr64[0] = bb;
    }
    const double bn = sqrt(bb), rn = sqrt(rr64[0]);
    const bool lead = blk == 0 && threadIdx.x == 0;
    if (!finite_d(rr64[0]) || rn <= s.tol * bn || bn == 0.0 || sweep >= kMaxSweeps) {
        if (lead) {
            sc->rr64 = rr64[0];
            if (!finite_d(rr64[0])) sc->numeric_bad = 1;
            sc->outer_done = 1;
        }
        return;
    }
    if (lead) {
        sc->rr64 = rr64[0];
        sc->sweeps = sweep + 1;
        sc->inner_done = 0;
        sc->inner_iters = 0;
        // ask the FP32 solve for the reduction still needed (with margin),
        // at least 1e-6 of ||b|| (relative to this sweep's start: a warm
        // start's residual may exceed ||b||)
        const double f = fmax(0.3 * s.tol * bn / rn, 1e-6 * fmin(1.0, bn / rn));
        sc->inner_eps2 = f * f;
    }
    if (s.precond == kPrecondMG) return;   // the V-cycle forms z and the partials
    const int J = s.J;
    const Chunk ch = my_chunk(s, blk);
    const float* __restrict__ r = s.cr;
    const uint8_t* __restrict__ code = s.stencil;
    double rz = 0.0, rr = 0.0;
    for (int c = ch.begin + threadIdx.x; c < ch.end; c += kBlock) {
        const uint8_t cc = code[c];
        const float rc = r[c];
        float z = rc;
        if (cc & 0x80) {
            float diag = 1.0f, off = 0.0f;
            if (cc & 1) {
                const float w = inv_n(code[c - J]);
                diag += w * w;
                off += w * r[c - J];
            }
            if (cc & 4) {
                const float w = inv_n(code[c - 1]);
                diag += w * w;
                off += w * r[c - 1];
            }
            const float wc = inv_n(cc);
            if (cc & 2) off += wc * r[c + J];
            if (cc & 8) off += wc * r[c + 1];
            z = diag * rc + off;
        }
        s.cu[c] = z;
        rz += (double)rc * (double)z;
        rr += (double)rc * (double)rc;
    }
    double v[2] = {rz, rr};
    put_partials<2>(s, kSlotCG, v, blk);
}

// KA: p = z + beta p_old; q = A' p; p.q        (z lives in cu)
__global__ void __launch_bounds__(kBlock) k_pcg_a(const SceneDev* scenes, const int2* map, int it)
{
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    SceneScalars* sc = s.sc;
    cptr<SceneScalars> scr = cst(sc);   // values earlier kernels wrote: scalar loads
    const int nb = s.pcg_blocks;
    // one test after the uniform loads, so that they issue together
    if (!s.active | (scr->outer_done != 0) | (scr->inner_done != 0) | (sb.block >= nb)) return;
    const int J = s.J, C = s.I * s.J;
    const bool first = it == 0;
    gptr<const float> z = gbl<const float>(s.cu);
    gptr<const float> pold = gbl<const float>((it & 1) ? s.cp : s.cs1);
    gptr<float> pnew = gbl((it & 1) ? s.cs1 : s.cp);
    gptr<float> q = gbl(s.cw);
    gptr<const uint8_t> code = gbl<const uint8_t>(s.stencil);
    const Chunk ch = my_chunk(s, sb.block);
    double pq = 0.0;
    float beta = 0.f;
    for (int base = ch.begin; base < ch.end; base += kPass) {
        // loads first (they depend neither on beta nor on the stencil code:
        // clamped indices, masked afterwards), so they issue together
        uint8_t cc[kItems];
        float zc[kItems][5], pc[kItems][5];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int c = base + threadIdx.x + k * kBlock;
            const bool in = c < ch.end;
            const int c0 = in ? c : ch.begin;
            const uint8_t code_c = code[c0];
            const int offs[5] = {0, -J, J, -1, 1};
            float zl[5], pl[5];
#pragma unroll
            for (int e = 0; e < 5; e++) {
                const int idx = min(max(c0 + offs[e], 0), C - 1);
                zl[e] = z[idx];
                pl[e] = first ? 0.f : pold[idx];
            }
            cc[k] = code_c;   // lanes past the chunk store nothing
#pragma unroll
            for (int e = 0; e < 5; e++) {
                // 0/1 factors rather than selects keep the loads unconditional
                const bool need = in && (e == 0 || ((cc[k] & 0x80) && (cc[k] & (1 << (e - 1)))));
                const float m = need ? 1.f : 0.f;
                zc[k][e] = zl[e] * m;
                pc[k][e] = pl[e] * m;
            }
        }
        if (base == ch.begin) {
            double t[2];
            sum_partials<2>(s, kSlotCG, nb, t);   // rho = r.z, ||r||^2 of the current residual
            const double rho = t[0], rr = t[1];
            const bool lead = sb.block == 0 && threadIdx.x == 0;
            const double rr0 = first ? rr : scr->rr0;
            if (first && lead) sc->rr0 = rr;
            bool stop = false;
            if (!finite_d(rho) || !finite_d(rr)) {
                stop = true;
                if (lead) {
                    sc->numeric_bad = 1;
                    sc->outer_done = 1;
                }
            } else if (rr == 0.0 || (!first && rr <= scr->inner_eps2 * rr0) || it >= s.max_inner) {
                stop = true;
            }
            if (stop) {
                if (lead) {
                    sc->inner_done = 1;
                    sc->rr = rr;
                }
                return;
            }
            if (lead) sc->gamma[it & 1] = rho;
            beta = first ? 0.f : (float)(rho / scr->gamma[(it - 1) & 1]);
        }
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int c = base + threadIdx.x + k * kBlock;
            if (c >= ch.end) continue;
            float P[5];
#pragma unroll
            for (int e = 0; e < 5; e++) P[e] = first ? zc[k][e] : zc[k][e] + beta * pc[k][e];
            // q = A' p summed in FP64 (integer weights times FP32 values:
            // exact but for the rare exponent spread), rounded once: for a
            // smooth p, n p_c - sum p_k cancels to far below |p|, and an
            // FP32 sum's error (about eps32 * 8 |p|) would enter the
            // residual recurrence and separate it from the true residual
            double qd = (double)P[0];
            if (cc[k] & 0x80) {
                qd = (double)((cc[k] >> 4) & 7) * (double)P[0];
                if (cc[k] & 1) qd -= (double)P[1];
                if (cc[k] & 2) qd -= (double)P[2];
                if (cc[k] & 4) qd -= (double)P[3];
                if (cc[k] & 8) qd -= (double)P[4];
            }
            pnew[c] = P[0];
            q[c] = (float)qd;
            pq += (double)P[0] * qd;
        }
    }
    double v[1] = {pq};
    put_partials<1>(s, kSlotPQ, v, sb.block);
}

// KB: r = r_old - alpha q; x += alpha p (FP64); z = M r; r.z, r.r (with the
// IPP; with the V-cycle, r and x only)
__global__ void __launch_bounds__(kBlock) k_pcg_b(const SceneDev* scenes, const int2* map, int it)
{
    const SceneBlock sb = scene_block(map, true);
    const SceneDev& s = scenes[sb.scene];
    SceneScalars* sc = s.sc;
    cptr<SceneScalars> scr = cst(sc);   // values earlier kernels wrote: scalar loads
    const int nb = s.pcg_blocks;
    // one test after the uniform loads, so that they issue together
    if (!s.active | (scr->outer_done != 0) | (scr->inner_done != 0) | (sb.block >= nb)) return;
    const int J = s.J, C = s.I * s.J;
    gptr<const float> rold = gbl<const float>((it & 1) ? s.cr2 : s.cr);
    gptr<float> rnew = gbl((it & 1) ? s.cr : s.cr2);
    gptr<const float> pnew = gbl<const float>((it & 1) ? s.cs1 : s.cp);
    gptr<const float> q = gbl<const float>(s.cw);
    gptr<float> z = gbl(s.cu);
    gptr<double> x = gbl(s.x64);
    gptr<const uint8_t> code = gbl<const uint8_t>(s.stencil);
    const Chunk ch = my_chunk(s, sb.block);
    if (s.precond == kPrecondMG) {
        // r and x only; the V-cycle that follows forms z, r.z and r.r.  A
        // pass's loads (kItems cells a thread, clamped to the chunk) are
        // issued before its stores -- and the first pass's before the p.q
        // reduction -- since the compiler cannot move a load of r_old above
        // a store to r_new.
        float a = 0.f;
        for (int base = ch.begin; base < ch.end; base += kPass) {
            float ro[kItems], qc[kItems], pc[kItems];
            double xc[kItems];
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                const int c = min(base + (int)threadIdx.x + k * kBlock, ch.end - 1);
                ro[k] = rold[c];
                qc[k] = q[c];
                xc[k] = x[c];
                pc[k] = pnew[c];
            }
            if (base == ch.begin) {
                double t[1];
                sum_partials<1>(s, kSlotPQ, nb, t);
                const double pq = t[0];
                const bool lead = sb.block == 0 && threadIdx.x == 0;
                if (!(pq > 0.0) || !finite_d(pq)) {
                    if (lead) sc->inner_done = 1;
                    return;
                }
                a = (float)(scr->gamma[it & 1] / pq);
                if (lead) {
                    sc->inner_iters = it + 1;
                    sc->total_iters += 1;
                }
            }
#pragma unroll
            for (int k = 0; k < kItems; k++) {
                const int c = base + (int)threadIdx.x + k * kBlock;
                if (c < ch.end) {
                    rnew[c] = ro[k] - a * qc[k];
                    x[c] = xc[k] + (double)a * (double)pc[k];
                }
            }
        }
        return;
    }
    double rz = 0.0, rr = 0.0;
    float alpha = 0.f;
    for (int base = ch.begin; base < ch.end; base += kPass) {
        uint8_t cc[kItems], cl[kItems][2];
        float rv[kItems][5], qv[kItems][5], pv[kItems];
        double xv[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int c = base + threadIdx.x + k * kBlock;
            const bool in = c < ch.end;
            const int c0 = in ? c : ch.begin;
            const uint8_t code_c = code[c0];
            const int offs[5] = {0, -J, J, -1, 1};
            float rl[5], ql[5];
#pragma unroll
            for (int e = 0; e < 5; e++) {
                const int idx = min(max(c0 + offs[e], 0), C - 1);
                rl[e] = rold[idx];
                ql[e] = q[idx];
            }
            const uint8_t code_im = code[max(c0 - J, 0)], code_jm = code[max(c0 - 1, 0)];
            const float pl = pnew[c0];
            const double xl = x[c0];
            cc[k] = code_c;   // lanes past the chunk store nothing
#pragma unroll
            for (int e = 0; e < 5; e++) {
                const bool need = in && (e == 0 || ((cc[k] & 0x80) && (cc[k] & (1 << (e - 1)))));
                const float m = need ? 1.f : 0.f;
                rv[k][e] = rl[e] * m;
                qv[k][e] = ql[e] * m;
            }
            cl[k][0] = ((cc[k] & 0x81) == 0x81) ? code_im : 0;
            cl[k][1] = ((cc[k] & 0x84) == 0x84) ? code_jm : 0;
            pv[k] = pl;
            xv[k] = xl;
        }
        if (base == ch.begin) {
            double t[1];
            sum_partials<1>(s, kSlotPQ, nb, t);
            const double pq = t[0];
            const bool lead = sb.block == 0 && threadIdx.x == 0;
            if (!(pq > 0.0) || !finite_d(pq)) {   // breakdown: end this sweep
                if (lead) sc->inner_done = 1;
                return;
            }
            alpha = (float)(scr->gamma[it & 1] / pq);
            if (lead) {
                sc->inner_iters = it + 1;
                sc->total_iters += 1;
            }
        }
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            const int c = base + threadIdx.x + k * kBlock;
            if (c >= ch.end) continue;
            float R[5];
#pragma unroll
            for (int e = 0; e < 5; e++) R[e] = rv[k][e] - alpha * qv[k][e];
            float zz = R[0];
            if (cc[k] & 0x80) {
                float diag = 1.0f, off = 0.0f;
                if (cc[k] & 1) {
                    const float w = inv_n(cl[k][0]);
                    diag += w * w;
                    off += w * R[1];
                }
                if (cc[k] & 4) {
                    const float w = inv_n(cl[k][1]);
                    diag += w * w;
                    off += w * R[3];
                }
                const float wc = inv_n(cc[k]);
                if (cc[k] & 2) off += wc * R[2];
                if (cc[k] & 8) off += wc * R[4];
                zz = diag * R[0] + off;
            }
            rnew[c] = R[0];
            z[c] = zz;
            x[c] = xv[k] + (double)alpha * (double)pv[k];
            rz += (double)R[0] * (double)zz;
            rr += (double)R[0] * (double)R[0];
        }
    }
    double v[2] = {rz, rr};
    put_partials<2>(s, kSlotCG, v, sb.block);
}

// ||b||^2 partials for an externally supplied right-hand side; x = 0.
__global__ void __launch_bounds__(kBlock) k_solve_begin(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    const int blk = solver_block();
    if (!s.active) return;
    if (blk >= s.pcg_blocks) return;
    const Chunk ch = my_chunk(s, blk);
    double bb = 0.0;
    for (int c = ch.begin + threadIdx.x; c < ch.end; c += kBlock) {
        const double b = s.b[c];
        bb += b * b;
        s.x64[c] = 0.0;
    }
    double v[1] = {bb};
    put_partials<1>(s, kSlotBB, v, blk);
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
    for (int l = 0; l < kMgMaxLevels; l++) sc->mg_stored[l] = 0;
}

}  // namespace fg

namespace fg {
#define FG_LAUNCH0(name) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc) { fg_launch(k_##name, g, dim3(kBlock), st, sc); }
#define FG_LAUNCH1(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, T a) { fg_launch(k_##name, g, dim3(kBlock), st, sc, a); }
FG_LAUNCH1(refine_residual, int, spec)
void launch_cg_init(dim3 g, hipStream_t st, const SceneDev* sc, int sweep, int spec)
{
    fg_launch(k_cg_init, g, dim3(kBlock), st, sc, sweep, spec);
}
void launch_pcg_a(dim3 g, hipStream_t st, const SceneDev* sc, const int2* map, int it)
{
    fg_launch(k_pcg_a, g, dim3(kBlock), st, sc, map, it);
}
void launch_pcg_b(dim3 g, hipStream_t st, const SceneDev* sc, const int2* map, int it)
{
    fg_launch(k_pcg_b, g, dim3(kBlock), st, sc, map, it);
}
FG_LAUNCH0(solve_begin)
void launch_solve_reset(dim3 g, hipStream_t st, const SceneDev* sc) { hipLaunchKernelGGL(k_solve_reset, g, dim3(64), 0, st, sc); }
}  // namespace fg
