// fg_visc.hip -- viscosity: LightViscosityModel (FlipSolver2dLib/
// viscositymodel.cpp:4-140) and the viscosity part of centeredParamsToGrid
// (flipsolver2d.cpp:1379-1431).
//
// k_visc_grid: a cell's viscosity is the quadratic-B-spline average of the
//   particles' viscosity over the cells i-2 .. i+1, j-2 .. j+1 (where the
//   spline at (x - i, y - j) is nonzero), summed in (cell, id) order; a
//   SOURCE cell takes its emitter's (afterTransfer, 396-397).
// The model (after the projection and updateVelocityFromSolids): one matrix
//   over the cells -- diagonal 1 on solid rows, 1 + 4 nu dt elsewhere;
//   symmetric couplings to (i+1, j) and (i, j+1) as the reference's
//   assembly leaves them, i.e. the later writer's value: nu of the far end
//   times dt unless that end is solid (then the near end's) -- solved twice,
//   for rho U and rho V on the cells' faces (i, j), by Jacobi-preconditioned
//   CG in FP64 from zero (Eigen's ConjugateGradient with its diagonal
//   preconditioner), both systems in the same launches:
//     k_visc_setup    matrix, b, x = 0, r = b, p = r / d; partials bb, rz
//     k_visc_spmv     q = A p; partials p.q
//     k_visc_update   x += alpha p, r -= alpha q; partials r.r, r.z
//     k_visc_dir      p = r / d + beta p
//     k_visc_red      one block a scene: the sums, alpha, beta, the stop test
//                     ||r||^2 < tol^2 ||b||^2 (a stopped system keeps alpha 0)
//     k_visc_apply    U = x / rho, V likewise (faces i < I, j < J)
// The second projection (flipsolver2d.cpp:450-453) reuses the solve.
#include "fg_kernels.h"

namespace fg {

__device__ __forceinline__ int visc_blocks(const SceneDev& s) { return (s.I * s.J + kBlock - 1) / kBlock; }

__global__ void __launch_bounds__(kBlock) k_visc_grid(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.visc_on) return;
    const int I = s.I, J = s.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= I * J) return;
    const int i = c / J, j = c - (c / J) * J;
    float wt = 1e-10f, vv = 0.0f;
    bool known = false;
    const int j0 = max(0, j - 2), j1 = min(J - 1, j + 1);
    for (int a = max(0, i - 2); a <= min(I - 1, i + 1); a++) {
        const int k0 = s.cellStart[a * J + j0], k1 = s.cellStart[a * J + j1 + 1];
        for (int k = k0; k < k1; k++) {
            const float2 p = ldg2(s.pos, k);
            const float w = qbspline(p.x - (float)i, p.y - (float)j);
            // NBFlipSolver's own (nbflipsolver.cpp:441-504): |w| > 1e-6, and
            // the sum divided only where some weight counted
            if (s.nb ? fabsf(w) > 1e-6f : w > 1e-9f) {
                wt += w;
                vv += w * s.pvisc[k];
                known = true;
            }
        }
    }
    if (s.nb) {   // its SOURCE cells are set in the combination (k_nb_combine_vel)
        s.cvisc[c] = known ? vv / wt : vv;
        return;
    }
    float v = vv / wt;
    if (s.mat[c] == M_SOURCE) {
        const int e = s.eid[c];
        if (e >= 0 && e < s.n_emitters) v = s.emitters[e].w;
    }
    s.cvisc[c] = v;
}

// the four per-block partials of one block (thread 0 stores them)
__device__ __forceinline__ void visc_put(const SceneDev& s, double a, double b, double c, double d)
{
    __shared__ double red[kBlock / 64];
    const int cb = visc_blocks(s);
    const double v[4] = {a, b, c, d};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double t = block_sum(v[q], red);
        if (threadIdx.x == 0) s.vpart[(size_t)q * cb + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kBlock) k_visc_setup(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.visc_on || (int)blockIdx.x >= visc_blocks(s)) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double bbu = 0, bbv = 0, rzu = 0, rzv = 0;
    if (c < C) {
        const int i = c / J, j = c - (c / J) * J;
        const double scale = s.dt;
        const bool sc = is_solid(s.mat[c]);
        const float vc = s.cvisc[c];
        const double d = sc ? 1.0 : 4.0 * (vc * scale) + 1.0;
        double e = 0.0, n = 0.0;
        if (i + 1 < I) {
            const bool sb = is_solid(s.mat[c + J]);
            if (!(sc && sb)) e = (sc ? s.cvisc[c + J] : vc) * scale;
        }
        if (j + 1 < J) {
            const bool sb = is_solid(s.mat[c + 1]);
            if (!(sc && sb)) n = (sc ? s.cvisc[c + 1] : vc) * scale;
        }
        s.vd[c] = d;
        s.ve[c] = e;
        s.vn[c] = n;
        const float rho = (float)s.rho;
        const double bu = rho * s.u[(size_t)i * J + j];        // fillRhs: float product
        const double bv = rho * s.v[(size_t)i * (J + 1) + j];
        s.vx[c] = 0.0;
        s.vx[C + c] = 0.0;
        s.vr[c] = bu;
        s.vr[C + c] = bv;
        const double zu = bu / d, zv = bv / d;
        s.vp[c] = zu;
        s.vp[C + c] = zv;
        bbu = bu * bu;
        bbv = bv * bv;
        rzu = bu * zu;
        rzv = bv * zv;
    }
    visc_put(s, bbu, bbv, rzu, rzv);
}

__global__ void __launch_bounds__(kBlock) k_visc_spmv(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.visc_on || (int)blockIdx.x >= visc_blocks(s)) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double pq[2] = {0.0, 0.0};
    if (c < C) {
        const int i = c / J, j = c - (c / J) * J;
        const double d = s.vd[c];
        const double e = s.ve[c], n = s.vn[c];
        const double ew = i > 0 ? s.ve[c - J] : 0.0, ns = j > 0 ? s.vn[c - 1] : 0.0;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double* p = s.vp + (size_t)q * C;
            double v = d * p[c];
            if (i + 1 < I) v += e * p[c + J];
            if (i > 0) v += ew * p[c - J];
            if (j + 1 < J) v += n * p[c + 1];
            if (j > 0) v += ns * p[c - 1];
            s.vq[(size_t)q * C + c] = v;
            pq[q] = p[c] * v;
        }
    }
    visc_put(s, pq[0], pq[1], 0.0, 0.0);
}

__global__ void __launch_bounds__(kBlock) k_visc_update(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.visc_on || (int)blockIdx.x >= visc_blocks(s)) return;
    const int C = s.I * s.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double rr[2] = {0.0, 0.0}, rz[2] = {0.0, 0.0};
    if (c < C) {
        const double d = s.vd[c];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double alpha = cst(s.vsc)[8 + q];
            const size_t k = (size_t)q * C + c;
            const double p = s.vp[k];
            s.vx[k] += alpha * p;
            const double r = s.vr[k] - alpha * s.vq[k];
            s.vr[k] = r;
            rr[q] = r * r;
            rz[q] = r * (r / d);
        }
    }
    visc_put(s, rr[0], rr[1], rz[0], rz[1]);
}

__global__ void __launch_bounds__(kBlock) k_visc_dir(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.visc_on || (int)blockIdx.x >= visc_blocks(s)) return;
    const int C = s.I * s.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const double d = s.vd[c];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const double beta = cst(s.vsc)[10 + q];
        const size_t k = (size_t)q * C + c;
        s.vp[k] = s.vr[k] / d + beta * s.vp[k];
    }
}

// what 0: after k_visc_setup; 1: after k_visc_spmv (alpha); 2: after
// k_visc_update (the stop test, beta).  vsc: bb 0-1, rz 2-3, pq 4-5, rr 6-7,
// alpha 8-9, beta 10-11, done 12-13, iterations 14-15.
__global__ void __launch_bounds__(kBlock) k_visc_red(const SceneDev* scenes, int what, double tol2)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.visc_on) return;
    __shared__ double red[kBlock / 64];
    const int cb = visc_blocks(s);
    const int nslot = what == 1 ? 2 : 4;
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < nslot; q++) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < cb; b += kBlock) acc += s.vpart[(size_t)q * cb + b];
        sum[q] = block_sum(acc, red);
    }
    if (threadIdx.x != 0) return;
    double* v = s.vsc;
    for (int q = 0; q < 2; q++) {
        if (what == 0) {
            v[q] = sum[q];
            v[2 + q] = sum[2 + q];
            v[6 + q] = sum[q];
            v[12 + q] = sum[q] == 0.0 ? 1.0 : 0.0;   // a zero right-hand side: x = 0
            v[14 + q] = 0.0;
            v[8 + q] = 0.0;
            v[10 + q] = 0.0;
        } else if (what == 1) {
            v[4 + q] = sum[q];
            v[8 + q] = v[12 + q] != 0.0 ? 0.0 : v[2 + q] / sum[q];
        } else {
            v[10 + q] = 0.0;
            if (v[12 + q] != 0.0) continue;
            v[6 + q] = sum[q];
            v[14 + q] += 1.0;
            if (sum[q] < tol2 * v[q]) {
                v[12 + q] = 1.0;
            } else {
                v[10 + q] = sum[2 + q] / v[2 + q];
                v[2 + q] = sum[2 + q];
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_visc_apply(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.visc_on) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const int i = c / J, j = c - (c / J) * J;
    const float rho = (float)s.rho;
    s.u[(size_t)i * J + j] = (float)(s.vx[c] / rho);          // applyResult
    s.v[(size_t)i * (J + 1) + j] = (float)(s.vx[C + c] / rho);
}

// how many scenes of the launch still iterate (either system)
__global__ void k_visc_flags(const SceneDev* scenes, int n, int32_t* out)
{
    int open = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const SceneDev& s = scenes[k];
        if (s.active && s.visc_on && (s.vsc[12] == 0.0 || s.vsc[13] == 0.0)) open++;
    }
    for (int o = 32; o > 0; o >>= 1) open += __shfl_down(open, o, 64);
    if (threadIdx.x == 0) out[0] = open;
}

}  // namespace fg

namespace fg {
#define FG_VLAUNCH(name) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc) { fg_launch(k_##name, g, dim3(kBlock), st, sc); }
FG_VLAUNCH(visc_grid)
FG_VLAUNCH(visc_setup)
FG_VLAUNCH(visc_spmv)
FG_VLAUNCH(visc_update)
FG_VLAUNCH(visc_dir)
FG_VLAUNCH(visc_apply)
void launch_visc_red(dim3 g, hipStream_t st, const SceneDev* sc, int what, double tol2)
{
    fg_launch(k_visc_red, g, dim3(kBlock), st, sc, what, tol2);
}
void launch_visc_flags(hipStream_t st, const SceneDev* sc, int n, int32_t* out)
{
    hipLaunchKernelGGL(k_visc_flags, dim3(1), dim3(64), 0, st, sc, n, out);
}
}  // namespace fg
