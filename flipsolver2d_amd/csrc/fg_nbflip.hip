// fg_nbflip.hip -- the grid kernels of NBFlipSolver::step
// (FlipSolver2dLib/nbflipsolver.cpp:26-64).
//
// NB-FLIP keeps particles only in a band below the liquid surface; the deep
// interior is carried by semi-Lagrangian advection of the velocity grid and
// the level set, and the two are combined once the particle grid exists:
//   k_nb_euler        the Euler pass of advect() (66-113, with
//                     FlipSolver::eulerAdvectionThread, flipsolver2d.cpp:
//                     340-351): every lattice point (i, j) is traced back by
//                     RK4 over -dt once and U, V and the level set are lerped
//                     there (the three grids share the point's trace);
//   k_nb_combine_sdf  updateGridFromSources' distance and combineLevelset
//                     (320-349, 358-369): min(particle SDF, source distance,
//                     advected + 1);
//   k_nb_combine_vel  FlipSolver::afterTransfer's source faces, then
//                     combineVelocityGrid (371-398): a face keeps the particle
//                     value where the combined level set there lies above
//                     the combination band and takes the advected value
//                     elsewhere; applyBodyForces (gridUpdate, 213-226) is
//                     fused in, as it is the next write of those faces;
//   k_nb_materials    updateMaterials on the final level set (gridUpdate).
// The band prune of advect() runs inside k_advect, the band reseed inside
// the reseed kernels, and the level-set extrapolations are the BFS band
// kernels (outside and inside variants).
#include "fg_kernels.h"

namespace fg {

__global__ void __launch_bounds__(kBlock) k_nb_euler(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.nb) return;
    const int I = s.I, J = s.J;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= (I + 1) * (J + 1)) return;
    const int i = k / (J + 1), j = k - (k / (J + 1)) * (J + 1);
    const float2 q = rk4(s.uf, s.vf, I, J, make_float2((float)i, (float)j), -s.dt);
    if (j < J) s.adv_u[(size_t)i * J + j] = lerp_grid(s.uf, I + 1, J, 0.5f, 0.f, q.x, q.y);
    if (i < I) s.adv_v[(size_t)i * (J + 1) + j] = lerp_grid(s.vf, I, J + 1, 0.f, 0.5f, q.x, q.y);
    if (i < I && j < J) s.adv_sdf[(size_t)i * J + j] = lerp_grid(s.fsdf, I, J, 0.f, 0.f, q.x, q.y);
    // with viscosity the viscosity grid too (offset (1/2, 1/2))
    if (s.visc_on && i < I && j < J) s.adv_visc[(size_t)i * J + j] = lerp_grid(s.cvisc, I, J, 0.5f, 0.5f, q.x, q.y);
}

__global__ void __launch_bounds__(kBlock) k_nb_combine_sdf(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.nb) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= s.I * s.J) return;
    float f = s.fsdf[c];
    const float d = s.src_sdf[c];
    f = (f < d) ? f : d;                      // std::min(dist, sdf)
    const float a = s.adv_sdf[c] + 1.f;
    s.fsdf[c] = (f < a) ? f : a;              // std::min(advected + h, sdf)
}

// The particle grid is the extrapolated, saved one (su/sv, which G2P reads
// as it was); the working grid u/v gets the combination plus the body force.
__global__ void __launch_bounds__(kBlock) k_nb_combine_vel(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.nb) return;
    const int I = s.I, J = s.J;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const float du = s.force * s.g_i, dv = s.force * s.g_j;
    auto source = [&](int i, int j, float4* em) -> bool {
        if (!s.has_sources || i >= I || j >= J) return false;
        const int c = i * J + j;
        if (s.mat[c] != M_SOURCE) return false;
        const int e = s.eid[c];
        if (e < 0 || e >= s.n_emitters) return false;
        *em = s.emitters[e];
        return em->z != 0.f;
    };
    if (k < (I + 1) * J) {
        const int i = k / J, j = k - (k / J) * J;
        float part = s.su[k];
        float4 em;
        if (source(i, j, &em)) {
            part = em.x;
            s.uvalS[k] = 1;
        }
        const float sdf = lerp_grid(s.fsdf, I, J, 0.f, 0.f, (float)i, 0.5f + (float)j);
        s.u[k] = (sdf > kNbCombinationBand ? part : s.adv_u[k]) + du;
    }
    if (k < I * (J + 1)) {
        const int i = k / (J + 1), j = k - (k / (J + 1)) * (J + 1);
        float part = s.sv[k];
        float4 em;
        if (source(i, j, &em)) {
            part = em.y;
            s.vvalS[k] = 1;
        }
        const float sdf = lerp_grid(s.fsdf, I, J, 0.f, 0.f, 0.5f + (float)i, (float)j);
        s.v[k] = (sdf > kNbCombinationBand ? part : s.adv_v[k]) + dv;
    }
    // the viscosity grid: afterTransfer's SOURCE cells (flipsolver2d.cpp:396-397),
    // updateGridFromSources' (nbflipsolver.cpp:338-346, to both grids),
    // combineCenteredGrids at the cell centre (400-419)
    if (s.visc_on && k < I * J) {
        const int i = k / J, j = k - (k / J) * J;
        float vis = s.cvisc[k], adv = s.adv_visc[k];
        auto ev = [&](int e) { return (e >= 0 && e < s.n_emitters) ? s.emitters[e].w : 0.f; };
        if (s.mat[k] == M_SOURCE && s.eid[k] >= 0) vis = ev(s.eid[k]);
        const int sid = s.src_id[k];
        if (sid >= 0) {
            vis = ev(sid);
            adv = ev(sid);
        }
        const float sdf = lerp_grid(s.fsdf, I, J, 0.f, 0.f, 0.5f + (float)i, (float)((double)j + 0.5));
        s.cvisc[k] = sdf > kNbCombinationBand ? vis : adv;
    }
}

// FlipSolver::updateMaterials, flipsolver2d.cpp:1053-1075
__global__ void __launch_bounds__(kBlock) k_nb_materials(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.nb) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= s.I * s.J) return;
    int8_t m = s.mat[c];
    if (s.fsdf[c] < 0.f) {
        if (is_empty(m)) m = M_FLUID;
    } else {
        if (is_strict_fluid(m)) m = M_EMPTY;
    }
    s.mat[c] = m;
}

}  // namespace fg

namespace fg {
void launch_nb_euler(dim3 g, hipStream_t st, const SceneDev* sc) { fg_launch(k_nb_euler, g, dim3(kBlock), st, sc); }
void launch_nb_combine_sdf(dim3 g, hipStream_t st, const SceneDev* sc)
{
    fg_launch(k_nb_combine_sdf, g, dim3(kBlock), st, sc);
}
void launch_nb_combine_vel(dim3 g, hipStream_t st, const SceneDev* sc)
{
    fg_launch(k_nb_combine_vel, g, dim3(kBlock), st, sc);
}
void launch_nb_materials(dim3 g, hipStream_t st, const SceneDev* sc)
{
    fg_launch(k_nb_materials, g, dim3(kBlock), st, sc);
}
}  // namespace fg
