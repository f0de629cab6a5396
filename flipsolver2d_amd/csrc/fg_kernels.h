// fg_kernels.h -- host-side launchers for the substep kernels.  Each .hip
// file defines the launchers of its own kernels, so kernels never cross a
// translation unit.  `grid.y` is the number of scenes in the launch; blocks
// beyond a scene's extent exit at once.
#pragma once

#include <hip/hip_ext.h>

#include <string>

#include "fg_device.h"

namespace fg {

// Events armed here are attached to the next launch (hipExtLaunchKernelGGL),
// so they time exactly that kernel on its stream.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
LaunchEvents& next_launch_events();

// Sets the message fg_last_error() returns and passes `code` through.
int report_error(int code, const std::string& msg);

template <typename F, typename... Args>
inline void fg_launch(F kernel, dim3 grid, dim3 block, hipStream_t st, Args... args)
{
    LaunchEvents& ev = next_launch_events();
    if (ev.start) {
        hipExtLaunchKernelGGL(kernel, grid, block, 0, st, ev.start, ev.stop, 0, args...);
        ev = LaunchEvents{};
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    }
}

// The same with `shm` bytes of dynamic LDS.
template <typename F, typename... Args>
inline void fg_launch_shm(F kernel, dim3 grid, dim3 block, size_t shm, hipStream_t st, Args... args)
{
    LaunchEvents& ev = next_launch_events();
    if (ev.start) {
        hipExtLaunchKernelGGL(kernel, grid, block, shm, st, ev.start, ev.stop, 0, args...);
        ev = LaunchEvents{};
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shm, st, args...);
    }
}

#define FG_DECLARE_LAUNCH(name, ...) void launch_##name(dim3 grid, hipStream_t st, const SceneDev* scenes, ##__VA_ARGS__)

// fg_particles.hip
// fg_mg.hip
FG_DECLARE_LAUNCH(mg_mark, int l);
FG_DECLARE_LAUNCH(mg_galerkin, int l);
FG_DECLARE_LAUNCH(mg_sym, int l);
FG_DECLARE_LAUNCH(mg_symmark, int l, int nsym);
FG_DECLARE_LAUNCH(mg_down0, const int2* map, int which);
FG_DECLARE_LAUNCH(mg_resid, const int2* map, int l);
FG_DECLARE_LAUNCH(mg_restrict, const int2* map, int l);
FG_DECLARE_LAUNCH(mg_coarse);
FG_DECLARE_LAUNCH(mg_prolong, const int2* map, int l);
FG_DECLARE_LAUNCH(mg_post, const int2* map, int l);
FG_DECLARE_LAUNCH(mg_rr, const int2* map, int l);
FG_DECLARE_LAUNCH(mg_pp, const int2* map, int l);
// max_j: the largest J of the launch's scenes (sizes the LDS of k_mg_up0);
// ring: k_mg_up0 (its LDS window) rather than k_mg_up0_wide
FG_DECLARE_LAUNCH(mg_up0, const int2* map, int which, int max_j, bool ring);

FG_DECLARE_LAUNCH(max_v2);
FG_DECLARE_LAUNCH(advect, int alive_from_key);
FG_DECLARE_LAUNCH(scan_local);
FG_DECLARE_LAUNCH(scan_top);
FG_DECLARE_LAUNCH(scan_add);
FG_DECLARE_LAUNCH(scatter);
FG_DECLARE_LAUNCH(fixup);
FG_DECLARE_LAUNCH(gather);
FG_DECLARE_LAUNCH(g2p_count);
FG_DECLARE_LAUNCH(finalize, int what);
constexpr int kScanTileCells = kBlock * 16;

// fg_grid.hip
FG_DECLARE_LAUNCH(stencil);
FG_DECLARE_LAUNCH(p2g);
FG_DECLARE_LAUNCH(sdf_materials);
FG_DECLARE_LAUNCH(sdf_exact);
FG_DECLARE_LAUNCH(extrapolate, int which);
// outside: extrapolateLevelsetOutside; else extrapolateLevelsetInside
FG_DECLARE_LAUNCH(levelset_init, bool outside);
FG_DECLARE_LAUNCH(levelset_band, int band, bool outside);
FG_DECLARE_LAUNCH(forces);
FG_DECLARE_LAUNCH(rhs);
// friction: also updateVelocityFromSolids (0: the second projection's plain application)
FG_DECLARE_LAUNCH(apply_solids, int friction);
FG_DECLARE_LAUNCH(density_rhs);
FG_DECLARE_LAUNCH(density_adjust);
constexpr int kBfsTile = 32;

// fg_init.hip
// Polygons of a scene's set-up, by set: 0 sinks, 1 sources, 2 solids,
// 3 initial fluid (firstFrameInit's pass order); polygon k has the vertices
// verts[voff[k] .. voff[k+1]) in metres.
struct PolySet {
    int32_t first, n;
};
struct InitGeo {
    PolySet set[4];
    const int32_t* voff;
    const float2* verts;
    const float* visc;      // per polygon: Emitter::viscosity of sources and fluids (0 otherwise)
    float dx;
};
void launch_init_cells(dim3 grid, hipStream_t st, const SceneDev* scenes, const InitGeo& geo);
void launch_seed(dim3 grid, hipStream_t st, const SceneDev* scenes, uint32_t seed);
FG_DECLARE_LAUNCH(clear_counts);
FG_DECLARE_LAUNCH(after_transfer);
FG_DECLARE_LAUNCH(reseed_count);
FG_DECLARE_LAUNCH(reseed_write);
FG_DECLARE_LAUNCH(reseed_commit);

// fg_visc.hip: the light viscosity model (cells; red: one block a scene)
FG_DECLARE_LAUNCH(visc_grid);
FG_DECLARE_LAUNCH(visc_setup);
FG_DECLARE_LAUNCH(visc_spmv);
FG_DECLARE_LAUNCH(visc_update);
FG_DECLARE_LAUNCH(visc_dir);
FG_DECLARE_LAUNCH(visc_apply);
FG_DECLARE_LAUNCH(visc_red, int what, double tol2);
void launch_visc_flags(hipStream_t st, const SceneDev* scenes, int n, int32_t* out);

// fg_nbflip.hip: NBFlipSolver's grid passes (nb_euler over the (I+1) x (J+1)
// lattice points, combine_vel over faces, the others over cells)
FG_DECLARE_LAUNCH(nb_euler);
FG_DECLARE_LAUNCH(nb_combine_sdf);
FG_DECLARE_LAUNCH(nb_combine_vel);
FG_DECLARE_LAUNCH(nb_materials);

// fg_pcg.hip
FG_DECLARE_LAUNCH(refine_residual, int spec);
FG_DECLARE_LAUNCH(cg_init, int sweep, int spec);
FG_DECLARE_LAUNCH(pcg_a, const int2* map, int it);
FG_DECLARE_LAUNCH(pcg_b, const int2* map, int it);
FG_DECLARE_LAUNCH(solve_begin);
FG_DECLARE_LAUNCH(solve_reset);

}  // namespace fg
