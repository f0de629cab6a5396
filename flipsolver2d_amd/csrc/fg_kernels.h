// fg_kernels.h -- host-side launchers for the substep kernels.  Each .hip
// file defines the launchers of its own kernels, so kernels never cross a
// translation unit.  `grid.y` is the number of scenes in the launch; blocks
// beyond a scene's extent exit at once.
#pragma once

#include "fg_device.h"

namespace fg {

#define FG_DECLARE_LAUNCH(name, ...) void launch_##name(dim3 grid, hipStream_t st, const SceneDev* scenes, ##__VA_ARGS__)

// fg_particles.hip
FG_DECLARE_LAUNCH(max_v2);
FG_DECLARE_LAUNCH(advect, int alive_from_key);
FG_DECLARE_LAUNCH(scan_local);
FG_DECLARE_LAUNCH(scan_top);
FG_DECLARE_LAUNCH(scan_add);
FG_DECLARE_LAUNCH(scatter);
FG_DECLARE_LAUNCH(fixup_gather);
FG_DECLARE_LAUNCH(commit_count);
FG_DECLARE_LAUNCH(g2p_count);
constexpr int kScanTileCells = kBlock * 16;

// fg_grid.hip
FG_DECLARE_LAUNCH(stencil);
FG_DECLARE_LAUNCH(p2g);
FG_DECLARE_LAUNCH(sdf_materials);
FG_DECLARE_LAUNCH(extrapolate, int which);
FG_DECLARE_LAUNCH(levelset_init);
FG_DECLARE_LAUNCH(levelset_band, int band);
FG_DECLARE_LAUNCH(forces);
FG_DECLARE_LAUNCH(rhs);
FG_DECLARE_LAUNCH(apply_solids);
constexpr int kBfsTile = 32;

// fg_pcg.hip
FG_DECLARE_LAUNCH(refine_update);
FG_DECLARE_LAUNCH(refine_residual);
FG_DECLARE_LAUNCH(inner_init);
FG_DECLARE_LAUNCH(pcg_a, int par);
FG_DECLARE_LAUNCH(pcg_b, int par);
FG_DECLARE_LAUNCH(solve_begin);
FG_DECLARE_LAUNCH(solve_reset);

}  // namespace fg
