// Host orchestration of the level-set kernels (levelset.hip): the terms of
// evolve_normal_WENO_godunov / evolve_kappa, reinit_sdf and the RK2 update of
// one batch of full-chip Micros.
#pragma once
#include <vector>

#include "plan.hpp"

namespace litho {

struct LsmOptions {
  double dx = 1.0, dy = 1.0, cfl = 0.7, curvature_weight = 0.01;
  int mask_type = 0;      // render of the masks: 0 Binary, 1 Gray (full_chip_lsm.h:14-15)
  int render_scale = 4;   // Gray: upscale per direction
};

// render_levelset_mask(phi, type, scale) (levelset_mask_renderer.cpp:7-33) of
// batch rows x cols tiles on the device.
void render_levelset_mask(Context& ctx, int batch, int rows, int cols, const float* phi, int type, int scale,
                          float* mask);

void ls_terms(Context& ctx, int batch, int rows, int cols, const float* phi, const float* vn, double dx, double dy,
              double b, float* delta, float* kappa);
double ls_dt(double cfl, double dx, double dy, double max_h1, double max_h2, double b);
void ls_reinit_sdf(Context& ctx, int batch, int rows, int cols, const float* phi, double dx, double dy, float* sdf);

// One RK2 update (full_chip_lsm.cpp:215-241) of `batch` n x n tiles; rects: host
// [batch][4] update windows {y0, y1, x0, x1}.  phi_out may alias phi.  With
// reinit the update window is re-initialised to the chamfer SDF of the result.
// mask_out (nullable) receives the render of o.mask_type.  dt_out (host,
// nullable) receives each tile's step; passing it synchronises the stream.
void lsm_step(Context& ctx, int batch, int n, const float* phi, const float* grad, const int* rects,
              const LsmOptions& o, bool reinit, float* phi_out, float* mask_out, double* dt_out);

}  // namespace litho
