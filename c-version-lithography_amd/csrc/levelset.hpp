// Level-set kernels (levelset.hip): WENO5-JS / Godunov normal-velocity term,
// curvature term, CFL step, RK2 stages of the full-chip LSM update and the
// chamfer signed-distance re-initialisation.  Fields are float32 arrays of
// `batch` row-major rows x cols tiles; stencil arithmetic is FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace litho {

// delta = Vn |grad phi| (Godunov upwind of WENO5 one-sided derivatives) and
// kappa = b (phi_xx phi_y^2 - 2 phi_x phi_y phi_xy + phi_yy phi_x^2) / |grad phi|^2;
// either output may be null.  (level_set_utils.cpp:150-278)
void launch_ls_terms(const float* phi, const float* vn, int batch, int rows, int cols, double dx, double dy,
                     double b, float* delta, float* kappa, hipStream_t s);

// Per tile: m = max |grad| over the update rectangle rects[b] = {y0, y1, x0, x1}
// (the velocity is grad inside it, 0 outside) and
// dt[b] = cfl / (m/dx + m/dy + 2 b (1/dx^2 + 1/dy^2)), or 1e-3 when that
// denominator is below 1e-9 (level_set_utils.cpp:280-288).
constexpr int kLsmDtBlocks = 64;
void launch_lsm_dt(const float* grad, const int* rects, int batch, int n, double dx, double dy, double cfl, double b,
                   float* partial_max /* [batch][kLsmDtBlocks] */, double* dt, hipStream_t s);

// One RK2 stage on square n x n tiles (full_chip_lsm.cpp:215-241), with
// Normal = kappa(phi_in) - Vn |grad phi_in|, Vn = grad inside rects[b]:
//   stage 1: out = phi_in + dt Normal inside the rectangle, phi_in outside;
//   stage 2: out = 0.5 base + 0.5 (phi_in + dt Normal) inside, base outside.
// mask_out (nullable) receives out >= 0 ? 1 : 0.
void launch_lsm_stage(int stage, const float* phi_in, const float* base, const float* grad, const int* rects,
                      const double* dt, int batch, int n, double dx, double dy, double b, float* out, float* mask_out,
                      hipStream_t s);

// render_levelset_mask (levelset_mask_renderer.cpp:7-33) of batch rows x cols
// tiles: type 0 Binary (phi >= 0), type 1 Gray (bilinear upscale by scale,
// zero threshold, mean over each pixel's scale x scale block).
void launch_render_mask(const float* phi, int batch, int rows, int cols, int type, int scale, float* mask,
                        hipStream_t s);

// Chamfer signed distance (level_set_utils.cpp:290-319 over OpenCV's
// distanceTransform(DIST_L2, 5)): inside = phi > level (or phi >= level when
// inclusive), sdf = dist_in dx - dist_out dx with the 5x5 chamfer distances in
// 16-bit fixed point.  With rects (nullable) only the rectangle of each tile is
// written to sdf (the LSM re-initialisation of the update window); mask_out
// (nullable) then receives sdf-updated tile >= 0 over the whole tile.
size_t chamfer_scratch_bytes(int batch, int rows, int cols);
void launch_reinit_sdf(const float* phi, int batch, int rows, int cols, double dx, double dy, float level, bool inclusive,
                       uint32_t* scratch, float* dist /* [batch][2][rows][cols] */, const int* rects, float* sdf,
                       float* mask_out, hipStream_t s);

}  // namespace litho
