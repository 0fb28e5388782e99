// Edge-placement-error (EPE) loss and gradient on the device:
// Loss::evaluate_sparse_epe (loss.cpp:125-163) and Gradient::epe_gradient
// (gradient.cpp:198-227) over a fixed set of edge points (EpsResult: y, x
// pixel coordinates and weight_epe), alone or added to the PE gradient.
#pragma once
#include "plan.hpp"

namespace litho {

// The edge points of one tile size, with their bilinear stencils
// (make_bilinear_stencil, epe_utils.cpp:38-64) and a pixel -> points table
// that sums each pixel's scattered contributions in point order.
struct EpSet {
  Plan* plan = nullptr;
  int count = 0;
  int touched = 0;                 // distinct stencil pixels
  DevMem pix, wts, wep;            // per point: 4 pixel indices (-1 unused), 4 stencil weights, weight_epe
  DevMem csr_pix, csr_off, csr_ep, csr_w;
  DevMem sg, part, red, extra, aerial, zero_target;
};

EpSet* make_ep_set(Plan& plan, int count, const double* points_yx, const double* weights);

struct EpeResult {
  double pe_loss = 0.0;            // sum (W - T)^2 pe_window (0 without a target)
  double epe_loss = 0.0;           // weighted squared EPE (pixel^2)
  double epe_abs_sum_nm = 0.0;
  int active = 0;                  // points with weight_epe > 0
};

// grad_mask = pe_weight * (PE adjoint of Gradient::pe_gradient, windowed) +
// epe_weight * Gradient::epe_gradient; target null: the EPE term alone.
// aerial / wafer / grad_mask: device N x N, nullable.
EpeResult epe_gradient(EpSet& set, const float* mask, const float* target, const float* pe_window, double pe_weight,
                       double epe_weight, double threshold, double alpha, double pixel_size_nm, float* grad_mask,
                       float* aerial, float* wafer);

}  // namespace litho
