// CTM optimizers with device-resident state: the reference's fixed-step
// gradient descent (CTM_Optimizer.cpp:109-173) and L-BFGS
// (CTM_Optimizer.cpp:189-227, solver LBFGS++).  The L-BFGS logic is a
// per-tile host state machine, so the same code drives one tile or a batch of
// equal-size tiles that each take their own line-search path.
#pragma once
#include <vector>

#include "plan.hpp"

namespace litho {

struct CtmOptions {
  int method = 0;          // 0 gradient descent, 1 L-BFGS
  int iterations = 50;
  double learning_rate = 0.9;
  int history = 10;
  double tolerance = 0.0;
  double threshold = 0.25;
  double alpha = 50.0;
};

struct CtmTileResult {
  int evaluations = 0;
  int iterations = 0;
  double final_pe = 0.0;
  double final_grad_norm = 0.0;
  int status = 1;          // 0 converged, 1 max iterations, 2 line search failed
  std::vector<double> history;   // PE per evaluation
};

// Strong-Wolfe line-search parameters (LBFGS++ LBFGSParam defaults).
struct LineSearchParams {
  int max_linesearch = 20;
  double min_step = 1e-20;
  double max_step = 1e20;
  double ftol = 1e-4;
  double wolfe = 0.9;
};

// Safeguarded cubic trial inside a zoom bracket (Nocedal & Wright eq. 3.59).
double zoom_trial(double a_lo, double f_lo, double g_lo, double a_hi, double f_hi, double g_hi);

// Runs the optimizer on `batch` tiles.  Device inputs: targets [batch][N][N],
// theta0 (nullable).  Outputs: masks (best-PE mask for GD, final mask for
// L-BFGS) and theta (nullable).
std::vector<CtmTileResult> run_ctm(Plan& plan, int batch, const float* targets, const float* theta0,
                                   const CtmOptions& opt, float* out_masks, float* out_theta);

}  // namespace litho
