// CTM optimizers with device-resident state: the reference's fixed-step
// gradient descent (CTM_Optimizer.cpp:109-173) and L-BFGS
// (CTM_Optimizer.cpp:189-227, solver LBFGS++).
//
// The optimizers run as a pool of slots of one plan: every slot holds one
// tile's optimizer state, and one step() advances every busy slot by one
// round (L-BFGS: the pair / direction update of the slots that accepted a
// step, then one batched evaluation of the slots with a pending trial; GD: one
// evaluation).  A finished slot is written out and can take the next tile at
// once, so the pool serves one tile, a fixed batch, a queue of tiles (batch
// driver) or a stream of requests (server) with the same code.  Each tile's
// trajectory does not depend on what shares its batch.
#pragma once
#include <memory>
#include <utility>
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
  double ftol = 1e-4;
  double wolfe = 0.9;
};

// Zoom trial of LBFGS++'s LineSearchNocedalWright: the minimiser of the
// quadratic through (step_lo, fx_lo) with slope dg_lo and (step_hi, fx_hi),
// or the bracket midpoint when that is not finite, not strictly inside the
// bracket, or within 1 % of the bracket width of an end.
double ls_quad_interp(double step_lo, double step_hi, double fx_lo, double fx_hi, double dg_lo);

// One tile for a pool slot: device N x N buffers.  The target is copied into
// the slot at admission; outputs are written when the tile finishes (best-PE
// mask for GD, final mask for L-BFGS).
struct CtmJob {
  long long id = -1;
  const float* target = nullptr;
  const float* theta0 = nullptr;   // nullable: initialize_theta(target)
  float* out_mask = nullptr;
  float* out_theta = nullptr;      // nullable
};

class CtmPool {
 public:
  // slots <= plan.capacity; the pool's evaluations run on the plan's stream.
  CtmPool(Plan& plan, int slots, const CtmOptions& opt);
  ~CtmPool();
  CtmPool(const CtmPool&) = delete;
  CtmPool& operator=(const CtmPool&) = delete;
  int slots() const;
  int busy() const;
  bool admit(const CtmJob& job);     // false when every slot is busy
  void step();                       // one round of every busy slot
  // Jobs finished since the last call, in completion order; their outputs are
  // complete once the plan's stream reaches the current point.
  std::vector<std::pair<long long, CtmTileResult>> take_finished();
  // Device bytes of one slot's optimizer state (admission budgets).
  static size_t slot_bytes(const Plan& plan, const CtmOptions& opt);

  struct Impl;

 private:
  std::unique_ptr<Impl> impl_;
};

void validate_ctm_options(const CtmOptions& opt);

// Runs the optimizer on `batch` tiles at once (one slot per tile).  Device
// inputs: targets [batch][N][N], theta0 (nullable).  Outputs: masks and theta
// (nullable), [batch][N][N].
std::vector<CtmTileResult> run_ctm(Plan& plan, int batch, const float* targets, const float* theta0,
                                   const CtmOptions& opt, float* out_masks, float* out_theta);

}  // namespace litho
