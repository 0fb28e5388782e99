// Batched CTM over independent tiles of several sizes.  The reference solves
// one tile per CTM_Optimizer (demo/demo_CTM_parallel_benchmark.cpp runs them
// side by side on CPU threads); here tiles of one size share a plan and are
// solved together in a pool of `chunk` optimizer slots whose evaluations replay
// one captured graph; a slot whose tile finished takes the next queued tile at
// the next round, so batches stay full until the queue drains.  The chunk size
// per tile size is chosen by timing batched evaluations (autotune) unless the
// caller fixes it.  Plans on distinct contexts (streams) are driven by one host
// thread each, so their rounds overlap on the GPU.
#pragma once
#include <vector>

#include "ctm_driver.hpp"

namespace litho {

struct BatchTile {
  int plan = 0;                  // index into the plan list (tile size)
  const float* target = nullptr; // device N x N
  const float* theta0 = nullptr; // device N x N, nullable (initialize_theta of the target)
  float* out_mask = nullptr;     // device N x N
  float* out_theta = nullptr;    // device N x N, nullable
};

// Tiles per batched solve for this plan: the largest batch whose per-tile
// evaluation time is within 3 % of the best among 1, 2, 4, ... up to `limit`
// (and the plan capacity): fewer, fuller rounds for the same tiles.  Times
// three warm evaluations per candidate; the result is cached in the plan.
// run_batch tunes its plans one after another before any of them runs.
int autotune_chunk(Plan& plan, const std::vector<const float*>& targets, int limit, const CtmOptions& opt);

std::vector<CtmTileResult> run_batch(const std::vector<Plan*>& plans, const BatchTile* tiles, int num_tiles,
                                     const CtmOptions& opt, int chunk, std::vector<int>* chunk_used);

}  // namespace litho
