// Full-chip Micro path on the device: Partition::build, the batched
// MicroGradientExchange and the full-chip CTM / LSM drivers.
#pragma once
#include <vector>

#include "../../include/litho_amd.h"
#include "lsm.hpp"
#include "plan.hpp"

namespace litho {

// Partition::build (partition.cpp:21-80).
void partition_build(int height, int width, int grid_rows, int grid_cols, int halo, int guard, int* tile_size,
                     std::vector<litho_tile>& tiles);
void validate_tiles(const litho_tile* tiles, int num_tiles, int tile_size);

// Device state of one partition for the exchange: tile table, loss windows,
// raw gradients, wafer images and reduction buffers of `num_tiles` Micros.
class Exchanger {
 public:
  Exchanger(Plan& plan, const litho_tile* tiles, int num_tiles);
  const int* table() const { return table_.as<int>(); }
  int num_tiles() const { return T_; }
  int n() const { return n_; }
  // Loss windows of build_micro_loss_windows for the device local targets.
  void prepare(const float* local_targets, const litho_exchange_options& o);
  // MicroGradientExchange::compute on device local masks (after prepare()).
  // raw_out / wafer_out: optional device copies; per_tile: host, optional.
  void run(const float* local_masks, const litho_exchange_options& o, float* mask_gradient, float* raw_out,
           float* wafer_out, litho_tile_loss* per_tile, double totals[3]);

 private:
  Plan& plan_;
  int T_, n_, max_strip_ = 0, max_core_ = 0;
  std::vector<litho_tile> tiles_;
  const float* targets_ = nullptr;
  DevMem table_, pe_win_, pen_win_, raw_, wafer_, guard_part_, core_part_, red_all_, cept_red_, core_red_;
  HostPinned* host_ = nullptr;
 public:
  ~Exchanger() { delete host_; }
  Exchanger(const Exchanger&) = delete;
  Exchanger& operator=(const Exchanger&) = delete;
};

// MicroImaging::compute(global_mask, threshold, alpha) (mirco_imaging.cpp:9-27):
// every Micro extracted from the global mask (periodic Halo), imaged in
// batches of the plan capacity, the Cores assembled into the global aerial /
// wafer images.  Outputs nullable; local_* are [T][n][n].
void micro_imaging(Plan& plan, const litho_tile* tiles, int num_tiles, int height, int width, const float* global_mask,
                   double threshold, double alpha, float* global_aerial, float* global_wafer, float* local_aerial,
                   float* local_wafer);

// compare_micro_imaging (full_chip_imaging_compare.cpp:19-71): the target's
// Micros (periodic extracts of the global target) and the optimized local
// Micros (as given, with their own Guard and Halo) imaged in batches, each set
// of Cores assembled into global aerial / wafer images, and the PE of both
// against the global target (FP64).  Outputs nullable: global images
// height x width, optimized_local_aerial [T][n][n], optimized_mask the
// Core-assembled optimized Micros.
struct MicroComparisonOutputs {
  float* optimized_mask = nullptr;
  float* target_aerial = nullptr;
  float* target_wafer = nullptr;
  float* optimized_aerial = nullptr;
  float* optimized_wafer = nullptr;
  float* optimized_local_aerial = nullptr;
};
struct MicroComparison {
  double target_pe = 0.0, optimized_pe = 0.0;
};
MicroComparison compare_micro_imaging(Plan& plan, const litho_tile* tiles, int num_tiles, int height, int width,
                                      const float* global_target, const float* optimized_local_masks,
                                      double threshold, double alpha, const MicroComparisonOutputs& out);

// evaluate_mask_continuity (mask_continuity.cpp:8-61) of device Micros
// [num_tiles][tile_size^2] against the device height x width target.
litho_mask_continuity mask_continuity(Context& ctx, const litho_tile* tiles, int num_tiles, int tile_size, int height,
                                      int width, const float* local_masks, const float* target);

struct FullChipResult {
  int best_iteration = 0;
  double best_loss = 0.0;
};

// FullChipCTMOptimizer::optimize (full_chip_ctm.cpp:25-152).
FullChipResult fullchip_ctm(Plan& plan, const litho_tile* tiles, int num_tiles, int height, int width,
                            const float* initial_pattern, const float* global_target,
                            const litho_exchange_options& options, int iterations, double learning_rate,
                            float* global_mask_out, float* local_masks_out, litho_fullchip_row* history);

// FullChipLSMOptimizer::optimize, Binary render (full_chip_lsm.cpp:39-255).
FullChipResult fullchip_lsm(Plan& plan, const litho_tile* tiles, int num_tiles, int height, int width,
                            const float* initial_binary, const float* global_target,
                            const litho_exchange_options& options, const LsmOptions& lsm, int iterations,
                            int reinit_interval, float* global_mask_out, float* global_phi_out,
                            litho_fullchip_row* history);

}  // namespace litho
