// Full-chip Micro tiling kernels (tiling.hip): periodic gather, Core scatter,
// loss windows, the four-neighbour Guard exchange (fusion + CEPT), Core
// metrics and the CTM theta step on the update window.
//
// Device tile table, kTileStride ints per Micro:
//   [0..3]  tile_global y0, y1, x0, x1     [4..7]   core_global
//   [8..11] core_local                     [12..15] update_local
//   [16]    number of neighbours
//   [20 + 20 j ..] neighbour slot j = 0..3 (left, right, top, bottom order of
//   the existing ones): exists, id, direction, -, my_guard_local(4),
//   neighbor_overlap_local(4), my_overlap_local(4), neighbor_guard_local(4)
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/litho_amd.h"

namespace litho {

constexpr int kTileStride = 100;
constexpr int kExchangeBlocks = 64;   // blocks per (Micro, neighbour slot) of the exchange kernels
constexpr int kMetricBlocks = 64;     // blocks per Micro of the Core-metric reduction

std::vector<int> tile_table(const litho_tile* tiles, int num_tiles);

void launch_extract(const float* global, int height, int width, const int* table, int num_tiles, int n, float* local,
                    hipStream_t s);
void launch_assemble(const float* local, const int* table, int num_tiles, int n, int width, float* global,
                     int max_core_elems, hipStream_t s);
// pe / penalty windows of build_micro_loss_windows (micro_gradient_exchange.cpp:65-93), PE and PEPenalty
void launch_windows(const float* targets, const int* table, int num_tiles, int n, int region_core_guard,
                    int loss_type, float* pe_win, float* pen_win, hipStream_t s);

struct ExchangeParams {
  int fusion, linear, cept;
  double w_owner, w_neighbor, cept_weight;
};
// Guard strips: mask_gradient[g] = fused raw gradient (+ cept_weight * 2 diff);
// partial[(t * 4 + slot) * kExchangeBlocks + blk] = sum diff^2 of the strip.
void launch_guard_exchange(const int* table, int num_tiles, int n, const float* masks, const float* raw,
                           float* mask_gradient, const ExchangeParams& p, int max_strip, double* partial,
                           hipStream_t s);
// Two-sided CEPT, Core side of neighbour slot `slot`: mask_gradient[ov] += cept_weight * 2 (m[ov] - m_nb[os]).
void launch_overlap_cept(const int* table, int num_tiles, int n, int slot, const float* masks, float* mask_gradient,
                         double cept_weight, int max_strip, hipStream_t s);
// Core metrics: partial[(t * kMetricBlocks + blk) * kCoreMetrics + {0, 1, 2}] = PE (loss window),
// background penalty and PE on the main pattern (target > 0.5) over the Core.
constexpr int kCoreMetrics = 3;
void launch_core_metrics(const int* table, int num_tiles, int n, const float* wafer, const float* targets,
                         const float* pe_win, const float* pen_win, double penalty_threshold, int max_core_elems,
                         double* partial, hipStream_t s);
// theta -= lr (-0.5 g sin theta) and mask = (1 + cos theta) / 2 on each update window.
void launch_theta_step(const int* table, int num_tiles, int n, const float* grad, double lr, float* theta,
                       float* mask, int max_update_elems, hipStream_t s);
void launch_mask_from_theta_tiles(const float* theta, size_t count, float* mask, hipStream_t s);
// Mask continuity partials, kContinuityValues per (Micro, neighbour slot, block
// of kExchangeBlocks): Guard-vs-neighbour-Core count, sum d^2, sum |d|, max |d|;
// Core seam pairs (right / bottom neighbours, on the assembled global mask and
// target): count, sum |jump|, sum |jump - target jump|, flat count, flat sum
// |jump|, foreground count, foreground sum |jump|.
constexpr int kContinuityValues = 11;
void launch_mask_continuity(const int* table, int num_tiles, int n, int width, const float* masks,
                            const float* global, const float* target, double* partial, hipStream_t s);
// partial[blk] = FP64 sum of (a - b)^2 over the grid-strided elements of block
// blk (blocks of 256 threads), summed in a fixed order.
void launch_sq_diff_partials(const float* a, const float* b, size_t n, double* partial, int blocks, hipStream_t s);

}  // namespace litho
