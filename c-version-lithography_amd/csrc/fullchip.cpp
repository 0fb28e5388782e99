// Full-chip Micro path (fullchip.hpp).
//
// Restates Partition::build (partition.cpp:21-80), MicroGradientExchange::compute
// for the PE / PEPenalty losses (micro_gradient_exchange.cpp:134-334),
// FullChipCTMOptimizer::optimize (full_chip_ctm.cpp:25-152) and
// FullChipLSMOptimizer::optimize with the Binary renderer (full_chip_lsm.cpp:39-255).
// Phase 1 of the exchange (imaging, loss and adjoint of every Micro) is one
// batched evaluation per chunk of plan capacity; phase 2 runs on the device
// (tiling.hip); the host only gathers the per-Micro FP64 scalars once per call.
#include "fullchip.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

#include "levelset.hpp"
#include "opt_kernels.hpp"
#include "tiling.hpp"

namespace litho {

namespace {

litho_rect make_rect(int y0, int y1, int x0, int x1) { return litho_rect{y0, y1, x0, x1}; }
litho_rect shifted(const litho_rect& r, int dy, int dx) { return make_rect(r.y0 + dy, r.y1 + dy, r.x0 + dx, r.x1 + dx); }
litho_rect expanded(const litho_rect& r, int w) { return make_rect(r.y0 - w, r.y1 + w, r.x0 - w, r.x1 + w); }
litho_rect intersection(const litho_rect& a, const litho_rect& b) {
  return make_rect(std::max(a.y0, b.y0), std::min(a.y1, b.y1), std::max(a.x0, b.x0), std::min(a.x1, b.x1));
}
int area(const litho_rect& r) { return std::max(0, r.y1 - r.y0) * std::max(0, r.x1 - r.x0); }
bool inside_tile(const litho_rect& r, int n) {
  return r.y0 >= 0 && r.x0 >= 0 && r.y1 <= n && r.x1 <= n && r.y0 <= r.y1 && r.x0 <= r.x1;
}

void check_options(const litho_exchange_options& o) {
  if (o.loss_type != 0 && o.loss_type != 1) throw Error(1, "MicroGradientOptions: loss_type must be PE or PEPenalty");
  if (o.loss_region != 0 && o.loss_region != 1) throw Error(1, "MicroGradientOptions: loss_region must be Core or CoreGuard");
  if (o.enable_grad_fusion) {
    if (o.fusion_mode != 0 && o.fusion_mode != 1) throw Error(1, "MicroGradientOptions: fusion_mode must be Average or Linear");
    if (!(o.fusion_weight_owner >= 0.0) || !(o.fusion_weight_neighbor >= 0.0) ||
        !(o.fusion_weight_owner + o.fusion_weight_neighbor > 0.0))
      throw Error(1, "MicroGradientOptions: fusion weights must be non-negative with a positive sum");
  }
  if (!std::isfinite(o.threshold) || !std::isfinite(o.alpha)) throw Error(1, "threshold and alpha must be finite");
}

}  // namespace

void partition_build(int height, int width, int grid_rows, int grid_cols, int halo, int guard, int* tile_size,
                     std::vector<litho_tile>& tiles) {
  if (height <= 0 || width <= 0 || grid_rows <= 0 || grid_cols <= 0 || grid_rows > height || grid_cols > width)
    throw Error(1, "Partition::build: invalid chip or grid size");
  if (halo < 0 || guard < 0 || guard > halo) throw Error(1, "Partition::build: need 0 <= guard <= halo");
  const int max_r = (height + grid_rows - 1) / grid_rows;
  const int max_c = (width + grid_cols - 1) / grid_cols;
  int ts = std::max(max_r, max_c) + 2 * halo;
  ts += ts % 2;
  *tile_size = ts;
  tiles.assign((size_t)grid_rows * grid_cols, litho_tile{});
  for (int r = 0; r < grid_rows; ++r)
    for (int c = 0; c < grid_cols; ++c) {
      litho_tile& t = tiles[(size_t)r * grid_cols + c];
      t.id = r * grid_cols + c;
      t.row = r;
      t.col = c;
      t.halo = halo;
      t.guard = guard;
      t.core_global = make_rect((int)((long long)r * height / grid_rows), (int)((long long)(r + 1) * height / grid_rows),
                                (int)((long long)c * width / grid_cols), (int)((long long)(c + 1) * width / grid_cols));
      const int oy = t.core_global.y0 - halo, ox = t.core_global.x0 - halo;
      t.tile_global = make_rect(oy, oy + ts, ox, ox + ts);
      t.core_local = shifted(t.core_global, -oy, -ox);
      t.update_local = expanded(t.core_local, guard);
      t.num_neighbors = 0;
    }
  const int dr[4] = {0, 0, -1, 1}, dc[4] = {-1, 1, 0, 0};
  for (litho_tile& t : tiles)
    for (int i = 0; i < 4; ++i) {
      const int r = t.row + dr[i], c = t.col + dc[i];
      if (r < 0 || r >= grid_rows || c < 0 || c >= grid_cols) continue;
      const litho_tile& nb = tiles[(size_t)r * grid_cols + c];
      litho_neighbor& e = t.neighbors[t.num_neighbors++];
      e.id = nb.id;
      e.direction = i;
      e.guard_global = intersection(expanded(t.core_global, guard), nb.core_global);
      e.my_guard_local = shifted(e.guard_global, -t.tile_global.y0, -t.tile_global.x0);
      e.neighbor_overlap_local = shifted(e.guard_global, -nb.tile_global.y0, -nb.tile_global.x0);
      e.overlap_global = intersection(t.core_global, expanded(nb.core_global, guard));
      e.my_overlap_local = shifted(e.overlap_global, -t.tile_global.y0, -t.tile_global.x0);
      e.neighbor_guard_local = shifted(e.overlap_global, -nb.tile_global.y0, -nb.tile_global.x0);
    }
}

void validate_tiles(const litho_tile* tiles, int num_tiles, int n) {
  if (tiles == nullptr || num_tiles <= 0) throw Error(1, "no Micros");
  for (int i = 0; i < num_tiles; ++i) {
    const litho_tile& t = tiles[i];
    if (t.id != i) throw Error(1, "Micro ids must equal their index");
    if (t.tile_global.y1 - t.tile_global.y0 != n || t.tile_global.x1 - t.tile_global.x0 != n)
      throw Error(1, "Micro size differs from the plan tile size (cache.N == meta.tile_size)");
    if (!inside_tile(t.core_local, n) || !inside_tile(t.update_local, n))
      throw Error(1, "Micro Core / update window outside the tile");
    if (t.num_neighbors < 0 || t.num_neighbors > 4) throw Error(1, "a Micro has at most four neighbours");
    for (int j = 0; j < t.num_neighbors; ++j) {
      const litho_neighbor& e = t.neighbors[j];
      if (e.id < 0 || e.id >= num_tiles || e.id == i) throw Error(1, "neighbour id out of range");
      if (e.direction < 0 || e.direction > 3) throw Error(1, "neighbour direction out of range");
      const litho_rect* rs[4] = {&e.my_guard_local, &e.neighbor_overlap_local, &e.my_overlap_local,
                                 &e.neighbor_guard_local};
      for (const litho_rect* r : rs)
        if (area(*r) > 0 && !inside_tile(*r, n)) throw Error(1, "neighbour strip outside the tile");
      if (area(e.my_guard_local) != area(e.neighbor_overlap_local) ||
          e.my_guard_local.y1 - e.my_guard_local.y0 != e.neighbor_overlap_local.y1 - e.neighbor_overlap_local.y0 ||
          area(e.my_overlap_local) != area(e.neighbor_guard_local) ||
          e.my_overlap_local.y1 - e.my_overlap_local.y0 != e.neighbor_guard_local.y1 - e.neighbor_guard_local.y0)
        throw Error(1, "neighbour strips of different shapes");
    }
  }
}

// ------------------------------------------------------------------ exchange
Exchanger::Exchanger(Plan& plan, const litho_tile* tiles, int num_tiles)
    : plan_(plan), T_(num_tiles), n_(plan.n), tiles_(tiles, tiles + num_tiles) {
  validate_tiles(tiles, num_tiles, n_);
  for (const litho_tile& t : tiles_) {
    max_core_ = std::max(max_core_, area(t.core_local));
    for (int j = 0; j < t.num_neighbors; ++j)
      max_strip_ = std::max({max_strip_, area(t.neighbors[j].my_guard_local), area(t.neighbors[j].my_overlap_local)});
  }
  const std::vector<int> tab = tile_table(tiles, num_tiles);
  const size_t plane = (size_t)n_ * n_;
  table_.alloc(tab.size() * sizeof(int));
  pe_win_.alloc((size_t)T_ * plane * sizeof(float));
  pen_win_.alloc((size_t)T_ * plane * sizeof(float));
  raw_.alloc((size_t)T_ * plane * sizeof(float));
  guard_part_.alloc((size_t)T_ * 4 * kExchangeBlocks * sizeof(double));
  core_part_.alloc((size_t)T_ * kMetricBlocks * kCoreMetrics * sizeof(double));
  red_all_.alloc((size_t)T_ * 4 * sizeof(double));
  cept_red_.alloc((size_t)T_ * sizeof(double));
  core_red_.alloc((size_t)T_ * kCoreMetrics * sizeof(double));
  host_ = new HostPinned((size_t)T_ * (5 + kCoreMetrics) * sizeof(double));
  LITHO_HIP(hipMemcpyAsync(table_.ptr, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, plan_.ctx->stream));
  LITHO_HIP(hipStreamSynchronize(plan_.ctx->stream));
}

void Exchanger::prepare(const float* local_targets, const litho_exchange_options& o) {
  check_options(o);
  if (local_targets == nullptr) throw Error(1, "MicroGradientExchange: target is required");
  targets_ = local_targets;
  launch_windows(local_targets, table(), T_, n_, o.loss_region, o.loss_type, pe_win_.as<float>(), pen_win_.as<float>(),
                 plan_.ctx->stream);
  LITHO_HIP(hipGetLastError());
}

void Exchanger::run(const float* masks, const litho_exchange_options& o, float* mask_gradient, float* raw_out,
                    float* wafer_out, litho_tile_loss* per_tile, double totals[3]) {
  if (targets_ == nullptr) throw Error(5, "MicroGradientExchange: prepare() first");
  if (masks == nullptr || mask_gradient == nullptr) throw Error(1, "MicroGradientExchange: null mask or gradient");
  check_options(o);
  hipStream_t s = plan_.ctx->stream;
  const size_t plane = (size_t)n_ * n_;
  const double pe_w = o.pe_weight;
  const double pen_w = o.loss_type == 1 ? o.penalty_weight : 0.0;   // micro_loss_weights (:95-99)
  float* raw = raw_out != nullptr ? raw_out : raw_.as<float>();
  float* wafer = wafer_out;
  if (wafer == nullptr && o.collect_core_metrics) {
    if (wafer_.bytes < (size_t)T_ * plane * sizeof(float)) wafer_.alloc((size_t)T_ * plane * sizeof(float));
    wafer = wafer_.as<float>();
  }
  // phase 1: imaging, region losses and adjoint of every Micro, in chunks of the plan capacity
  for (int c0 = 0; c0 < T_; c0 += plan_.capacity) {
    const int cnt = std::min(plan_.capacity, T_ - c0);
    const size_t off = (size_t)c0 * plane;
    EvalRequest rq;
    rq.mode = 0;
    rq.base = masks + off;
    rq.im.thr = (float)o.threshold;
    rq.im.alpha = (float)o.alpha;
    rq.im.wafer = wafer != nullptr ? wafer + off : nullptr;
    rq.im.target = targets_ + off;
    rq.im.pe_win = pe_win_.as<float>() + off;
    rq.im.pen_win = pen_win_.as<float>() + off;
    rq.im.pe_weight = (float)pe_w;
    rq.im.penalty = (float)pen_w;
    rq.im.pen_thr = (float)o.penalty_threshold;
    rq.ga.scale = (float)(4.0 * o.alpha);
    rq.ga.grad_mask = raw + off;
    run_eval(plan_, cnt, rq);
    LITHO_HIP(hipMemcpyAsync(red_all_.as<double>() + (size_t)c0 * 4, plan_.red.ptr, (size_t)cnt * 4 * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
  }
  // phase 2: mask_gradient = raw, then Guard fusion / CEPT per neighbour strip
  LITHO_HIP(hipMemcpyAsync(mask_gradient, raw, (size_t)T_ * plane * sizeof(float), hipMemcpyDeviceToDevice, s));
  ExchangeParams prm;
  prm.fusion = o.enable_grad_fusion ? 1 : 0;
  prm.linear = o.fusion_mode == 1 ? 1 : 0;
  prm.cept = o.enable_cept ? 1 : 0;
  prm.w_owner = o.fusion_weight_owner;
  prm.w_neighbor = o.fusion_weight_neighbor;
  prm.cept_weight = o.cept_weight;
  const bool any_strip = prm.fusion || prm.cept;
  if (any_strip) {
    launch_guard_exchange(table(), T_, n_, masks, raw, mask_gradient, prm, max_strip_, guard_part_.as<double>(), s);
    if (prm.cept && !o.cept_guard_only)
      for (int slot = 0; slot < 4; ++slot)
        launch_overlap_cept(table(), T_, n_, slot, masks, mask_gradient, o.cept_weight, max_strip_, s);
    launch_reduce_fixed(guard_part_.as<double>(), 4 * kExchangeBlocks, 1, cept_red_.as<double>(), 1, T_, s);
  }
  if (o.collect_core_metrics) {
    launch_core_metrics(table(), T_, n_, wafer, targets_, pe_win_.as<float>(), pen_win_.as<float>(),
                        o.penalty_threshold, max_core_, core_part_.as<double>(), s);
    launch_reduce_fixed(core_part_.as<double>(), kMetricBlocks, kCoreMetrics, core_red_.as<double>(), kCoreMetrics,
                        T_, s);
  }
  LITHO_HIP(hipGetLastError());
  // FP64 scalars of every Micro: one transfer
  double* h = host_->as<double>();
  LITHO_HIP(hipMemcpyAsync(h, red_all_.ptr, (size_t)T_ * 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  if (any_strip)
    LITHO_HIP(hipMemcpyAsync(h + (size_t)T_ * 4, cept_red_.ptr, (size_t)T_ * sizeof(double), hipMemcpyDeviceToHost, s));
  if (o.collect_core_metrics)
    LITHO_HIP(hipMemcpyAsync(h + (size_t)T_ * 5, core_red_.ptr, (size_t)T_ * kCoreMetrics * sizeof(double),
                             hipMemcpyDeviceToHost, s));
  LITHO_HIP(hipStreamSynchronize(s));
  double loss = 0.0, data = 0.0, cept = 0.0;
  for (int t = 0; t < T_; ++t) {
    litho_tile_loss r{};
    r.pe_loss = h[4 * t + 0];
    r.bg_penalty_raw = (o.loss_type == 1 || o.collect_core_metrics) ? h[4 * t + 1] : 0.0;
    r.penalty_loss = pen_w * r.bg_penalty_raw;
    r.data_loss = pe_w * r.pe_loss + r.penalty_loss;
    r.cept_loss = prm.cept ? h[(size_t)T_ * 4 + t] : 0.0;
    if (o.collect_core_metrics) {
      const double* cm = h + (size_t)T_ * 5 + (size_t)kCoreMetrics * t;
      r.core_pe_loss = cm[0];
      r.core_bg_penalty_raw = cm[1];
      // CoreLossMetrics (micro_gradient_exchange.cpp:225-243)
      r.core_pe_penalty_loss = o.pe_weight * cm[2] + o.penalty_weight * cm[1];
      r.core_data_loss = pe_w * cm[0] + pen_w * cm[1];
    }
    r.loss = r.data_loss + (prm.cept ? o.cept_weight * r.cept_loss : 0.0);
    loss += r.loss;
    data += r.data_loss;
    cept += r.cept_loss;
    if (per_tile != nullptr) per_tile[t] = r;
  }
  if (totals != nullptr) {
    totals[0] = loss;
    totals[1] = data;
    totals[2] = cept;
  }
}

// ------------------------------------------------------------------ drivers
namespace {

struct ChipBuffers {
  DevMem targets, masks, grad, state, best_masks, best_state, global;
};

void alloc_tiles(DevMem& m, int T, int n) { m.alloc((size_t)T * n * n * sizeof(float)); }

// One history row from the exchange's per-Micro scalars, summed in Micro
// order (full_chip_ctm.cpp:84-104, full_chip_lsm.cpp:152-182).
litho_fullchip_row history_row(int it, const double tot[3], const std::vector<litho_tile_loss>& per,
                               const litho_exchange_options& o, double min_dt, double max_dt) {
  litho_fullchip_row r{};
  r.iteration = it;
  r.loss = tot[0];
  r.data_loss = tot[1];
  r.cept_loss = tot[2];
  r.min_dt = min_dt;
  r.max_dt = max_dt;
  for (const litho_tile_loss& l : per) {
    r.pe_contribution += o.pe_weight * l.pe_loss;
    r.penalty_contribution += l.penalty_loss;
    r.pe_loss_raw += l.pe_loss;
    r.bg_penalty_raw += l.bg_penalty_raw;
    r.core_pe += l.core_pe_loss;
    r.core_bg_penalty_raw += l.core_bg_penalty_raw;
    r.core_pe_penalty += l.core_pe_penalty_loss;
    r.core_data_loss += l.core_data_loss;
  }
  r.cept_contribution = o.enable_cept ? o.cept_weight * tot[2] : 0.0;
  return r;
}

void record_best(int it, double loss, FullChipResult& res, bool& have, const float* masks, const float* state,
                 float* best_masks, float* best_state, size_t bytes, hipStream_t s) {
  if (have && !(loss < res.best_loss)) return;     // ties keep the first state (full_chip_ctm.h:33)
  have = true;
  res.best_iteration = it;
  res.best_loss = loss;
  LITHO_HIP(hipMemcpyAsync(best_masks, masks, bytes, hipMemcpyDeviceToDevice, s));
  if (best_state != nullptr) LITHO_HIP(hipMemcpyAsync(best_state, state, bytes, hipMemcpyDeviceToDevice, s));
}

void check_chip(const Plan& plan, const litho_tile* tiles, int T, int height, int width, const float* a,
                const float* b) {
  if (height <= 0 || width <= 0) throw Error(1, "chip size must be positive");
  if (a == nullptr || b == nullptr) throw Error(1, "initial pattern and target are required");
  validate_tiles(tiles, T, plan.n);
  for (int i = 0; i < T; ++i) {
    const litho_rect& c = tiles[i].core_global;
    if (c.y0 < 0 || c.x0 < 0 || c.y1 > height || c.x1 > width) throw Error(1, "Micro Core outside the chip");
  }
}

}  // namespace

namespace {

// Imaging::compute of T local tiles in batches of the plan capacity.
void image_tiles(Plan& plan, int T, const float* masks, double threshold, double alpha, float* aerial,
                 float* wafer) {
  const size_t plane = plan.tile_elems();
  for (int c = 0; c < T; c += plan.capacity) {
    const int b = std::min(plan.capacity, T - c);
    EvalRequest rq;
    rq.mode = 0;
    rq.base = masks + (size_t)c * plane;
    rq.adjoint = false;
    rq.im.thr = (float)threshold;
    rq.im.alpha = (float)alpha;
    rq.im.aerial = aerial + (size_t)c * plane;
    rq.im.wafer = wafer + (size_t)c * plane;
    run_eval(plan, b, rq);
  }
}

int max_core_elems(const litho_tile* tiles, int T) {
  int m = 0;
  for (int i = 0; i < T; ++i) m = std::max(m, area(tiles[i].core_local));
  return m;
}

// Loss::pe_loss (loss.cpp:16-23) of two device images, FP64, fixed order.
double pe_loss_device(const float* wafer, const float* target, size_t count, hipStream_t s) {
  constexpr int kBlocks = 1024;
  DevMem part(kBlocks * sizeof(double)), out(sizeof(double));
  launch_sq_diff_partials(wafer, target, count, part.as<double>(), kBlocks, s);
  launch_reduce_fixed(part.as<double>(), kBlocks, 1, out.as<double>(), 1, 1, s);
  double v = 0.0;
  LITHO_HIP(hipMemcpyAsync(&v, out.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
  LITHO_HIP(hipStreamSynchronize(s));
  return v;
}

}  // namespace

void micro_imaging(Plan& plan, const litho_tile* tiles, int T, int height, int width, const float* global_mask,
                   double threshold, double alpha, float* global_aerial, float* global_wafer, float* local_aerial,
                   float* local_wafer) {
  check_chip(plan, tiles, T, height, width, global_mask, global_mask);
  if (!std::isfinite(threshold) || !std::isfinite(alpha)) throw Error(1, "Imaging: threshold and alpha must be finite");
  const int n = plan.n;
  hipStream_t s = plan.ctx->stream;
  const std::vector<int> tab = tile_table(tiles, T);
  DevMem table(tab.size() * sizeof(int)), masks, la, lw;
  LITHO_HIP(hipMemcpyAsync(table.ptr, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, s));
  alloc_tiles(masks, T, n);
  float* aerial = local_aerial;
  float* wafer = local_wafer;
  if (aerial == nullptr) {
    alloc_tiles(la, T, n);
    aerial = la.as<float>();
  }
  if (wafer == nullptr) {
    alloc_tiles(lw, T, n);
    wafer = lw.as<float>();
  }
  const int max_core = max_core_elems(tiles, T);
  // every Micro reads the global mask with periodic Halo indices (mirco_imaging.cpp:15-17)
  launch_extract(global_mask, height, width, table.as<int>(), T, n, masks.as<float>(), s);
  image_tiles(plan, T, masks.as<float>(), threshold, alpha, aerial, wafer);
  // each Core writes the part of the chip it owns (mirco_imaging.cpp:23-25)
  if (global_aerial != nullptr) launch_assemble(aerial, table.as<int>(), T, n, width, global_aerial, max_core, s);
  if (global_wafer != nullptr) launch_assemble(wafer, table.as<int>(), T, n, width, global_wafer, max_core, s);
  LITHO_HIP(hipGetLastError());
  LITHO_HIP(hipStreamSynchronize(s));
}

MicroComparison compare_micro_imaging(Plan& plan, const litho_tile* tiles, int T, int height, int width,
                                      const float* global_target, const float* optimized_local_masks,
                                      double threshold, double alpha, const MicroComparisonOutputs& out) {
  check_chip(plan, tiles, T, height, width, global_target, optimized_local_masks);
  if (!std::isfinite(threshold) || !std::isfinite(alpha)) throw Error(1, "Imaging: threshold and alpha must be finite");
  const int n = plan.n;
  const size_t chip = (size_t)height * width;
  hipStream_t s = plan.ctx->stream;
  const std::vector<int> tab = tile_table(tiles, T);
  DevMem table(tab.size() * sizeof(int)), ta, tw, oa, ow, la, lw;
  LITHO_HIP(hipMemcpyAsync(table.ptr, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, s));
  // the global images every PE needs, in the caller's buffers or here; the
  // Cores tile the chip, and pixels outside every Core count as 0
  auto global_buf = [&](DevMem& m, float* given) {
    if (given != nullptr) {
      LITHO_HIP(hipMemsetAsync(given, 0, chip * sizeof(float), s));
      return given;
    }
    m.alloc(chip * sizeof(float));
    LITHO_HIP(hipMemsetAsync(m.ptr, 0, chip * sizeof(float), s));
    return m.as<float>();
  };
  float* t_ai = global_buf(ta, out.target_aerial);
  float* t_wi = global_buf(tw, out.target_wafer);
  float* o_ai = global_buf(oa, out.optimized_aerial);
  float* o_wi = global_buf(ow, out.optimized_wafer);
  // the target: every Micro's periodic extract imaged (full_chip_imaging_compare.cpp:51-52)
  micro_imaging(plan, tiles, T, height, width, global_target, threshold, alpha, t_ai, t_wi, nullptr, nullptr);
  // the optimized Micros as given, Guard and Halo included (:53)
  float* l_ai = out.optimized_local_aerial;
  if (l_ai == nullptr) {
    alloc_tiles(la, T, n);
    l_ai = la.as<float>();
  }
  alloc_tiles(lw, T, n);
  image_tiles(plan, T, optimized_local_masks, threshold, alpha, l_ai, lw.as<float>());
  const int max_core = max_core_elems(tiles, T);
  launch_assemble(l_ai, table.as<int>(), T, n, width, o_ai, max_core, s);
  launch_assemble(lw.as<float>(), table.as<int>(), T, n, width, o_wi, max_core, s);
  if (out.optimized_mask != nullptr) {
    LITHO_HIP(hipMemsetAsync(out.optimized_mask, 0, chip * sizeof(float), s));
    launch_assemble(optimized_local_masks, table.as<int>(), T, n, width, out.optimized_mask, max_core, s);
  }
  LITHO_HIP(hipGetLastError());
  MicroComparison r;
  // Loss::pe_loss of the Core-assembled wafer images against the global target (:62-63)
  r.target_pe = pe_loss_device(t_wi, global_target, chip, s);
  r.optimized_pe = pe_loss_device(o_wi, global_target, chip, s);
  return r;
}

litho_mask_continuity mask_continuity(Context& ctx, const litho_tile* tiles, int T, int tile_size, int height,
                                      int width, const float* local_masks, const float* target) {
  if (height <= 0 || width <= 0) throw Error(1, "chip size must be positive");
  if (local_masks == nullptr || target == nullptr) throw Error(1, "mask continuity: null masks or target");
  validate_tiles(tiles, T, tile_size);
  for (int i = 0; i < T; ++i) {
    const litho_rect& c = tiles[i].core_global;
    if (c.y0 < 0 || c.x0 < 0 || c.y1 > height || c.x1 > width) throw Error(1, "Micro Core outside the chip");
  }
  hipStream_t s = ctx.stream;
  const std::vector<int> tab = tile_table(tiles, T);
  DevMem table(tab.size() * sizeof(int)), global((size_t)height * width * sizeof(float));
  const size_t np = (size_t)T * 4 * kExchangeBlocks * kContinuityValues;
  DevMem part(np * sizeof(double));
  LITHO_HIP(hipMemcpyAsync(table.ptr, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, s));
  LITHO_HIP(hipMemsetAsync(global.ptr, 0, global.bytes, s));
  launch_assemble(local_masks, table.as<int>(), T, tile_size, width, global.as<float>(), max_core_elems(tiles, T), s);
  launch_mask_continuity(table.as<int>(), T, tile_size, width, local_masks, global.as<float>(), target,
                         part.as<double>(), s);
  LITHO_HIP(hipGetLastError());
  std::vector<double> h(np);
  LITHO_HIP(hipMemcpyAsync(h.data(), part.ptr, np * sizeof(double), hipMemcpyDeviceToHost, s));
  LITHO_HIP(hipStreamSynchronize(s));
  // fixed order: Micros, neighbour slots, blocks
  double v[kContinuityValues] = {0};
  for (size_t r = 0; r < (size_t)T * 4 * kExchangeBlocks; ++r)
    for (int i = 0; i < kContinuityValues; ++i) {
      const double x = h[r * kContinuityValues + i];
      v[i] = i == 3 ? std::max(v[i], x) : v[i] + x;
    }
  litho_mask_continuity m{};
  m.overlap_samples = (long long)v[0];
  m.overlap_sse = v[1];
  m.overlap_max_abs = v[3];
  if (m.overlap_samples > 0) {
    m.overlap_rmse = std::sqrt(v[1] / v[0]);
    m.overlap_mae = v[2] / v[0];
  }
  m.seam_samples = (long long)v[4];
  m.flat_seam_samples = (long long)v[7];
  m.foreground_seam_samples = (long long)v[9];
  if (m.seam_samples > 0) {
    m.seam_jump_mae = v[5] / v[4];
    m.seam_reference_error_mae = v[6] / v[4];
  }
  if (m.flat_seam_samples > 0) m.flat_seam_jump_mae = v[8] / v[7];
  if (m.foreground_seam_samples > 0) m.foreground_seam_jump_mae = v[10] / v[9];
  return m;
}

FullChipResult fullchip_ctm(Plan& plan, const litho_tile* tiles, int T, int height, int width,
                            const float* initial_pattern, const float* global_target,
                            const litho_exchange_options& options, int iterations, double learning_rate,
                            float* global_mask_out, float* local_masks_out, litho_fullchip_row* history) {
  check_chip(plan, tiles, T, height, width, initial_pattern, global_target);
  if (iterations < 0) throw Error(1, "FullChipCTMOptions: iterations must be non-negative");
  if (!std::isfinite(learning_rate)) throw Error(1, "FullChipCTMOptions: learning_rate must be finite");
  litho_exchange_options o = options;
  o.cept_guard_only = 1;                        // full_chip_ctm.h:12
  const int n = plan.n;
  const size_t plane = (size_t)n * n, bytes = (size_t)T * plane * sizeof(float);
  hipStream_t s = plan.ctx->stream;
  Exchanger ex(plan, tiles, T);
  ChipBuffers b;
  alloc_tiles(b.targets, T, n);
  alloc_tiles(b.masks, T, n);
  alloc_tiles(b.grad, T, n);
  alloc_tiles(b.state, T, n);
  alloc_tiles(b.best_masks, T, n);
  b.global.alloc((size_t)height * width * sizeof(float));
  float* theta = b.state.as<float>();
  float* masks = b.masks.as<float>();
  // theta0 = initialize_theta(initial_pattern) on the chip, then the periodic gather
  launch_init_theta(initial_pattern, b.global.as<float>(), (size_t)height * width, s);
  launch_extract(b.global.as<float>(), height, width, ex.table(), T, n, theta, s);
  launch_mask_from_theta_tiles(theta, (size_t)T * plane, masks, s);
  launch_extract(global_target, height, width, ex.table(), T, n, b.targets.as<float>(), s);
  ex.prepare(b.targets.as<float>(), o);
  FullChipResult res;
  bool have = false;
  int max_update = 0;
  for (int i = 0; i < T; ++i) max_update = std::max(max_update, area(tiles[i].update_local));
  double tot[3];
  std::vector<litho_tile_loss> per(T);
  for (int it = 0; it <= iterations; ++it) {
    if (it > 0)   // full_chip_ctm.cpp:133-141: theta step on Core + Guard
      launch_theta_step(ex.table(), T, n, b.grad.as<float>(), learning_rate, theta, masks, max_update, s);
    ex.run(masks, o, b.grad.as<float>(), nullptr, nullptr, history != nullptr ? per.data() : nullptr, tot);
    if (history != nullptr) history[it] = history_row(it, tot, per, o, 0.0, 0.0);
    record_best(it, tot[0], res, have, masks, nullptr, b.best_masks.as<float>(), nullptr, bytes, s);
  }
  if (global_mask_out != nullptr)
    launch_assemble(b.best_masks.as<float>(), ex.table(), T, n, width, global_mask_out, (int)plane, s);
  if (local_masks_out != nullptr)
    LITHO_HIP(hipMemcpyAsync(local_masks_out, b.best_masks.ptr, bytes, hipMemcpyDeviceToDevice, s));
  LITHO_HIP(hipGetLastError());
  LITHO_HIP(hipStreamSynchronize(s));
  return res;
}

FullChipResult fullchip_lsm(Plan& plan, const litho_tile* tiles, int T, int height, int width,
                            const float* initial_binary, const float* global_target,
                            const litho_exchange_options& options, const LsmOptions& lo, int iterations,
                            int reinit_interval, float* global_mask_out, float* global_phi_out,
                            litho_fullchip_row* history) {
  check_chip(plan, tiles, T, height, width, initial_binary, global_target);
  if (iterations < 0) throw Error(1, "FullChipLSMOptions: iterations must be non-negative");
  if (reinit_interval < 0) throw Error(1, "FullChipLSMOptions: reinit_interval must be non-negative");
  if (!(lo.dx > 0.0) || !(lo.dy > 0.0)) throw Error(1, "FullChipLSMOptions: dx/dy must be positive");
  if (lo.mask_type != 0 && lo.mask_type != 1) throw Error(1, "FullChipLSMOptions: mask_type must be Binary or Gray");
  if (lo.mask_type == 1 && lo.render_scale <= 0) throw Error(1, "FullChipLSMOptions: render_scale must be positive");
  litho_exchange_options o = options;
  o.loss_region = 1;                            // Core + Guard (full_chip_lsm.h:22)
  o.cept_guard_only = 1;
  o.collect_core_metrics = 1;
  const int n = plan.n;
  const size_t plane = (size_t)n * n, bytes = (size_t)T * plane * sizeof(float);
  hipStream_t s = plan.ctx->stream;
  Context& ctx = *plan.ctx;
  Exchanger ex(plan, tiles, T);
  ChipBuffers b;
  alloc_tiles(b.targets, T, n);
  alloc_tiles(b.masks, T, n);
  alloc_tiles(b.grad, T, n);
  alloc_tiles(b.state, T, n);
  alloc_tiles(b.best_masks, T, n);
  alloc_tiles(b.best_state, T, n);
  float* phi = b.state.as<float>();
  float* masks = b.masks.as<float>();
  // phi0 = reinit_sdf(binary - 0.5) per Micro (full_chip_lsm.cpp:141-146): inside = binary > 0.5
  launch_extract(initial_binary, height, width, ex.table(), T, n, b.grad.as<float>(), s);
  {
    char* ws = static_cast<char*>(ctx.workspace(chamfer_scratch_bytes(T, n, n)));
    uint32_t* scratch = reinterpret_cast<uint32_t*>(ws);
    float* dist = reinterpret_cast<float*>(ws + (size_t)T * 2 * plane * sizeof(uint32_t));
    launch_reinit_sdf(b.grad.as<float>(), T, n, n, lo.dx, lo.dy, 0.5f, false, scratch, dist, nullptr, phi, masks, s);
    if (lo.mask_type == 1) launch_render_mask(phi, T, n, n, 1, lo.render_scale, masks, s);
  }
  launch_extract(global_target, height, width, ex.table(), T, n, b.targets.as<float>(), s);
  ex.prepare(b.targets.as<float>(), o);
  std::vector<int> rects((size_t)T * 4);
  for (int i = 0; i < T; ++i) {
    const litho_rect& u = tiles[i].update_local;
    rects[4 * i + 0] = u.y0;
    rects[4 * i + 1] = u.y1;
    rects[4 * i + 2] = u.x0;
    rects[4 * i + 3] = u.x1;
  }
  std::vector<double> dt(T, 0.0);
  FullChipResult res;
  bool have = false;
  double tot[3];
  std::vector<litho_tile_loss> per(T);
  for (int it = 0; it <= iterations; ++it) {
    double mn = 0.0, mx = 0.0;
    if (it > 0) {
      const bool reinit = reinit_interval > 0 && it % reinit_interval == 0;
      lsm_step(ctx, T, n, phi, b.grad.as<float>(), rects.data(), lo, reinit, phi, masks, dt.data());
      mn = *std::min_element(dt.begin(), dt.end());
      mx = *std::max_element(dt.begin(), dt.end());
    }
    ex.run(masks, o, b.grad.as<float>(), nullptr, nullptr, history != nullptr ? per.data() : nullptr, tot);
    if (history != nullptr) history[it] = history_row(it, tot, per, o, mn, mx);
    record_best(it, tot[0], res, have, masks, phi, b.best_masks.as<float>(), b.best_state.as<float>(), bytes, s);
  }
  if (global_mask_out != nullptr)
    launch_assemble(b.best_masks.as<float>(), ex.table(), T, n, width, global_mask_out, (int)plane, s);
  if (global_phi_out != nullptr)
    launch_assemble(b.best_state.as<float>(), ex.table(), T, n, width, global_phi_out, (int)plane, s);
  LITHO_HIP(hipGetLastError());
  LITHO_HIP(hipStreamSynchronize(s));
  return res;
}

}  // namespace litho
