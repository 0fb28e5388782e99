// litho_cpp.hpp — the reference's C++ call surface over the C ABI (litho_amd.h).
//
// Header-only.  Class names, argument meaning, defaults and error behaviour
// follow wyhcoder/C-version-Lithography's Imaging / Gradient / Loss /
// CTM_Optimizer / Partition / MicroGradientExchange / FullChipCTMOptimizer /
// FullChipLSMOptimizer (imaging.h:16-39, gradient.h:12-55, loss.h:51-56,
// CTM_Optimizer.h:24-41, partition.h:67-86, micro_gradient_exchange.h:19-130,
// full_chip_ctm.h:10-58, full_chip_lsm.h:8-66), so the reference's callers and
// tests read the same.  Differences, all from the device design:
//  * litho::Matrix (row-major double) stands in for Eigen::MatrixXd;
//  * the device never stores the K electric fields: Gradient takes the mask
//    the image came from instead of Imaging::get_electric_field();
//  * LithoPrepare takes the optics as numbers (OpticsSpec) and K explicitly;
//  * tile sizes without a compiled device FFT (litho_supported_size) run the
//    direct-DFT pipeline, up to N = 4096.
// Invalid arguments raise std::invalid_argument, other failures
// std::runtime_error, with the library's message.
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "litho_amd.h"

namespace litho {

// ------------------------------------------------------------------ Matrix
class Matrix {
 public:
  Matrix() = default;
  Matrix(int rows, int cols, double value = 0.0) : r_(rows), c_(cols), d_((size_t)rows * cols, value) {}
  static Matrix Constant(int rows, int cols, double value) { return Matrix(rows, cols, value); }
  static Matrix Zero(int rows, int cols) { return Matrix(rows, cols, 0.0); }
  int rows() const { return r_; }
  int cols() const { return c_; }
  size_t size() const { return d_.size(); }
  double& operator()(int r, int c) { return d_[(size_t)r * c_ + c]; }
  double operator()(int r, int c) const { return d_[(size_t)r * c_ + c]; }
  double* data() { return d_.data(); }
  const double* data() const { return d_.data(); }
  double sum() const {
    double s = 0.0;
    for (double v : d_) s += v;
    return s;
  }
  double squaredNorm() const {
    double s = 0.0;
    for (double v : d_) s += v * v;
    return s;
  }
  double norm() const { return std::sqrt(squaredNorm()); }
  bool allFinite() const {
    for (double v : d_)
      if (!std::isfinite(v)) return false;
    return true;
  }

 private:
  int r_ = 0, c_ = 0;
  std::vector<double> d_;
};

namespace detail {

inline void check(int rc) {
  if (rc == LITHO_OK) return;
  const std::string m = litho_last_error();
  if (rc == LITHO_EINVAL) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

inline void hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device float32 buffer of n values.
class Buffer {
 public:
  explicit Buffer(size_t n) : n_(n) { hip(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(float)), "hipMalloc"); }
  ~Buffer() {
    if (p_ != nullptr) (void)hipFree(p_);
  }
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  float* get() const { return p_; }
  void upload(const double* src, size_t n, size_t offset = 0) {
    std::vector<float> h(src, src + n);
    hip(hipMemcpy(p_ + offset, h.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }
  void upload(const Matrix& m, size_t offset = 0) { upload(m.data(), m.size(), offset); }
  void fill(double v) {
    std::vector<float> h(n_, (float)v);
    hip(hipMemcpy(p_, h.data(), n_ * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }
  Matrix download(int rows, int cols, size_t offset = 0) const {
    std::vector<float> h((size_t)rows * cols);
    hip(hipMemcpy(h.data(), p_ + offset, h.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    Matrix m(rows, cols);
    for (size_t i = 0; i < h.size(); ++i) m.data()[i] = h[i];
    return m;
  }

 private:
  float* p_ = nullptr;
  size_t n_ = 0;
};

inline void same_shape(const Matrix& a, const Matrix& b, const char* what) {
  if (a.rows() != b.rows() || a.cols() != b.cols()) throw std::invalid_argument(what);
}

// Context + plan of one tile size, shared by the objects built from a cache.
struct PlanState {
  litho_ctx* ctx = nullptr;
  litho_plan* plan = nullptr;
  ~PlanState() {
    if (plan != nullptr) (void)litho_plan_destroy(plan);
    if (ctx != nullptr) (void)litho_ctx_destroy(ctx);
  }
};

}  // namespace detail

// --------------------------------------------------------- optics / cache
struct OpticsSpec {                 // SimulationParameters system / optics / source
  double pixel_size_nm = 4.0;
  double wavelength_nm = 193.0;
  double na = 1.35;
  double sigma_in = 0.6;            // annular source
  double sigma_out = 0.9;
  // Pupil::Params::zernike_coeffs: {fringe index 1..11, coefficient in waves},
  // summed in this order (pupil.cpp:5-46, zernike.cpp:46-53)
  std::vector<std::pair<int, double>> zernike_coeffs;
};

// The part of ImagingCache (litho_prepare.h:18-47) the hot path reads, resident
// on the device: the SOCS kernels' frequency window and eigenvalues.
struct ImagingCache {
  int N = 0;
  int K = 0;
  int win_rows = 0, win_cols = 0, win_y0 = 0, win_x0 = 0;
  double source_weight_sum = 0.0;
  std::vector<double> socs_eigenvalues;
  std::shared_ptr<detail::PlanState> device;
  litho_plan* plan() const { return device->plan; }
  litho_ctx* ctx() const { return device->ctx; }
  int capacity = 1;
};

// LithoPrepare(grid, pupil, source, use_socs = true, compact_socs = true)
// (litho_prepare.cpp:16-54,225-385) with K as a parameter; the SOCS kernels
// are computed on the device in FP64.  Like the reference, K is clamped to the
// decomposition's rank, min(K, Ns, window pixels), and K <= 0 keeps all Ns
// (litho_prepare.cpp:302-304).  batch_capacity: tiles per batched call (the
// number of Micros for the full-chip optimizers).
class LithoPrepare {
 public:
  LithoPrepare(int n, const OpticsSpec& optics = {}, int num_kernels = 36, int batch_capacity = 1, int device = 0) {
    c_.N = n;
    c_.capacity = batch_capacity;
    int npts = 0;
    detail::check(litho_socs_window(n, optics.pixel_size_nm, optics.wavelength_nm, optics.na, optics.sigma_in,
                                    optics.sigma_out, &c_.win_rows, &c_.win_cols, &c_.win_y0, &c_.win_x0, &npts,
                                    &c_.source_weight_sum));
    num_kernels = std::min({num_kernels > 0 ? num_kernels : npts, npts, c_.win_rows * c_.win_cols});
    c_.K = num_kernels;
    c_.device = std::make_shared<detail::PlanState>();
    detail::check(litho_ctx_create(device, &c_.device->ctx));
    std::vector<double> kern((size_t)num_kernels * c_.win_rows * c_.win_cols * 2);
    c_.socs_eigenvalues.assign(num_kernels, 0.0);
    std::vector<int> zi;
    std::vector<double> zc;
    for (const auto& z : optics.zernike_coeffs) {
      zi.push_back(z.first);
      zc.push_back(z.second);
    }
    detail::check(litho_socs_precompute_zernike(c_.device->ctx, n, optics.pixel_size_nm, optics.wavelength_nm,
                                                optics.na, optics.sigma_in, optics.sigma_out, zi.data(), zc.data(),
                                                (int)zi.size(), num_kernels, kern.data(),
                                                c_.socs_eigenvalues.data(), nullptr));
    detail::check(litho_plan_create(c_.device->ctx, n, num_kernels, c_.win_rows, c_.win_cols, c_.win_y0, c_.win_x0,
                                    kern.data(), c_.socs_eigenvalues.data(), c_.source_weight_sum, batch_capacity,
                                    &c_.device->plan));
  }
  const ImagingCache& cache() const { return c_; }

 private:
  ImagingCache c_;
};

// ------------------------------------------------------------------ Imaging
struct Imaging_Result {
  Matrix aerial_image;
  Matrix wafer_image;
};

class Imaging {
 public:
  explicit Imaging(const ImagingCache& cache) : cache_(cache), m_(sz()), a_(sz()), w_(sz()) {}
  Imaging(const Imaging&) = delete;
  Imaging& operator=(const Imaging&) = delete;
  Imaging_Result compute(const Matrix& mask, double threshold = 0.25, double alpha = 50.0) {
    if (mask.rows() != cache_.N || mask.cols() != cache_.N)
      throw std::invalid_argument("Imaging::compute: mask size must match cache.N");
    if (!mask.allFinite() || !std::isfinite(threshold) || !std::isfinite(alpha))
      throw std::invalid_argument("Imaging: mask, threshold, and alpha must be finite");
    m_.upload(mask);
    detail::check(litho_socs_forward(cache_.plan(), 1, m_.get(), a_.get(), w_.get(), threshold, alpha));
    return {a_.download(cache_.N, cache_.N), w_.download(cache_.N, cache_.N)};
  }

 private:
  size_t sz() const { return (size_t)cache_.N * cache_.N; }
  const ImagingCache& cache_;
  detail::Buffer m_, a_, w_;
};

// ------------------------------------------------------------------ Loss
struct Loss {
  // Loss::pe_loss (loss.cpp:16-23)
  static double pe_loss(const Matrix& wafer, const Matrix& target) {
    detail::same_shape(wafer, target, "pe_loss: size mismatch");
    double s = 0.0;
    for (size_t i = 0; i < wafer.size(); ++i) {
      const double d = wafer.data()[i] - target.data()[i];
      s += d * d;
    }
    return s;
  }
  static double pe_loss(const Matrix& wafer, const Matrix& target, const Matrix& window) {
    detail::same_shape(wafer, target, "pe_loss: size mismatch");
    detail::same_shape(wafer, window, "pe_loss: window size mismatch");
    double s = 0.0;
    for (size_t i = 0; i < wafer.size(); ++i) {
      const double d = wafer.data()[i] - target.data()[i];
      s += d * d * window.data()[i];
    }
    return s;
  }
  // Loss::background_penalty_loss (loss.cpp:53-64)
  static double background_penalty_loss(const Matrix& wafer, const Matrix& target, double penalty_threshold,
                                         const Matrix* window = nullptr) {
    detail::same_shape(wafer, target, "background_penalty_loss: size mismatch");
    double s = 0.0;
    for (size_t i = 0; i < wafer.size(); ++i) {
      if (!(target.data()[i] < 0.5)) continue;
      const double e = std::max(wafer.data()[i] - penalty_threshold, 0.0);
      s += (window != nullptr ? window->data()[i] : 1.0) * e * e;
    }
    return s;
  }
  // Loss::pe_loss_with_penalty (loss.cpp:25-51)
  static double pe_loss_with_penalty(const Matrix& wafer, const Matrix& target, double penalty,
                                     double penalty_threshold) {
    return pe_loss(wafer, target) + penalty * background_penalty_loss(wafer, target, penalty_threshold);
  }
};

// ------------------------------------------------------------------ Gradient
// Gradient::pe_gradient / pe_gradient_with_penalty (gradient.cpp:121-196):
// dPE/dM of the image of `mask` (the reference passes that image's wafer and
// electric fields; the device rebuilds the fields from the mask).
class Gradient {
 public:
  Gradient(const ImagingCache& cache, double threshold, double alpha)
      : cache_(cache), thr_(threshold), alpha_(alpha), m_(sz()), t_(sz()), pw_(sz()), bw_(sz()), g_(sz()) {}
  Matrix pe_gradient(const Matrix& mask, const Matrix& target) { return run(mask, target, nullptr, 1.0, 0.0, 0.01, nullptr); }
  Matrix pe_gradient(const Matrix& mask, const Matrix& target, const Matrix& loss_window) {
    return run(mask, target, &loss_window, 1.0, 0.0, 0.01, nullptr);
  }
  Matrix pe_gradient_with_penalty(const Matrix& mask, const Matrix& target, double penalty,
                                  double penalty_threshold = 0.01) {
    return run(mask, target, nullptr, 1.0, penalty, penalty_threshold, nullptr);
  }
  Matrix pe_gradient_with_penalty(const Matrix& mask, const Matrix& target, double penalty, double penalty_threshold,
                                  const Matrix& pe_window, double pe_weight, const Matrix& penalty_window) {
    return run(mask, target, &pe_window, pe_weight, penalty, penalty_threshold, &penalty_window);
  }
  double last_pe_loss() const { return pe_; }

 private:
  size_t sz() const { return (size_t)cache_.N * cache_.N; }
  Matrix run(const Matrix& mask, const Matrix& target, const Matrix* pe_win, double pe_weight, double penalty,
             double pen_thr, const Matrix* pen_win) {
    const int n = cache_.N;
    if (mask.rows() != n || mask.cols() != n) throw std::invalid_argument("Gradient: mask size must match cache.N");
    detail::same_shape(mask, target, "Gradient: target size mismatch");
    m_.upload(mask);
    t_.upload(target);
    const float* pw = nullptr;
    if (pe_win != nullptr) {
      detail::same_shape(mask, *pe_win, "Gradient: window size mismatch");
      pw_.upload(*pe_win);
      pw = pw_.get();
    }
    const float* bw = nullptr;
    if (penalty != 0.0) {               // the penalty acts on its window (all ones unless given)
      if (pen_win != nullptr) {
        detail::same_shape(mask, *pen_win, "Gradient: penalty window size mismatch");
        bw_.upload(*pen_win);
      } else {
        bw_.fill(1.0);
      }
      bw = bw_.get();
    }
    double pe = 0.0, bg = 0.0;
    detail::check(litho_pe_gradient(cache_.plan(), 1, m_.get(), t_.get(), pw, bw, pe_weight, penalty, pen_thr, thr_,
                                    alpha_, nullptr, nullptr, g_.get(), &pe, &bg));
    pe_ = pe;
    return g_.download(n, n);
  }
  const ImagingCache& cache_;
  double thr_, alpha_, pe_ = 0.0;
  detail::Buffer m_, t_, pw_, bw_, g_;
};

// ------------------------------------------------------------------ CTM
// CTM_Optimizer (CTM_Optimizer.h:24-41): target and resist parameters are
// given directly (the reference reads them from its LithographySimulator).
class CTM_Optimizer {
 public:
  CTM_Optimizer(const ImagingCache& cache, const Matrix& target, int iterations, double learning_rate,
                std::string optimizer_method = "gradient_descent", int lbfgs_history_size = 10,
                double gradient_tolerance = 0.0, double threshold = 0.25, double alpha = 50.0)
      : cache_(cache), target_(target), it_(iterations), lr_(learning_rate), method_(std::move(optimizer_method)),
        m_(lbfgs_history_size), tol_(gradient_tolerance), thr_(threshold), alpha_(alpha) {
    if (method_ != "gradient_descent" && method_ != "lbfgs")
      throw std::invalid_argument("CTM optimizer must be gradient_descent or lbfgs");
    if (target.rows() != cache.N || target.cols() != cache.N)
      throw std::invalid_argument("CTM target size must match cache.N");
  }
  // optimize() returns the best-PE mask (gradient descent) or the final mask (L-BFGS).
  Matrix optimize(bool verbose = false, bool save_intermediate = false) {
    (void)verbose;
    (void)save_intermediate;
    const int n = cache_.N;
    detail::Buffer t((size_t)n * n), mask((size_t)n * n);
    t.upload(target_);
    litho_ctm_options o{method_ == "lbfgs" ? 1 : 0, it_, lr_, m_, tol_, thr_, alpha_};
    litho_ctm_info info{};
    history_.assign(4096, 0.0);
    detail::check(litho_ctm_optimize(cache_.plan(), t.get(), nullptr, &o, mask.get(), nullptr, history_.data(),
                                     (int)history_.size(), &info));
    history_.resize(std::min<size_t>((size_t)info.evaluations, history_.size()));
    info_ = info;
    return mask.download(n, n);
  }
  Matrix optimize_lbfgs(bool verbose = true) {
    method_ = "lbfgs";
    return optimize(verbose, false);
  }
  const std::vector<double>& error_history() const { return history_; }
  const litho_ctm_info& info() const { return info_; }
  // CTM_Optimizer::initialize_theta (CTM_Optimizer.cpp:66-73)
  static Matrix initialize_theta(const Matrix& initial_mask) {
    constexpr double pi = 3.14159265358979323846;
    Matrix th(initial_mask.rows(), initial_mask.cols());
    for (size_t i = 0; i < th.size(); ++i) th.data()[i] = initial_mask.data()[i] > 0.0 ? 0.3 * pi : 0.7 * pi;
    return th;
  }

 private:
  const ImagingCache& cache_;
  Matrix target_;
  int it_;
  double lr_;
  std::string method_;
  int m_;
  double tol_, thr_, alpha_;
  std::vector<double> history_;
  litho_ctm_info info_{};
};

// ------------------------------------------------------------------ full chip
struct PartitionMeta {
  int height = 0, width = 0, grid_rows = 0, grid_cols = 0, halo = 0, guard = 0, tile_size = 0;
  std::vector<litho_tile> tiles;
};

struct Partition {
  // Partition::build (partition.cpp:21-80)
  static PartitionMeta build(int height, int width, int grid_rows, int grid_cols, int halo, int guard) {
    PartitionMeta m{height, width, grid_rows, grid_cols, halo, guard, 0, {}};
    m.tiles.resize((size_t)std::max(grid_rows, 0) * std::max(grid_cols, 0));
    detail::check(litho_partition_build(height, width, grid_rows, grid_cols, halo, guard, &m.tile_size,
                                        m.tiles.data()));
    return m;
  }
  // Partition::extract_local_tile (partition.cpp:83-95): periodic halo gather
  static Matrix extract_local_tile(const Matrix& global, const litho_tile& tile) {
    const int n = tile.tile_global.y1 - tile.tile_global.y0;
    Matrix out(n, n);
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int gy = ((tile.tile_global.y0 + y) % global.rows() + global.rows()) % global.rows();
        const int gx = ((tile.tile_global.x0 + x) % global.cols() + global.cols()) % global.cols();
        out(y, x) = global(gy, gx);
      }
    return out;
  }
};

enum class MicroLossType { PE, PEPenalty };
enum class MicroLossRegion { Core, CoreGuard };
enum class GradientFusionMode { Average, Linear };

struct MicroGradientOptions {       // micro_gradient_exchange.h:19-40 (PE modes)
  MicroLossType loss_type = MicroLossType::PE;
  MicroLossRegion loss_region = MicroLossRegion::Core;
  double pe_weight = 1.0;
  double penalty_weight = 2.0;
  double penalty_threshold = 0.01;
  bool collect_core_metrics = false;
  bool enable_cept = false;
  bool cept_guard_only = false;
  double cept_weight = 1.1;
  bool enable_grad_fusion = false;
  GradientFusionMode fusion_mode = GradientFusionMode::Average;
  double fusion_weight_owner = 0.5;
  double fusion_weight_neighbor = 0.5;
};

namespace detail {
inline litho_exchange_options to_abi(const MicroGradientOptions& o, double thr, double alpha) {
  return litho_exchange_options{o.loss_type == MicroLossType::PEPenalty ? 1 : 0,
                                o.loss_region == MicroLossRegion::CoreGuard ? 1 : 0,
                                o.pe_weight,
                                o.penalty_weight,
                                o.penalty_threshold,
                                o.collect_core_metrics ? 1 : 0,
                                o.enable_cept ? 1 : 0,
                                o.cept_guard_only ? 1 : 0,
                                o.cept_weight,
                                o.enable_grad_fusion ? 1 : 0,
                                o.fusion_mode == GradientFusionMode::Linear ? 1 : 0,
                                o.fusion_weight_owner,
                                o.fusion_weight_neighbor,
                                thr,
                                alpha};
}
inline void check_chip_cache(const PartitionMeta& meta, const ImagingCache& cache) {
  if (cache.N != meta.tile_size) throw std::invalid_argument("cache.N must equal meta.tile_size");
}
}  // namespace detail

// MicroImaging (mirco_imaging.h): every Micro imaged from its periodic extract
// of the global mask, the Cores assembled (mirco_imaging.cpp:9-27).  Built
// from the cache rather than an Imaging object: the Micros run as batches of
// the cache's plan capacity.
struct MicroImagingResult {
  std::vector<Matrix> local_aerial_images, local_wafer_images;
  Matrix global_aerial_image, global_wafer_image;
};
class MicroImaging {
 public:
  MicroImaging(const ImagingCache& cache, const PartitionMeta& meta) : cache_(cache), meta_(meta) {
    detail::check_chip_cache(meta, cache);
  }
  MicroImagingResult compute(const Matrix& global_mask, double threshold = 0.25, double alpha = 50.0) {
    if (global_mask.rows() != meta_.height || global_mask.cols() != meta_.width)
      throw std::invalid_argument("MicroImaging: mask size must match the partition");
    const size_t g = (size_t)meta_.height * meta_.width, t = (size_t)cache_.N * cache_.N;
    const size_t T = meta_.tiles.size();
    detail::Buffer m(g), ga(g), gw(g), la(T * t), lw(T * t);
    m.upload(global_mask);
    detail::check(litho_micro_imaging(cache_.plan(), meta_.tiles.data(), (int)T, meta_.height, meta_.width, m.get(),
                                      threshold, alpha, ga.get(), gw.get(), la.get(), lw.get()));
    MicroImagingResult r;
    for (size_t i = 0; i < T; ++i) {
      r.local_aerial_images.push_back(la.download(cache_.N, cache_.N, i * t));
      r.local_wafer_images.push_back(lw.download(cache_.N, cache_.N, i * t));
    }
    r.global_aerial_image = ga.download(meta_.height, meta_.width);
    r.global_wafer_image = gw.download(meta_.height, meta_.width);
    return r;
  }

 private:
  const ImagingCache& cache_;
  const PartitionMeta& meta_;
};

struct CoreLossMetrics {             // micro_gradient_exchange.h:64-73 (EPE fields stay 0: no EPE mode here)
  double pe_loss = 0.0, bg_penalty_raw = 0.0, pe_penalty_loss = 0.0, epe_loss = 0.0, epe_abs_sum_nm = 0.0;
  int ep_count = 0;
  double epe_mean = 0.0, data_loss = 0.0;
};

struct MicroGradientData {
  double loss = 0.0, data_loss = 0.0, pe_loss = 0.0, penalty_loss = 0.0, bg_penalty_raw = 0.0, cept_loss = 0.0;
  CoreLossMetrics core;
  Matrix raw_mask_gradient;
  Matrix mask_gradient;
};

struct MicroGradientExchangeResult {
  double loss = 0.0, data_loss = 0.0, cept_loss = 0.0;
  std::vector<MicroGradientData> locals;
};

class MicroGradientExchange {
 public:
  MicroGradientExchange(const PartitionMeta& meta, const ImagingCache& cache, double threshold = 0.25,
                        double alpha = 50.0)
      : meta_(meta), cache_(cache), thr_(threshold), alpha_(alpha) {
    detail::check_chip_cache(meta, cache);
  }
  // local masks (one per Micro, full tile) against the global target
  MicroGradientExchangeResult compute(const std::vector<Matrix>& local_masks, const Matrix& global_target,
                                      const MicroGradientOptions& options = {}) {
    const int n = meta_.tile_size, T = (int)meta_.tiles.size();
    if ((int)local_masks.size() != T) throw std::invalid_argument("one local mask per Micro is required");
    const size_t plane = (size_t)n * n;
    detail::Buffer m(plane * T), t(plane * T), mg(plane * T), raw(plane * T);
    for (int i = 0; i < T; ++i) {
      if (local_masks[i].rows() != n || local_masks[i].cols() != n)
        throw std::invalid_argument("local mask size must equal meta.tile_size");
      m.upload(local_masks[i], plane * i);
      t.upload(Partition::extract_local_tile(global_target, meta_.tiles[i]), plane * i);
    }
    const litho_exchange_options o = detail::to_abi(options, thr_, alpha_);
    std::vector<litho_tile_loss> per(T);
    double tot[3];
    detail::check(litho_tile_exchange(cache_.plan(), meta_.tiles.data(), T, m.get(), t.get(), &o, mg.get(), raw.get(),
                                      nullptr, per.data(), tot));
    MicroGradientExchangeResult r;
    r.loss = tot[0];
    r.data_loss = tot[1];
    r.cept_loss = tot[2];
    for (int i = 0; i < T; ++i) {
      MicroGradientData d;
      d.loss = per[i].loss;
      d.data_loss = per[i].data_loss;
      d.pe_loss = per[i].pe_loss;
      d.penalty_loss = per[i].penalty_loss;
      d.bg_penalty_raw = per[i].bg_penalty_raw;
      d.cept_loss = per[i].cept_loss;
      d.core.pe_loss = per[i].core_pe_loss;
      d.core.bg_penalty_raw = per[i].core_bg_penalty_raw;
      d.core.pe_penalty_loss = per[i].core_pe_penalty_loss;
      d.core.data_loss = per[i].core_data_loss;
      d.raw_mask_gradient = raw.download(n, n, plane * i);
      d.mask_gradient = mg.download(n, n, plane * i);
      r.locals.push_back(std::move(d));
    }
    return r;
  }
  // a global mask: every Micro extracts its consistent local mask (CEPT is then zero)
  MicroGradientExchangeResult compute(const Matrix& global_mask, const Matrix& global_target,
                                      const MicroGradientOptions& options = {}) {
    std::vector<Matrix> locals;
    for (const litho_tile& t : meta_.tiles) locals.push_back(Partition::extract_local_tile(global_mask, t));
    return compute(locals, global_target, options);
  }

 private:
  const PartitionMeta& meta_;
  const ImagingCache& cache_;
  double thr_, alpha_;
};

struct FullChipCTMOptions {         // full_chip_ctm.h:10-17
  int iterations = 50;
  double learning_rate = 0.9;
  MicroGradientOptions gradient_options{};
};
struct FullChipCTMIteration {       // full_chip_ctm.h:19-29
  int iteration = 0;
  double loss = 0.0, data_loss = 0.0, cept_loss = 0.0;
  double pe_contribution = 0.0, penalty_contribution = 0.0, epe_contribution = 0.0, cept_contribution = 0.0;
  double pe_loss_raw = 0.0, bg_penalty_raw = 0.0, epe_loss_raw = 0.0, epe_abs_sum_nm = 0.0;
  int epe_point_count = 0;
};
struct FullChipCTMResult {
  int best_iteration = 0;
  double best_loss = 0.0;
  Matrix global_mask;
  std::vector<FullChipCTMIteration> history;
};

class FullChipCTMOptimizer {
 public:
  FullChipCTMOptimizer(const PartitionMeta& meta, const ImagingCache& cache, double threshold = 0.25,
                       double alpha = 85.0)
      : meta_(meta), cache_(cache), thr_(threshold), alpha_(alpha) {
    detail::check_chip_cache(meta, cache);
  }
  FullChipCTMResult optimize(const Matrix& global_target, const FullChipCTMOptions& options = {}) {
    return optimize(global_target, global_target, options);
  }
  FullChipCTMResult optimize(const Matrix& initial_pattern, const Matrix& global_target,
                             const FullChipCTMOptions& options = {}) {
    detail::same_shape(initial_pattern, global_target, "initial pattern and target sizes differ");
    const size_t g = (size_t)meta_.height * meta_.width;
    detail::Buffer p(g), t(g), out(g);
    p.upload(initial_pattern);
    t.upload(global_target);
    const litho_exchange_options o = detail::to_abi(options.gradient_options, thr_, alpha_);
    std::vector<litho_fullchip_row> h((size_t)(options.iterations + 1));
    FullChipCTMResult r;
    detail::check(litho_fullchip_ctm(cache_.plan(), meta_.tiles.data(), (int)meta_.tiles.size(), meta_.height,
                                     meta_.width, p.get(), t.get(), &o, options.iterations, options.learning_rate,
                                     out.get(), nullptr, h.data(), &r.best_iteration, &r.best_loss));
    r.global_mask = out.download(meta_.height, meta_.width);
    for (const litho_fullchip_row& x : h)
      r.history.push_back({x.iteration, x.loss, x.data_loss, x.cept_loss, x.pe_contribution, x.penalty_contribution,
                           x.epe_contribution, x.cept_contribution, x.pe_loss_raw, x.bg_penalty_raw, x.epe_loss_raw,
                           x.epe_abs_sum_nm, x.epe_point_count});
    return r;
  }

 private:
  const PartitionMeta& meta_;
  const ImagingCache& cache_;
  double thr_, alpha_;
};

enum class LevelSetMaskType { Binary, Gray };   // levelset_mask_renderer.h

// render_levelset_mask(phi, type, scale) (levelset_mask_renderer.cpp:7-33) on
// device 0.  Gray: bilinear upscale by `scale` (cv::resize INTER_LINEAR
// semantics), zero threshold, mean over each pixel's scale x scale block.
inline Matrix render_levelset_mask(const Matrix& phi, LevelSetMaskType type, int scale = 4) {
  // one context for these calls, created on first use and kept for the process
  static litho_ctx* ctx = [] {
    litho_ctx* c = nullptr;
    detail::check(litho_ctx_create(0, &c));
    return c;
  }();
  const size_t n = (size_t)phi.rows() * phi.cols();
  detail::Buffer p(n), m(n);
  p.upload(phi);
  detail::check(litho_render_levelset_mask(ctx, 1, phi.rows(), phi.cols(), p.get(),
                                           type == LevelSetMaskType::Gray ? 1 : 0, scale, m.get()));
  return m.download(phi.rows(), phi.cols());
}

struct FullChipLSMOptions {         // full_chip_lsm.h:10-25
  int iterations = 200;
  double dx = 1.0, dy = 1.0;
  double cfl = 0.7;
  double curvature_weight = 0.01;
  int reinit_interval = 20;
  LevelSetMaskType mask_type = LevelSetMaskType::Binary;
  int render_scale = 4;             // Gray: upscale per direction
  MicroGradientOptions gradient_options{};
};
struct FullChipLSMIteration {       // full_chip_lsm.h: the columns of demo_full_chip_lsm.cpp:170-184
  int iteration = 0;
  double loss = 0.0, data_loss = 0.0, cept_loss = 0.0, min_dt = 0.0, max_dt = 0.0;
  double pe_contribution = 0.0, penalty_contribution = 0.0, epe_contribution = 0.0, cept_contribution = 0.0;
  CoreLossMetrics core;
  double total_pe_loss = 0.0, total_bg_penalty_raw = 0.0, total_epe_loss = 0.0, total_epe_abs_sum_nm = 0.0;
  int total_ep_count = 0;
  double epe_mean = 0.0;
};
struct FullChipLSMResult {
  int best_iteration = 0;
  double best_loss = 0.0;
  Matrix global_mask, global_phi;
  std::vector<FullChipLSMIteration> history;
};

class FullChipLSMOptimizer {
 public:
  FullChipLSMOptimizer(const PartitionMeta& meta, const ImagingCache& cache, double threshold = 0.25,
                       double alpha = 85.0)
      : meta_(meta), cache_(cache), thr_(threshold), alpha_(alpha) {
    detail::check_chip_cache(meta, cache);
  }
  FullChipLSMResult optimize(const Matrix& initial_binary, const Matrix& global_target,
                             const FullChipLSMOptions& options = {}) {
    detail::same_shape(initial_binary, global_target, "initial binary and target sizes differ");
    const size_t g = (size_t)meta_.height * meta_.width;
    detail::Buffer b(g), t(g), mask(g), phi(g);
    b.upload(initial_binary);
    t.upload(global_target);
    const litho_exchange_options o = detail::to_abi(options.gradient_options, thr_, alpha_);
    const litho_lsm_options lo{options.dx, options.dy, options.cfl, options.curvature_weight,
                               options.mask_type == LevelSetMaskType::Gray ? 1 : 0, options.render_scale};
    std::vector<litho_fullchip_row> h((size_t)(options.iterations + 1));
    FullChipLSMResult r;
    detail::check(litho_fullchip_lsm(cache_.plan(), meta_.tiles.data(), (int)meta_.tiles.size(), meta_.height,
                                     meta_.width, b.get(), t.get(), &o, &lo, options.iterations,
                                     options.reinit_interval, mask.get(), phi.get(), h.data(), &r.best_iteration,
                                     &r.best_loss));
    r.global_mask = mask.download(meta_.height, meta_.width);
    r.global_phi = phi.download(meta_.height, meta_.width);
    for (const litho_fullchip_row& x : h) {
      FullChipLSMIteration it;
      it.iteration = x.iteration;
      it.loss = x.loss;
      it.data_loss = x.data_loss;
      it.cept_loss = x.cept_loss;
      it.min_dt = x.min_dt;
      it.max_dt = x.max_dt;
      it.pe_contribution = x.pe_contribution;
      it.penalty_contribution = x.penalty_contribution;
      it.epe_contribution = x.epe_contribution;
      it.cept_contribution = x.cept_contribution;
      it.core.pe_loss = x.core_pe;
      it.core.bg_penalty_raw = x.core_bg_penalty_raw;
      it.core.pe_penalty_loss = x.core_pe_penalty;
      it.core.data_loss = x.core_data_loss;
      it.total_pe_loss = x.pe_loss_raw;
      it.total_bg_penalty_raw = x.bg_penalty_raw;
      it.epe_mean = x.epe_mean;
      r.history.push_back(it);
    }
    return r;
  }

 private:
  const PartitionMeta& meta_;
  const ImagingCache& cache_;
  double thr_, alpha_;
};

}  // namespace litho
