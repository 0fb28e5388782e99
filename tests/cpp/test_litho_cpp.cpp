// C++ tests of the reference call surface (include/litho_cpp.hpp) on the GPU,
// following the reference's tests/test_ctm_optimizer.cpp (test_synthetic:
// 64 x 64 and 65 x 65 grids, a 19 x 19 square at (23, 23), annular 0.6 / 0.9 source, NA
// 1.35, 4 nm pixels, threshold 0.25, alpha 85) and the consistency checks of
// tests/test_micro_gradient_exchange.cpp / test_full_chip_ctm.cpp.
// Exit status 0 when every check holds; each failure prints its message.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/litho_cpp.hpp"

namespace {

int failures = 0;

void require(bool condition, const std::string& message) {
  if (!condition) {
    std::printf("FAIL: %s\n", message.c_str());
    ++failures;
  }
}

litho::Matrix square_target(int n) {
  litho::Matrix t(n, n);
  for (int y = 23; y < 42; ++y)
    for (int x = 23; x < 42; ++x) t(y, x) = 1.0;
  return t;
}

litho::Matrix mask_of(const litho::Matrix& theta) {
  litho::Matrix m(theta.rows(), theta.cols());
  for (size_t i = 0; i < m.size(); ++i) m.data()[i] = 0.5 * (1.0 + std::cos(theta.data()[i]));
  return m;
}

// check_pe_gradient of the reference test: central finite difference of the PE
// loss along the normalised descent direction in theta.
void check_pe_gradient(const litho::ImagingCache& cache, const litho::Matrix& target, bool with_penalty) {
  const double threshold = 0.25, alpha = with_penalty ? 10.0 : 85.0;
  const double penalty = 2.0, penalty_threshold = 0.01;
  const litho::Matrix theta = litho::CTM_Optimizer::initialize_theta(target);
  litho::Imaging imaging(cache);
  litho::Gradient gradient(cache, threshold, alpha);
  const litho::Matrix mask = mask_of(theta);
  const auto baseline = imaging.compute(mask, threshold, alpha);
  if (with_penalty)
    require(litho::Loss::pe_loss_with_penalty(baseline.wafer_image, target, penalty, penalty_threshold) >
                litho::Loss::pe_loss(baseline.wafer_image, target),
            "penalty test has no active penalty pixels");
  const litho::Matrix gm = with_penalty ? gradient.pe_gradient_with_penalty(mask, target, penalty, penalty_threshold)
                                        : gradient.pe_gradient(mask, target);
  litho::Matrix gt(theta.rows(), theta.cols());
  for (size_t i = 0; i < gt.size(); ++i) gt.data()[i] = -0.5 * gm.data()[i] * std::sin(theta.data()[i]);
  require(gt.norm() > 1e-8, "PE gradient is unexpectedly zero");
  const double nrm = gt.norm();
  auto evaluate = [&](double h) {
    litho::Matrix th = theta;
    for (size_t i = 0; i < th.size(); ++i) th.data()[i] -= h * gt.data()[i] / nrm;
    const auto r = imaging.compute(mask_of(th), threshold, alpha);
    return with_penalty ? litho::Loss::pe_loss_with_penalty(r.wafer_image, target, penalty, penalty_threshold)
                        : litho::Loss::pe_loss(r.wafer_image, target);
  };
  // The device images are FP32: their rounding (~1e-7 relative in PE) divided
  // by a 1e-3 step reaches the 1e-3 bound, so the step is 1e-2 here.  The
  // central difference's truncation at 1e-2 is below 1e-5 relative for both
  // cases (FP64 oracle: 9.1e-6 without, 6.8e-8 with the penalty).
  const double step = 1e-2;
  const double numeric = (evaluate(step) - evaluate(-step)) / (2.0 * step);
  const double analytic = -nrm;
  std::printf("%s gradient: finite difference %.6f, analytic %.6f\n", with_penalty ? "penalty PE" : "PE", numeric,
              analytic);
  // the reference's bound (test_ctm_optimizer.cpp check_pe_gradient)
  require(std::abs(numeric - analytic) <= 1e-3 * std::max(1.0, std::abs(numeric)),
          with_penalty ? "penalty PE gradient disagrees with finite differences"
                       : "PE gradient disagrees with finite differences");
}

void test_synthetic_grid(int n);

void test_synthetic() {
  // even and odd optical grids, as the reference test (65 has no compiled
  // FFT and runs the direct-DFT pipeline)
  for (int n : {64, 65}) test_synthetic_grid(n);
}

void test_synthetic_grid(int n) {
  std::printf("grid %d:\n", n);
  litho::OpticsSpec optics;   // 4 nm, 193 nm, NA 1.35, annular 0.6 / 0.9
  litho::LithoPrepare prepare(n, optics, 24);
  const litho::ImagingCache& cache = prepare.cache();
  require(cache.N == n, "imaging cache and grid dimensions differ");
  const litho::Matrix target = square_target(n);
  check_pe_gradient(cache, target, false);
  check_pe_gradient(cache, target, true);

  litho::CTM_Optimizer gradient_descent(cache, target, 3, 0.9, "gradient_descent", 10, 1e100, 0.25, 85.0);
  const litho::Matrix gd_mask = gradient_descent.optimize(false, false);
  require(gradient_descent.error_history().size() == 1, "gradient descent did not stop on gradient norm");
  require(gd_mask.allFinite(), "gradient descent returned a non-finite mask");

  litho::CTM_Optimizer lbfgs_early(cache, target, 3, 0.9, "lbfgs", 10, 1e100, 0.25, 85.0);
  const litho::Matrix lbfgs_early_mask = lbfgs_early.optimize(false, false);
  require(lbfgs_early.error_history().size() == 1, "L-BFGS did not stop on gradient norm");
  require(lbfgs_early_mask.allFinite(), "L-BFGS returned a non-finite mask");
  double diff = 0.0;
  for (size_t i = 0; i < gd_mask.size(); ++i) diff = std::max(diff, std::abs(gd_mask.data()[i] - lbfgs_early_mask.data()[i]));
  require(diff <= 1e-12, "both methods should return the same initial mask when stopped immediately");

  litho::CTM_Optimizer lbfgs_step(cache, target, 1, 0.9, "lbfgs", 10, 0.0, 0.25, 85.0);
  const litho::Matrix lbfgs_step_mask = lbfgs_step.optimize(false, false);
  const auto& h = lbfgs_step.error_history();
  require(h.size() >= 2, "L-BFGS did not evaluate a line-search step");
  require(!h.empty() && h.back() <= h.front(), "L-BFGS line search did not reduce PE");
  require(lbfgs_step_mask.allFinite(), "L-BFGS step returned a non-finite mask");
  std::printf("CTM: GD early stop %zu evaluation, L-BFGS one step %zu evaluations, PE %.4f -> %.4f\n",
              gradient_descent.error_history().size(), h.size(), h.front(), h.back());

  bool threw = false;
  try {
    litho::CTM_Optimizer bad(cache, target, 3, 0.9, "newton");
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  require(threw, "an unknown optimizer method must raise std::invalid_argument");
}

void test_full_chip() {
  // 96 x 96 chip, 2 x 2 Micros, halo 16, guard 8 -> tile 80
  const litho::PartitionMeta meta = litho::Partition::build(96, 96, 2, 2, 16, 8);
  require(meta.tile_size == 80 && meta.tiles.size() == 4, "Partition::build geometry");
  require(meta.tiles[0].num_neighbors == 2 && meta.tiles[0].neighbors[0].direction == 1,
          "Partition::build neighbour order (right, then bottom, for the top-left Micro)");
  litho::LithoPrepare prepare(meta.tile_size, litho::OpticsSpec{}, 8, (int)meta.tiles.size());
  const litho::ImagingCache& cache = prepare.cache();
  litho::Matrix target(96, 96);
  for (int y = 20; y < 44; ++y)
    for (int x = 30; x < 70; ++x) target(y, x) = 1.0;
  for (int y = 60; y < 84; ++y)
    for (int x = 10; x < 50; ++x) target(y, x) = 1.0;

  // a global mask gives consistent local masks: CEPT loss and its gradient vanish
  litho::MicroGradientExchange exchange(meta, cache, 0.225, 50.0);
  litho::MicroGradientOptions o;
  o.enable_cept = true;
  const auto ex = exchange.compute(target, target, o);
  require(ex.cept_loss == 0.0, "CEPT loss of consistent local masks must be zero");
  double sum = 0.0;
  for (const auto& l : ex.locals) sum += l.loss;
  require(std::abs(sum - ex.loss) <= 1e-9 * std::max(1.0, ex.loss), "total loss is the sum of the Micros' losses");
  for (const auto& l : ex.locals) {
    double d = 0.0;
    for (size_t i = 0; i < l.mask_gradient.size(); ++i)
      d = std::max(d, std::abs(l.mask_gradient.data()[i] - l.raw_mask_gradient.data()[i]));
    require(d == 0.0, "without fusion and with zero CEPT the mask gradient equals the raw gradient");
  }

  litho::FullChipCTMOptions co;
  co.iterations = 3;
  co.gradient_options.enable_cept = true;
  co.gradient_options.enable_grad_fusion = true;
  litho::FullChipCTMOptimizer ctm(meta, cache, 0.225, 50.0);
  const auto rc = ctm.optimize(target, co);
  require(rc.history.size() == 4, "FullChipCTM history has iterations + 1 entries");
  require(rc.global_mask.allFinite() && rc.best_loss <= rc.history[0].loss, "FullChipCTM best state");
  std::printf("FullChipCTM: loss %.4f -> best %.4f (iteration %d)\n", rc.history[0].loss, rc.best_loss,
              rc.best_iteration);

  // MicroImaging: each local image equals Imaging::compute of the Micro's
  // extract; the global images hold the Cores
  litho::MicroImaging micro(cache, meta);
  const auto mi = micro.compute(target, 0.225, 50.0);
  litho::Imaging single(cache);
  const litho::Matrix first = litho::Partition::extract_local_tile(target, meta.tiles[0]);
  const auto ref0 = single.compute(first, 0.225, 50.0);
  double dmax = 0.0, amax = 0.0;
  for (size_t i = 0; i < ref0.aerial_image.size(); ++i) {
    dmax = std::max(dmax, std::abs(mi.local_aerial_images[0].data()[i] - ref0.aerial_image.data()[i]));
    amax = std::max(amax, std::abs(ref0.aerial_image.data()[i]));
  }
  const litho_rect& core = meta.tiles[0].core_local;
  const litho_rect& cg = meta.tiles[0].core_global;
  require(mi.local_aerial_images.size() == meta.tiles.size() && dmax <= 1e-6 * amax &&
              mi.global_aerial_image(cg.y0, cg.x0) == mi.local_aerial_images[0](core.y0, core.x0),
          "MicroImaging local images and Core assembly");

  litho::FullChipLSMOptions lo;
  lo.iterations = 3;
  lo.reinit_interval = 2;
  lo.gradient_options.enable_cept = true;
  litho::FullChipLSMOptimizer lsm(meta, cache, 0.225, 50.0);
  const auto rl = lsm.optimize(target, target, lo);
  require(rl.history.size() == 4 && rl.history[0].min_dt == 0.0 && rl.history[1].min_dt > 0.0,
          "FullChipLSM history and CFL steps");
  bool binary = true;
  for (size_t i = 0; i < rl.global_mask.size(); ++i)
    binary = binary && (rl.global_mask.data()[i] == 0.0 || rl.global_mask.data()[i] == 1.0);
  require(binary, "FullChipLSM Binary render returns a 0/1 mask");
  std::printf("FullChipLSM: loss %.4f -> best %.4f (update %d), dt %.4f..%.4f\n", rl.history[0].loss, rl.best_loss,
              rl.best_iteration, rl.history[1].min_dt, rl.history[1].max_dt);

  lo.mask_type = litho::LevelSetMaskType::Gray;
  const auto rg = lsm.optimize(target, target, lo);
  bool quarters = true, fractional = false;
  for (size_t i = 0; i < rg.global_mask.size(); ++i) {
    const double v = rg.global_mask.data()[i] * 16.0;
    quarters = quarters && v == (double)(int)v && v >= 0.0 && v <= 16.0;
    fractional = fractional || (v > 0.0 && v < 16.0);
  }
  require(quarters && fractional, "FullChipLSM Gray render returns coverages in sixteenths, some partial");
  std::printf("FullChipLSM Gray: loss %.4f -> best %.4f\n", rg.history[0].loss, rg.best_loss);
}

// render_levelset_mask closed forms: scale 1 is the Binary render; a field
// linear in x with its zero at x = c covers, per pixel x, the fraction of its
// subpixel centres (x * s + k + 0.5) / s - 0.5 >= c (clamped to the tile).
void test_render() {
  litho::Matrix phi(3, 16);
  const double c = 7.0;
  for (int y = 0; y < 3; ++y)
    for (int x = 0; x < 16; ++x) phi(y, x) = x - c;
  const auto b = litho::render_levelset_mask(phi, litho::LevelSetMaskType::Binary);
  const auto g1 = litho::render_levelset_mask(phi, litho::LevelSetMaskType::Gray, 1);
  const auto g4 = litho::render_levelset_mask(phi, litho::LevelSetMaskType::Gray, 4);
  bool same = true, ramp = true;
  for (int x = 0; x < 16; ++x) {
    same = same && b(1, x) == g1(1, x) && b(1, x) == (x >= c ? 1.0 : 0.0);
    int inside = 0;
    for (int k = 0; k < 4; ++k) {
      double u = (x * 4 + k + 0.5) / 4.0 - 0.5;
      u = u < 0.0 ? 0.0 : (u > 15.0 ? 15.0 : u);
      inside += u - c >= 0.0 ? 1 : 0;
    }
    ramp = ramp && g4(1, x) == inside / 4.0;
  }
  require(same, "render_levelset_mask: Gray at scale 1 equals Binary");
  require(ramp, "render_levelset_mask: Gray coverage of a linear field");
  bool threw = false;
  try {
    (void)litho::render_levelset_mask(phi, litho::LevelSetMaskType::Gray, 0);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  require(threw, "render_levelset_mask: scale 0 is rejected");
}

}  // namespace

int main() {
  try {
    test_synthetic();
    test_full_chip();
    test_render();
  } catch (const std::exception& e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 1;
  }
  std::printf(failures == 0 ? "all checks passed\n" : "%d checks failed\n", failures);
  return failures == 0 ? 0 : 1;
}
