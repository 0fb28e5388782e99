// Host side of the level-set path (lsm.hpp).
#include "lsm.hpp"

#include <cmath>
#include <cstring>

#include "levelset.hpp"

namespace litho {

namespace {

void check_tile(int batch, int rows, int cols) {
  if (batch <= 0 || batch > 65535) throw Error(1, "level set: batch must be in [1, 65535]");
  if (rows < 3 || cols < 3) throw Error(1, "level set: tiles need at least 3 rows and 3 columns");
  if (cols > 4096) throw Error(1, "level set: at most 4096 columns per tile");
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

void check_render(int type, int scale) {
  if (type != 0 && type != 1) throw Error(1, "render_levelset_mask: type must be Binary (0) or Gray (1)");
  if (type == 1 && scale <= 0) throw Error(1, "render_levelset_mask: render_scale must be positive");
}

}  // namespace

void render_levelset_mask(Context& ctx, int batch, int rows, int cols, const float* phi, int type, int scale,
                          float* mask) {
  if (batch <= 0 || batch > 65535 || rows <= 0 || cols <= 0)
    throw Error(1, "render_levelset_mask: empty tile or batch out of [1, 65535]");
  if (phi == nullptr || mask == nullptr) throw Error(1, "render_levelset_mask: null buffer");
  check_render(type, scale);
  launch_render_mask(phi, batch, rows, cols, type, scale, mask, ctx.stream);
  LITHO_HIP(hipGetLastError());
}

void ls_terms(Context& ctx, int batch, int rows, int cols, const float* phi, const float* vn, double dx, double dy,
              double b, float* delta, float* kappa) {
  check_tile(batch, rows, cols);
  if (phi == nullptr || (delta != nullptr && vn == nullptr)) throw Error(1, "level set: null input");
  if (!(dx > 0.0) || !(dy > 0.0)) throw Error(1, "evolve_normal_WENO_godunov: dx/dy must be positive");
  launch_ls_terms(phi, vn, batch, rows, cols, dx, dy, b, delta, kappa, ctx.stream);
  LITHO_HIP(hipGetLastError());
}

double ls_dt(double cfl, double dx, double dy, double max_h1, double max_h2, double b) {
  const double normal = max_h1 / dx + max_h2 / dy;
  const double kap = b * (1.0 / (dx * dx) + 1.0 / (dy * dy)) * 2.0;
  const double den = normal + kap;
  return den < 1e-9 ? 1e-3 : cfl / den;
}

void ls_reinit_sdf(Context& ctx, int batch, int rows, int cols, const float* phi, double dx, double dy, float* sdf) {
  check_tile(batch, rows, cols);
  if (phi == nullptr || sdf == nullptr) throw Error(1, "reinit_sdf: null buffer");
  const size_t plane = (size_t)rows * cols;
  char* ws = static_cast<char*>(ctx.workspace(chamfer_scratch_bytes(batch, rows, cols)));
  uint32_t* scratch = reinterpret_cast<uint32_t*>(ws);
  float* dist = reinterpret_cast<float*>(ws + (size_t)batch * 2 * plane * sizeof(uint32_t));
  launch_reinit_sdf(phi, batch, rows, cols, dx, dy, 0.f, false, scratch, dist, nullptr, sdf, nullptr, ctx.stream);
  LITHO_HIP(hipGetLastError());
}

void lsm_step(Context& ctx, int batch, int n, const float* phi, const float* grad, const int* rects,
              const LsmOptions& o, bool reinit, float* phi_out, float* mask_out, double* dt_out) {
  check_tile(batch, n, n);
  if (phi == nullptr || grad == nullptr || rects == nullptr || phi_out == nullptr)
    throw Error(1, "lsm_step: null buffer");
  if (!(o.dx > 0.0) || !(o.dy > 0.0)) throw Error(1, "lsm_step: dx/dy must be positive");
  check_render(o.mask_type, o.render_scale);
  for (int b = 0; b < batch; ++b) {
    const int* r = rects + 4 * b;
    if (r[0] < 0 || r[1] > n || r[2] < 0 || r[3] > n || r[0] > r[1] || r[2] > r[3])
      throw Error(1, "lsm_step: update window outside the tile");
  }
  const size_t plane = (size_t)n * n;
  // workspace: rects | dt | partial maxima | phi1 | chamfer scratch + distances
  const size_t off_dt = align256((size_t)batch * 4 * sizeof(int));
  const size_t off_part = off_dt + align256((size_t)batch * sizeof(double));
  const size_t off_phi1 = off_part + align256((size_t)batch * kLsmDtBlocks * sizeof(float));
  const size_t off_ch = off_phi1 + align256((size_t)batch * plane * sizeof(float));
  const size_t total = off_ch + (reinit ? chamfer_scratch_bytes(batch, n, n) : 0);
  char* ws = static_cast<char*>(ctx.workspace(total));
  hipStream_t s = ctx.stream;
  int* d_rects = reinterpret_cast<int*>(ws);
  double* d_dt = reinterpret_cast<double*>(ws + off_dt);
  float* part = reinterpret_cast<float*>(ws + off_part);
  float* phi1 = reinterpret_cast<float*>(ws + off_phi1);
  LITHO_HIP(hipMemcpyAsync(d_rects, rects, (size_t)batch * 4 * sizeof(int), hipMemcpyHostToDevice, s));
  launch_lsm_dt(grad, d_rects, batch, n, o.dx, o.dy, o.cfl, o.curvature_weight, part, d_dt, s);
  launch_lsm_stage(1, phi, phi, grad, d_rects, d_dt, batch, n, o.dx, o.dy, o.curvature_weight, phi1, nullptr, s);
  // the Binary render comes with the last kernel that writes phi_out; the
  // Gray render reads the finished phi_out
  float* binary_mask = o.mask_type == 0 ? mask_out : nullptr;
  launch_lsm_stage(2, phi1, phi, grad, d_rects, d_dt, batch, n, o.dx, o.dy, o.curvature_weight, phi_out,
                   reinit ? nullptr : binary_mask, s);
  if (reinit) {
    uint32_t* scratch = reinterpret_cast<uint32_t*>(ws + off_ch);
    float* dist = reinterpret_cast<float*>(ws + off_ch + (size_t)batch * 2 * plane * sizeof(uint32_t));
    // signs = (phi >= 0) - 0.5: inside means phi_out >= 0 (full_chip_lsm.cpp:236-239)
    launch_reinit_sdf(phi_out, batch, n, n, o.dx, o.dy, 0.f, true, scratch, dist, d_rects, phi_out, binary_mask, s);
  }
  if (o.mask_type == 1 && mask_out != nullptr)
    launch_render_mask(phi_out, batch, n, n, 1, o.render_scale, mask_out, s);
  LITHO_HIP(hipGetLastError());
  if (dt_out != nullptr) {
    LITHO_HIP(hipMemcpyAsync(dt_out, d_dt, (size_t)batch * sizeof(double), hipMemcpyDeviceToHost, s));
    LITHO_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace litho
