// EPE loss and gradient on the device (epe.hpp).
//
// One evaluation: the forward image (aerial kept), one thread per edge point
// for the bilinear samples of the aerial image and of its gradient magnitude
// (central differences, one-sided second-order at the tile edges, as
// LevelSetUtils::_gradient_1d with d = 1), the point's loss and its dL/dI
// weight 2 w (I - tau) / G^2; the pixel -> points table sums each stencil
// pixel's contributions in point order (the reference's bilinear_scatter
// order) into a dL/dI image that the adjoint adds to the PE term.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "epe.hpp"
#include "opt_kernels.hpp"

namespace litho {

namespace {

__device__ __forceinline__ double d1(const float* __restrict__ m, int n, int y, int x, bool along_x) {
  // LevelSetUtils::_gradient_1d (level_set_utils.cpp:228-255) at one pixel, d = 1
  auto at = [&](int yy, int xx) { return (double)m[(size_t)yy * n + xx]; };
  const int i = along_x ? x : y;
  auto v = [&](int k) { return along_x ? at(y, k) : at(k, x); };
  if (i == 0) return (-3.0 * v(0) + 4.0 * v(1) - v(2)) / 2.0;
  if (i == n - 1) return (3.0 * v(n - 1) - 4.0 * v(n - 2) + v(n - 3)) / 2.0;
  return (v(i + 1) - v(i - 1)) / 2.0;
}

__global__ void k_epe_sample(const float* __restrict__ aerial, int n, int count, const int* __restrict__ pix,
                             const double* __restrict__ wts, const double* __restrict__ wep, double thr,
                             double pixel_nm, double* __restrict__ sg, double* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double w = wep[i];
  double g = 0.0, loss = 0.0, absnm = 0.0, act = 0.0;
  if (w > 0.0) {
    double iv = 0.0, gm = 0.0;
    for (int j = 0; j < 4; ++j) {
      const int p = pix[4 * i + j];
      if (p < 0) continue;
      const double sw = wts[4 * i + j];
      const int y = p / n, x = p - y * n;
      iv += sw * (double)aerial[p];
      const double gx = d1(aerial, n, y, x, true), gy = d1(aerial, n, y, x, false);
      gm += sw * fmax(sqrt(gx * gx + gy * gy), 1e-6);
    }
    const double mag = fmax(gm, 1e-6);
    const double signed_epe = (iv - thr) / mag;
    loss = w * signed_epe * signed_epe;
    absnm = fabs(signed_epe) * pixel_nm;
    act = 1.0;
    g = 2.0 * w * (iv - thr) / (mag * mag);
  }
  sg[i] = g;
  part[3 * (size_t)i + 0] = loss;
  part[3 * (size_t)i + 1] = absnm;
  part[3 * (size_t)i + 2] = act;
}

__global__ void k_epe_scatter(int touched, const int* __restrict__ csr_pix, const int* __restrict__ csr_off,
                              const int* __restrict__ csr_ep, const double* __restrict__ csr_w,
                              const double* __restrict__ sg, double scale, float* __restrict__ extra) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= touched) return;
  double v = 0.0;
  for (int e = csr_off[q]; e < csr_off[q + 1]; ++e) v += csr_w[e] * sg[csr_ep[e]];
  extra[csr_pix[q]] = (float)(scale * v);
}

struct Entry {
  int y, x;
  double w;
};

// make_bilinear_stencil (epe_utils.cpp:38-64): clamp into the pixel centres,
// the (up to) four neighbouring pixels with separable weights, duplicates at
// the edges merged.
int bilinear_stencil(int h, int w, double y, double x, Entry (&out)[4]) {
  const double yy = std::clamp(y, 0.0, (double)(h - 1)), xx = std::clamp(x, 0.0, (double)(w - 1));
  const int y0 = (int)std::floor(yy), x0 = (int)std::floor(xx);
  const int y1 = std::min(y0 + 1, h - 1), x1 = std::min(x0 + 1, w - 1);
  const double wy1 = y1 == y0 ? 0.0 : yy - y0, wx1 = x1 == x0 ? 0.0 : xx - x0;
  const double wy0 = 1.0 - wy1, wx0 = 1.0 - wx1;
  int c = 0;
  const int py[2] = {y0, y1}, px[2] = {x0, x1};
  const double wy[2] = {wy0, wy1}, wx[2] = {wx0, wx1};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const double wt = wy[a] * wx[b];
      if (wt <= 0.0) continue;
      int ex = -1;
      for (int i = 0; i < c; ++i)
        if (out[i].y == py[a] && out[i].x == px[b]) ex = i;
      if (ex >= 0)
        out[ex].w += wt;
      else
        out[c++] = Entry{py[a], px[b], wt};
    }
  return c;
}

}  // namespace

EpSet* make_ep_set(Plan& plan, int count, const double* yx, const double* weights) {
  if (count < 0) throw Error(1, "EPE: negative point count");
  if (count > 0 && (yx == nullptr || weights == nullptr)) throw Error(1, "EPE: null points or weights");
  const int n = plan.n;
  for (int i = 0; i < count; ++i)
    if (!std::isfinite(yx[2 * i]) || !std::isfinite(yx[2 * i + 1]) || !std::isfinite(weights[i]))
      throw Error(1, "EPE: non-finite edge point or weight");
  auto* set = new EpSet();
  try {
    set->plan = &plan;
    set->count = count;
    std::vector<int> pix((size_t)4 * count, -1);
    std::vector<double> wts((size_t)4 * count, 0.0);
    struct Trip {
      int p, ep, k;
      double w;
    };
    std::vector<Trip> trips;
    for (int i = 0; i < count; ++i) {
      Entry e[4];
      const int c = bilinear_stencil(n, n, yx[2 * i], yx[2 * i + 1], e);
      for (int j = 0; j < c; ++j) {
        pix[4 * i + j] = e[j].y * n + e[j].x;
        wts[4 * i + j] = e[j].w;
        if (weights[i] > 0.0) trips.push_back(Trip{e[j].y * n + e[j].x, i, j, e[j].w});
      }
    }
    std::stable_sort(trips.begin(), trips.end(), [](const Trip& a, const Trip& b) { return a.p < b.p; });
    std::vector<int> cpix, coff(1, 0), cep;
    std::vector<double> cw;
    for (size_t t = 0; t < trips.size(); ++t) {
      if (t == 0 || trips[t].p != trips[t - 1].p) {
        if (t > 0) coff.push_back((int)t);
        cpix.push_back(trips[t].p);
      }
      cep.push_back(trips[t].ep);
      cw.push_back(trips[t].w);
    }
    coff.push_back((int)trips.size());
    if (cpix.empty()) coff.assign(1, 0);
    set->touched = (int)cpix.size();
    hipStream_t s = plan.ctx->stream;
    auto up = [&](DevMem& m, const void* src, size_t bytes) {
      m.alloc(std::max<size_t>(bytes, 8));
      if (bytes) LITHO_HIP(hipMemcpyAsync(m.ptr, src, bytes, hipMemcpyHostToDevice, s));
    };
    up(set->pix, pix.data(), pix.size() * sizeof(int));
    up(set->wts, wts.data(), wts.size() * sizeof(double));
    up(set->wep, weights, (size_t)count * sizeof(double));
    up(set->csr_pix, cpix.data(), cpix.size() * sizeof(int));
    up(set->csr_off, coff.data(), coff.size() * sizeof(int));
    up(set->csr_ep, cep.data(), cep.size() * sizeof(int));
    up(set->csr_w, cw.data(), cw.size() * sizeof(double));
    set->sg.alloc(std::max<size_t>((size_t)count * sizeof(double), 8));
    set->part.alloc(std::max<size_t>((size_t)count * 3 * sizeof(double), 24));
    set->red.alloc(3 * sizeof(double));
    const size_t plane = plan.tile_elems() * sizeof(float);
    set->extra.alloc(plane);
    set->aerial.alloc(plane);
    // pixels outside every stencil stay 0 in the dL/dI image for good
    LITHO_HIP(hipMemsetAsync(set->extra.ptr, 0, plane, s));
    LITHO_HIP(hipStreamSynchronize(s));
  } catch (...) {
    delete set;
    throw;
  }
  return set;
}

EpeResult epe_gradient(EpSet& set, const float* mask, const float* target, const float* pe_window, double pe_weight,
                       double epe_weight, double threshold, double alpha, double pixel_nm, float* grad_mask,
                       float* aerial, float* wafer) {
  Plan& plan = *set.plan;
  if (mask == nullptr) throw Error(1, "EPE: null mask");
  if (!std::isfinite(threshold) || !std::isfinite(alpha) || !(alpha > 0.0))
    throw Error(1, "EPE: threshold must be finite and alpha positive");
  if (!std::isfinite(pe_weight) || !std::isfinite(epe_weight) || !std::isfinite(pixel_nm))
    throw Error(1, "EPE: non-finite weight or pixel size");
  hipStream_t s = plan.ctx->stream;
  const size_t plane = plan.tile_elems();
  // 1. forward image
  EvalRequest f;
  f.mode = 0;
  f.base = mask;
  f.adjoint = false;
  f.im.thr = (float)threshold;
  f.im.alpha = (float)alpha;
  f.im.aerial = set.aerial.as<float>();
  run_eval(plan, 1, f);
  // 2. point samples, losses and dL/dI weights
  EpeResult r;
  std::vector<double> tot(3, 0.0);
  if (set.count > 0) {
    hipLaunchKernelGGL(k_epe_sample, dim3((set.count + 255) / 256), dim3(256), 0, s, set.aerial.as<float>(), plan.n,
                       set.count, set.pix.as<int>(), set.wts.as<double>(), set.wep.as<double>(), threshold, pixel_nm,
                       set.sg.as<double>(), set.part.as<double>());
    launch_reduce_fixed(set.part.as<double>(), set.count, 3, set.red.as<double>(), 3, 1, s);
    // 3. the dL/dI image, pre-divided by the adjoint's 4 alpha: the EPE term's
    // own factor is 2 (gradient.cpp:225)
    if (set.touched > 0)
      hipLaunchKernelGGL(k_epe_scatter, dim3((set.touched + 255) / 256), dim3(256), 0, s, set.touched,
                         set.csr_pix.as<int>(), set.csr_off.as<int>(), set.csr_ep.as<int>(), set.csr_w.as<double>(),
                         set.sg.as<double>(), epe_weight * 2.0 / (4.0 * alpha), set.extra.as<float>());
    LITHO_HIP(hipMemcpyAsync(tot.data(), set.red.ptr, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  // 4. forward + adjoint with the PE term (weight 0 against a zero target when none is given)
  const float* tgt = target;
  if (tgt == nullptr) {
    if (set.zero_target.ptr == nullptr) {
      set.zero_target.alloc(plane * sizeof(float));
      LITHO_HIP(hipMemsetAsync(set.zero_target.ptr, 0, plane * sizeof(float), s));
    }
    tgt = set.zero_target.as<float>();
  }
  EvalRequest g;
  g.mode = 0;
  g.base = mask;
  g.adjoint = true;
  g.im.thr = (float)threshold;
  g.im.alpha = (float)alpha;
  g.im.aerial = aerial;
  g.im.wafer = wafer;
  g.im.target = tgt;
  g.im.pe_win = target != nullptr ? pe_window : nullptr;
  g.im.pe_weight = target != nullptr ? (float)pe_weight : 0.f;
  g.im.extra_dldi = set.count > 0 ? set.extra.as<float>() : nullptr;
  g.ga.scale = (float)(4.0 * alpha);
  g.ga.grad_mask = grad_mask;
  run_eval(plan, 1, g);
  const double* red = fetch_reductions(plan, 1);
  r.pe_loss = target != nullptr ? red[0] : 0.0;
  r.epe_loss = tot[0];
  r.epe_abs_sum_nm = tot[1];
  r.active = (int)tot[2];
  return r;
}

}  // namespace litho
