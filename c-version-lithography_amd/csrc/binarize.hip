// Imbinarize::binarize_advanced (imbinarize.cpp:84-190) on the device.
//
// * Peaks (_peak_values, :63-80): pixels >= the maximum of their in-image
//   3 x 3 neighbourhood and > peak_min_intensity.
// * Threshold (_quantile, :44-58): the floor / ceil order statistics of the
//   peak values at q (n - 1), found exactly by a two-level 16-bit radix
//   select over order-preserving keys of the float values (two histogram
//   passes, no sort), then interpolated in FP64 on the host.
// * Connected components (cv::connectedComponentsWithStats, 8-connectivity):
//   lock-free union-find over pixel indices (each pixel unions with its
//   already-scanned neighbours, links always point to a smaller index, so
//   every component's root is its smallest pixel index), then path
//   compression, an area count per root and the area filter.
// * Keep-out zone (cv::dilate with the MORPH_ELLIPSE element of radius r):
//   per pixel the distance to the nearest target pixel of its row within r,
//   then a vertical pass over the element's 2r + 1 rows.
// All decisions are integer / comparison work: the result does not depend on
// thread scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "binarize.hpp"

namespace litho {

namespace {

constexpr int kBins = 65536;

__device__ __forceinline__ uint32_t float_key(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

float key_float(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float f;
  std::memcpy(&f, &b, sizeof(f));
  return f;
}

__device__ __forceinline__ bool is_peak(const float* __restrict__ g, int h, int w, int y, int x, double floor_v,
                                        float& v) {
  v = g[(size_t)y * w + x];
  if (!((double)v > floor_v)) return false;
  float m = v;
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
      const int yy = y + dy, xx = x + dx;
      if (yy >= 0 && yy < h && xx >= 0 && xx < w) m = fmaxf(m, g[(size_t)yy * w + xx]);
    }
  return v >= m;
}

// Adds cnt to hist[key] for every lane with has set, one atomic per distinct
// key of the wave: plateaus make many pixels peaks with the same value, and
// per-pixel atomics on one bin would serialise.  Wave-uniform control flow:
// every lane of the wave must call it.
__device__ __forceinline__ void wave_hist_add(uint32_t* __restrict__ hist, uint32_t key, uint32_t cnt, bool has) {
  const int lane = (int)__lane_id();   // the hardware lane (blocks may be 2-D)
  unsigned long long pending = __ballot(has);
  while (pending != 0ull) {
    const int leader = __ffsll(pending) - 1;
    const uint32_t lkey = (uint32_t)__shfl((int)key, leader, 64);
    const bool same = has && key == lkey;
    uint32_t v = same ? cnt : 0u;
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
    if (lane == leader) atomicAdd(&hist[lkey], v);
    has = has && !same;
    pending = __ballot(has);
  }
}

__global__ void __launch_bounds__(256) k_peak_hist(const float* __restrict__ g, int h, int w, double floor_v,
                                                   uint32_t* __restrict__ hist) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  float v = 0.f;
  const bool peak = x < w && y < h && is_peak(g, h, w, y, x, floor_v, v);
  wave_hist_add(hist, float_key(v) >> 16, 1u, peak);
}

// low-16-bit histograms of the peaks whose high 16 bits are bin_a / bin_b
__global__ void __launch_bounds__(256) k_peak_hist_low(const float* __restrict__ g, int h, int w, double floor_v,
                                                       uint32_t bin_a, uint32_t bin_b, uint32_t* __restrict__ hist_a,
                                                       uint32_t* __restrict__ hist_b) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  float v = 0.f;
  const bool peak = x < w && y < h && is_peak(g, h, w, y, x, floor_v, v);
  const uint32_t k = float_key(v);
  wave_hist_add(hist_a, k & 0xffffu, 1u, peak && (k >> 16) == bin_a);
  wave_hist_add(hist_b, k & 0xffffu, 1u, peak && (k >> 16) == bin_b);
}

__global__ void k_threshold(const float* __restrict__ g, size_t n, double thr, uint8_t* __restrict__ out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (double)g[i] >= thr ? 1 : 0;
}

__device__ __forceinline__ int ccl_find(int* L, int x) {
  int p = __hip_atomic_load(&L[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (p != x) {
    x = p;
    p = __hip_atomic_load(&L[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return x;
}

// Links the roots of a and b, the larger root to the smaller; retried when
// another thread relinked a root meanwhile (every retry sees a smaller root,
// so the loop ends).
__device__ __forceinline__ void ccl_union(int* L, int a, int b) {
  for (;;) {
    a = ccl_find(L, a);
    b = ccl_find(L, b);
    if (a == b) return;
    if (a < b) {
      const int old = atomicMin(&L[b], a);
      if (old == b) return;
      b = old;
    } else {
      const int old = atomicMin(&L[a], b);
      if (old == a) return;
      a = old;
    }
  }
}

__global__ void k_ccl_init(const uint8_t* __restrict__ m, size_t n, int* __restrict__ L, int* __restrict__ area) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    L[i] = m[i] ? (int)i : -1;
    area[i] = 0;
  }
}

__global__ void k_ccl_union(const uint8_t* __restrict__ m, int h, int w, int* L) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  const int i = y * w + x;
  if (!m[i]) return;
  // the neighbours that precede the pixel in raster order: left, up-left, up, up-right
  if (x > 0 && m[i - 1]) ccl_union(L, i, i - 1);
  if (y > 0) {
    if (x > 0 && m[i - w - 1]) ccl_union(L, i, i - w - 1);
    if (m[i - w]) ccl_union(L, i, i - w);
    if (x + 1 < w && m[i - w + 1]) ccl_union(L, i, i - w + 1);
  }
}

// One pointer-jumping pass over the union-find forest, out[i] = in[in[i]]:
// every pass halves the depth of every chain (the merge phase links roots in
// whatever order the threads ran, so chains across a large component can be
// long).  changed is set when some label moved.
__global__ void k_ccl_jump(size_t n, const int* __restrict__ in, int* __restrict__ out, int* __restrict__ changed) {
  bool moved = false;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int p = in[i];
    const int q = p >= 0 ? in[p] : p;
    out[i] = q;
    moved = moved || q != p;
  }
  if (__ballot(moved) != 0ull && (threadIdx.x & 63) == 0) atomicOr(changed, 1);
}

// Pixel areas per root (labels fully compressed: L[i] is the root), one
// atomic per distinct root of a wave.
__global__ void __launch_bounds__(256) k_ccl_area(size_t n, const int* __restrict__ L, int* __restrict__ area) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t n_up = (n + stride - 1) / stride * stride;   // whole waves take part in every round
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_up; i += stride) {
    const int r = i < n ? L[i] : -1;
    wave_hist_add(reinterpret_cast<uint32_t*>(area), (uint32_t)(r < 0 ? 0 : r), 1u, r >= 0);
  }
}

// the area filter: keep pixels whose component has area >= thr
__global__ void k_ccl_keep(size_t n, const int* __restrict__ L, const int* __restrict__ area, int thr,
                           uint8_t* __restrict__ out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int r = L[i];
    out[i] = r >= 0 && area[r] >= thr ? 1 : 0;
  }
}

// distance (<= r, else r + 1) to the nearest pixel of tgt in the same row
__global__ void k_row_dist(const float* __restrict__ t, int h, int w, int r, uint8_t* __restrict__ d) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  const float* row = t + (size_t)y * w;
  int best = r + 1;
  for (int j = -r; j <= r; ++j) {
    const int xx = x + j;
    if (xx >= 0 && xx < w && (double)row[xx] > 0.5) best = min(best, abs(j));
  }
  d[(size_t)y * w + x] = (uint8_t)best;
}

// keep_out(y, x) = some row y + dy of the element with a target pixel within
// its half-width dx(dy) of x
__global__ void k_keep_out(const uint8_t* __restrict__ d, int h, int w, int r, const int* __restrict__ half,
                           uint8_t* __restrict__ k) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  uint8_t on = 0;
  for (int dy = -r; dy <= r; ++dy) {
    const int yy = y + dy, hw = half[dy + r];
    if (yy >= 0 && yy < h && hw >= 0 && (int)d[(size_t)yy * w + x] <= hw) on = 1;
  }
  k[(size_t)y * w + x] = on;
}

// sraf_initial = filtered & !tgt, sraf_final = sraf_initial & !keep,
// pre_final = tgt | sraf_final; optional float copies of the stages
__global__ void k_sraf(const uint8_t* __restrict__ filtered, const float* __restrict__ t,
                       const uint8_t* __restrict__ keep, size_t n, uint8_t* __restrict__ pre_final,
                       const uint8_t* __restrict__ initial, float* __restrict__ o_initial,
                       float* __restrict__ o_filtered, float* __restrict__ o_sraf_i, float* __restrict__ o_sraf_f,
                       float* __restrict__ o_keep) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const uint8_t tg = (double)t[i] > 0.5 ? 1 : 0;
    const uint8_t kp = keep != nullptr ? keep[i] : 0;
    const uint8_t si = filtered[i] & (uint8_t)(1 - tg);
    const uint8_t sf = si & (uint8_t)(1 - kp);
    pre_final[i] = tg | sf;
    if (o_initial) o_initial[i] = (float)initial[i];
    if (o_filtered) o_filtered[i] = (float)filtered[i];
    if (o_sraf_i) o_sraf_i[i] = (float)si;
    if (o_sraf_f) o_sraf_f[i] = (float)sf;
    if (o_keep) o_keep[i] = (float)kp;
  }
}

__global__ void k_u8_to_float(const uint8_t* __restrict__ a, size_t n, float* __restrict__ out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)a[i];
}

int grid1(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

// bin holding rank `rank` (0-based) of a histogram, and the rank inside it
void locate(const std::vector<uint32_t>& hist, long long rank, uint32_t* bin, long long* inside) {
  long long acc = 0;
  for (uint32_t b = 0; b < (uint32_t)hist.size(); ++b) {
    if (acc + hist[b] > rank) {
      *bin = b;
      *inside = rank - acc;
      return;
    }
    acc += hist[b];
  }
  throw Error(5, "binarize: rank beyond the histogram");
}

}  // namespace

void ellipse_half_widths(int r, int* dx_out) {
  // getStructuringElement(MORPH_ELLIPSE, Size(2r+1, 2r+1)): ksize 1 is a rect
  if (r == 0) {
    dx_out[0] = 0;
    return;
  }
  const int c = r;
  const double inv_r2 = 1.0 / ((double)r * r);
  for (int dy = -r; dy <= r; ++dy) {
    const int dx = (int)std::nearbyint(c * std::sqrt((double)(r * r - dy * dy) * inv_r2));
    // columns [max(c - dx, 0), min(c + dx + 1, 2r + 1)) are c +- dx for dx <= r
    dx_out[dy + r] = std::min(dx, r);
  }
}

BinarizeInfo binarize_advanced(Context& ctx, int h, int w, const float* gray, const float* target,
                               double peak_min, double q, int area_thr, int spacing, const BinarizeOutputs& out) {
  if (h <= 0 || w <= 0) throw Error(1, "binarize_advanced: empty image");
  if ((long long)h * w >= (1ll << 31)) throw Error(1, "binarize_advanced: image above 2^31 pixels");
  if (gray == nullptr || target == nullptr || out.final_mask == nullptr)
    throw Error(1, "binarize_advanced: null gray mask, target or final mask");
  if (spacing < 0 || spacing > 254) throw Error(1, "binarize_advanced: min_spacing_pixels must be in [0, 254]");
  if (!std::isfinite(peak_min) || !std::isfinite(q)) throw Error(1, "binarize_advanced: non-finite parameter");
  hipStream_t s = ctx.stream;
  const size_t n = (size_t)h * w;
  const dim3 blk(32, 8), grd((w + 31) / 32, (h + 7) / 8);
  BinarizeInfo info;
  // peaks: high-16-bit histogram of the keys
  DevMem hist(3 * (size_t)kBins * sizeof(uint32_t));
  uint32_t* h_hi = hist.as<uint32_t>();
  LITHO_HIP(hipMemsetAsync(hist.ptr, 0, hist.bytes, s));
  hipLaunchKernelGGL(k_peak_hist, grd, blk, 0, s, gray, h, w, peak_min, h_hi);
  std::vector<uint32_t> hh(kBins), ha(kBins), hb(kBins);
  LITHO_HIP(hipMemcpyAsync(hh.data(), h_hi, kBins * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  LITHO_HIP(hipStreamSynchronize(s));
  long long np = 0;
  for (uint32_t c : hh) np += c;
  info.num_peaks = np;
  if (np == 0) {   // imbinarize.cpp:108-117: every mask empty
    for (float* o : {out.final_mask, out.initial_mask, out.filtered_mask, out.sraf_initial, out.sraf_final,
                     out.keep_out})
      if (o != nullptr) LITHO_HIP(hipMemsetAsync(o, 0, n * sizeof(float), s));
    LITHO_HIP(hipStreamSynchronize(s));
    return info;
  }
  const double qq = std::min(std::max(q, 0.0), 1.0);
  const double pos = qq * (double)(np - 1);
  const long long lo = (long long)std::floor(pos), hi = (long long)std::ceil(pos);
  uint32_t bin_a = 0, bin_b = 0;
  long long in_a = 0, in_b = 0;
  locate(hh, lo, &bin_a, &in_a);
  locate(hh, hi, &bin_b, &in_b);
  hipLaunchKernelGGL(k_peak_hist_low, grd, blk, 0, s, gray, h, w, peak_min, bin_a, bin_b, h_hi + kBins,
                     h_hi + 2 * kBins);
  LITHO_HIP(hipMemcpyAsync(ha.data(), h_hi + kBins, kBins * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  LITHO_HIP(hipMemcpyAsync(hb.data(), h_hi + 2 * kBins, kBins * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  LITHO_HIP(hipStreamSynchronize(s));
  uint32_t low_a = 0, low_b = 0;
  long long dummy = 0;
  locate(ha, in_a, &low_a, &dummy);
  locate(hb, in_b, &low_b, &dummy);
  const double v_lo = (double)key_float((bin_a << 16) | low_a);
  const double v_hi = (double)key_float((bin_b << 16) | low_b);
  const double thr = lo == hi ? v_lo : v_lo + (v_hi - v_lo) * (pos - (double)lo);
  info.threshold = thr;

  DevMem initial(n), filtered(n), pre(n), fin(n), keep(spacing > 0 ? n : 0), dist(spacing > 0 ? n : 0);
  DevMem labels(n * sizeof(int)), labels2(n * sizeof(int)), area(n * sizeof(int)), half((2 * 256 + 1) * sizeof(int));
  DevMem changed(sizeof(int));
  auto components = [&](const uint8_t* m, uint8_t* kept) {
    hipLaunchKernelGGL(k_ccl_init, dim3(grid1(n)), dim3(256), 0, s, m, n, labels.as<int>(), area.as<int>());
    hipLaunchKernelGGL(k_ccl_union, grd, blk, 0, s, m, h, w, labels.as<int>());
    // compress by pointer jumping until no label moves (at most ~log2 of the
    // longest chain passes; checked every four passes)
    int* a = labels.as<int>();
    int* b = labels2.as<int>();
    for (int round = 0;; ++round) {
      if (round > 16) throw Error(5, "binarize: label compression did not converge");
      LITHO_HIP(hipMemsetAsync(changed.ptr, 0, sizeof(int), s));
      for (int k = 0; k < 4; ++k) {
        hipLaunchKernelGGL(k_ccl_jump, dim3(grid1(n)), dim3(256), 0, s, n, a, b, changed.as<int>());
        std::swap(a, b);
      }
      int moved = 0;
      LITHO_HIP(hipMemcpyAsync(&moved, changed.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
      LITHO_HIP(hipStreamSynchronize(s));
      if (!moved) break;
    }
    hipLaunchKernelGGL(k_ccl_area, dim3(grid1(n)), dim3(256), 0, s, n, a, area.as<int>());
    hipLaunchKernelGGL(k_ccl_keep, dim3(grid1(n)), dim3(256), 0, s, n, a, area.as<int>(), area_thr, kept);
  };
  hipLaunchKernelGGL(k_threshold, dim3(grid1(n)), dim3(256), 0, s, gray, n, thr, initial.as<uint8_t>());
  components(initial.as<uint8_t>(), filtered.as<uint8_t>());
  if (spacing > 0) {
    std::vector<int> hw(2 * spacing + 1);
    ellipse_half_widths(spacing, hw.data());
    LITHO_HIP(hipMemcpyAsync(half.ptr, hw.data(), hw.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_row_dist, grd, blk, 0, s, target, h, w, spacing, dist.as<uint8_t>());
    hipLaunchKernelGGL(k_keep_out, grd, blk, 0, s, dist.as<uint8_t>(), h, w, spacing, half.as<int>(),
                       keep.as<uint8_t>());
  }
  hipLaunchKernelGGL(k_sraf, dim3(grid1(n)), dim3(256), 0, s, filtered.as<uint8_t>(), target,
                     spacing > 0 ? keep.as<uint8_t>() : nullptr, n, pre.as<uint8_t>(), initial.as<uint8_t>(),
                     out.initial_mask, out.filtered_mask, out.sraf_initial, out.sraf_final, out.keep_out);
  components(pre.as<uint8_t>(), fin.as<uint8_t>());
  hipLaunchKernelGGL(k_u8_to_float, dim3(grid1(n)), dim3(256), 0, s, fin.as<uint8_t>(), n, out.final_mask);
  LITHO_HIP(hipGetLastError());
  LITHO_HIP(hipStreamSynchronize(s));   // the scratch buffers are freed on return
  return info;
}

}  // namespace litho
