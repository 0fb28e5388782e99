// Level-set kernels for gfx950: WENO5-JS / Godunov normal term, curvature,
// CFL step and RK2 stages of the full-chip LSM update, and the 5x5 chamfer
// signed-distance re-initialisation.
//
// Restates source/utils/level_set_utils.cpp:
//   pad_edge (:19-43)            -> clamped indices
//   weno5_1d / der_weno5 (:45-148), evolve_normal_WENO_godunov (:150-225)
//                                -> ls_normal_term
//   _gradient_1d / evolve_kappa (:228-278) -> ls_kappa_term
//   get_dt_normal_kappa (:280-288) -> k_lsm_dt_final
//   reinit_sdf (:290-319)        -> k_chamfer + k_sdf_combine
// and the RK2 update of source/optimizer/full_chip_lsm.cpp:208-242.
//
// The stencils read float32 phi and compute in FP64 (the reference's
// arithmetic; the WENO epsilon 1e-6 max v^2 + 1e-99 is not representable in
// FP32).  Edge padding by 3 (twice, as the reference pads the already padded
// field) reduces to clamping every stencil index into the tile.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "levelset.hpp"

namespace litho {

namespace {

__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// 1/x for the WENO weights: hardware reciprocal refined by two Newton-Raphson
// steps r <- r + r(1 - x r), which square the relative error each time (well
// below one ulp of double from the estimate's ~2^-26 start).  x lies in
// [1e-198, ~1e300] here, so the range scaling of a full IEEE division is not
// needed; the result differs from 1/x by at most an ulp or so.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
}

// WENO5-JS combination of five one-sided differences (level_set_utils.cpp:56-111).
__device__ __forceinline__ double weno5(double v1, double v2, double v3, double v4, double v5) {
  const double c13 = 1.0 / 3.0, c76 = 7.0 / 6.0, c116 = 11.0 / 6.0, c16 = 1.0 / 6.0, c56 = 5.0 / 6.0;
  const double cs1 = 13.0 / 12.0, cs2 = 0.25;
  const double d1 = c13 * v1 - c76 * v2 + c116 * v3;
  const double d2 = -c16 * v2 + c56 * v3 + c13 * v4;
  const double d3 = c13 * v3 + c56 * v4 - c16 * v5;
  const double t1 = v1 - 2.0 * v2 + v3, t2 = v2 - 2.0 * v3 + v4, t3 = v3 - 2.0 * v4 + v5;
  const double e1 = v1 - 4.0 * v2 + 3.0 * v3, e2 = v2 - v4, e3 = 3.0 * v3 - 4.0 * v4 + v5;
  const double s1 = cs1 * t1 * t1 + cs2 * e1 * e1;
  const double s2 = cs1 * t2 * t2 + cs2 * e2 * e2;
  const double s3 = cs1 * t3 * t3 + cs2 * e3 * e3;
  const double vmax2 = fmax(fmax(fmax(v1 * v1, v2 * v2), fmax(v3 * v3, v4 * v4)), v5 * v5);
  const double eps = 1e-6 * vmax2 + 1e-99;
  // (s + eps)^2 >= 1e-198, and the weights stay below ~1e198
  const double a1 = 0.1 * rcp_nr((s1 + eps) * (s1 + eps));
  const double a2 = 0.6 * rcp_nr((s2 + eps) * (s2 + eps));
  const double a3 = 0.3 * rcp_nr((s3 + eps) * (s3 + eps));
  return (a1 * d1 + a2 * d2 + a3 * d3) * rcp_nr(a1 + a2 + a3);
}

// phi of one Micro at absolute (yy, xx), clamped into the tile (the
// reference's edge padding), read from global memory.
struct GlobalPhi {
  const float* t;
  int rows, cols;
  __device__ __forceinline__ double operator()(int yy, int xx) const {
    return (double)t[(size_t)clampi(yy, rows - 1) * cols + clampi(xx, cols - 1)];
  }
};

// The same from an LDS copy of rows [y0 - 3, y0 + TY + 3) x columns
// [x0 - 3, x0 + TX + 3) filled with the clamped values: every stencil access
// lies within 3 pixels of its output pixel.
template <int W>
struct LdsPhi {
  const float* s;
  int y0, x0;
  __device__ __forceinline__ double operator()(int yy, int xx) const {
    return (double)s[(yy - y0 + 3) * W + (xx - x0 + 3)];
  }
};

// f(o) = phi at offset o in [-3, 3] along the line through the pixel.
// minus: v_m = (f[j+m-2] - f[j+m-3]) / d, m = 0..4; plus: v_q = (f[j+4-q] - f[j+3-q]) / d.
template <class F>
__device__ __forceinline__ void weno_pair(const F& f, double inv_d, double& minus, double& plus) {
  double f_[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) f_[o] = f(o - 3);    // f[j-3] .. f[j+3]
  double d[6];
#pragma unroll
  for (int o = 0; o < 6; ++o) d[o] = (f_[o + 1] - f_[o]) * inv_d;  // d[o] = f[j+o-2] - f[j+o-3]
  minus = weno5(d[0], d[1], d[2], d[3], d[4]);
  plus = weno5(d[5], d[4], d[3], d[2], d[1]);
}

// Godunov |grad phi| for the sign of vn (level_set_utils.cpp:179-214).
__device__ __forceinline__ double godunov(double vn, double pxm, double pxp, double pym, double pyp) {
  double gx = 0.0, gy = 0.0;
  if (vn > 0.0) {
    const double a = fmax(pxm, 0.0), b = fmin(pxp, 0.0), c = fmax(pym, 0.0), e = fmin(pyp, 0.0);
    gx = fmax(a * a, b * b);
    gy = fmax(c * c, e * e);
  } else if (vn < 0.0) {
    const double a = fmin(pxm, 0.0), b = fmax(pxp, 0.0), c = fmin(pym, 0.0), e = fmax(pyp, 0.0);
    gx = fmax(a * a, b * b);
    gy = fmax(c * c, e * e);
  }
  return sqrt(gx + gy);
}

// |grad phi| normal term at (y, x): Vn |grad phi|.
template <class P>
__device__ __forceinline__ double ls_normal_term(const P& phi, int y, int x, double vn, double inv_dx, double inv_dy) {
  if (vn == 0.0) return 0.0;
  double pxm, pxp, pym, pyp;
  weno_pair([&](int o) { return phi(y, x + o); }, inv_dx, pxm, pxp);
  weno_pair([&](int o) { return phi(y + o, x); }, inv_dy, pym, pyp);
  return vn * godunov(vn, pxm, pxp, pym, pyp);
}

// _gradient_1d at index j of a line of length n (n >= 3) given a getter;
// inv_d2 = 1 / (2 d) (exact for the unit grid spacing of the reference).
template <class G>
__device__ __forceinline__ double grad1(const G& g, int j, int n, double inv_d2) {
  if (j == 0) return (-3.0 * g(0) + 4.0 * g(1) - g(2)) * inv_d2;
  if (j == n - 1) return (3.0 * g(n - 1) - 4.0 * g(n - 2) + g(n - 3)) * inv_d2;
  return (g(j + 1) - g(j - 1)) * inv_d2;
}

// evolve_kappa at (y, x) (level_set_utils.cpp:257-278).  The one-sided
// differences at the tile edges keep every access inside the tile and
// within 3 pixels of (y, x).
template <class P>
__device__ __forceinline__ double ls_kappa_term(const P& phi, int rows, int cols, int y, int x, double dx, double dy,
                                                double b) {
  const double inv_dx2 = 1.0 / (2.0 * dx), inv_dy2 = 1.0 / (2.0 * dy);
  auto px_at = [&](int yy, int xx) { return grad1([&](int k) { return phi(yy, k); }, xx, cols, inv_dx2); };
  auto py_at = [&](int yy, int xx) { return grad1([&](int k) { return phi(k, xx); }, yy, rows, inv_dy2); };
  const double px = px_at(y, x), py = py_at(y, x);
  const double pxy = grad1([&](int k) { return px_at(k, x); }, y, rows, inv_dy2);
  const double pxx = grad1([&](int k) { return px_at(y, k); }, x, cols, inv_dx2);
  const double pyy = grad1([&](int k) { return py_at(k, x); }, y, rows, inv_dy2);
  const double g2 = px * px + py * py;
  const double num = pxx * py * py - 2.0 * px * py * pxy + pyy * px * px;
  return b * (g2 > 2.220446049250313e-16 ? num / g2 : 0.0);
}

__device__ __forceinline__ bool in_rect(const int* r, int y, int x) {
  return y >= r[0] && y < r[1] && x >= r[2] && x < r[3];
}

}  // namespace

// ---------------------------------------------------------------- terms
__global__ void __launch_bounds__(256)
k_ls_terms(const float* __restrict__ phi, const float* __restrict__ vn, int rows, int cols, double dx, double dy,
           double b, float* __restrict__ delta, float* __restrict__ kappa) {
  const int t = blockIdx.y;
  const size_t plane = (size_t)rows * cols;
  const GlobalPhi tile{phi + t * plane, rows, cols};
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < plane; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / cols), x = (int)(p - (size_t)y * cols);
    if (delta != nullptr)
      delta[t * plane + p] = (float)ls_normal_term(tile, y, x, (double)vn[t * plane + p], 1.0 / dx, 1.0 / dy);
    if (kappa != nullptr) kappa[t * plane + p] = (float)ls_kappa_term(tile, rows, cols, y, x, dx, dy, b);
  }
}

void launch_ls_terms(const float* phi, const float* vn, int batch, int rows, int cols, double dx, double dy,
                     double b, float* delta, float* kappa, hipStream_t s) {
  const size_t plane = (size_t)rows * cols;
  const int blocks = (int)((plane + 255) / 256 < 1024 ? (plane + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_ls_terms, dim3(blocks, batch), dim3(256), 0, s, phi, vn, rows, cols, dx, dy, b, delta, kappa);
}

// ---------------------------------------------------------------- CFL step
__global__ void __launch_bounds__(256)
k_lsm_absmax(const float* __restrict__ grad, const int* __restrict__ rects, int n, float* __restrict__ partial) {
  const int t = blockIdx.y;
  const int* r = rects + 4 * t;
  const int h = r[1] - r[0], w = r[3] - r[2];
  const size_t cnt = (h > 0 && w > 0) ? (size_t)h * w : 0;
  const float* g = grad + (size_t)t * n * n;
  float m = 0.f;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
    const int yy = r[0] + (int)(q / w), xx = r[2] + (int)(q % w);
    m = fmaxf(m, fabsf(g[(size_t)yy * n + xx]));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  __shared__ float sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[t * gridDim.x + blockIdx.x] = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
}

__global__ void k_lsm_dt_final(const float* __restrict__ partial, int nblk, double dx, double dy, double cfl,
                               double b, double* __restrict__ dt) {
  const int t = blockIdx.x;
  if (threadIdx.x != 0) return;
  float m = 0.f;
  for (int i = 0; i < nblk; ++i) m = fmaxf(m, partial[t * nblk + i]);
  const double normal = (double)m / dx + (double)m / dy;
  const double kap = b * (1.0 / (dx * dx) + 1.0 / (dy * dy)) * 2.0;
  const double den = normal + kap;
  dt[t] = den < 1e-9 ? 1e-3 : cfl / den;
}

void launch_lsm_dt(const float* grad, const int* rects, int batch, int n, double dx, double dy, double cfl, double b,
                   float* partial_max, double* dt, hipStream_t s) {
  hipLaunchKernelGGL(k_lsm_absmax, dim3(kLsmDtBlocks, batch), dim3(256), 0, s, grad, rects, n, partial_max);
  hipLaunchKernelGGL(k_lsm_dt_final, dim3(batch), dim3(64), 0, s, partial_max, kLsmDtBlocks, dx, dy, cfl, b, dt);
}

// ---------------------------------------------------------------- RK2 stage
// One block per 32 x 8 pixels of one Micro: the clamped 38 x 14 neighbourhood
// of phi_in goes to LDS once, and every stencil of the block reads it there.
constexpr int kStageTX = 32, kStageTY = 8;
constexpr int kStageW = kStageTX + 6, kStageH = kStageTY + 6;

template <int STAGE>
__global__ void __launch_bounds__(kStageTX * kStageTY)
k_lsm_stage(const float* __restrict__ phi_in, const float* base, const float* __restrict__ grad,
            const int* __restrict__ rects, const double* __restrict__ dt, int n, double dx, double dy, double b,
            float* out, float* __restrict__ mask_out) {
  // out may alias base (an in-place update): both are touched at the thread's
  // own pixel only; phi_in is never written by this launch
  __shared__ float sp[kStageH * kStageW];
  const int t = blockIdx.z;
  const size_t plane = (size_t)n * n;
  const float* tile = phi_in + t * plane;
  const int x0 = blockIdx.x * kStageTX, y0 = blockIdx.y * kStageTY;
  for (int e = threadIdx.x; e < kStageH * kStageW; e += kStageTX * kStageTY) {
    const int ly = e / kStageW, lx = e - ly * kStageW;
    sp[e] = tile[(size_t)clampi(y0 + ly - 3, n - 1) * n + clampi(x0 + lx - 3, n - 1)];
  }
  __syncthreads();
  const int x = x0 + (int)threadIdx.x % kStageTX, y = y0 + (int)threadIdx.x / kStageTX;
  if (x >= n || y >= n) return;
  const size_t p = (size_t)y * n + x;
  const LdsPhi<kStageW> phi{sp, y0, x0};
  const int* r = rects + 4 * t;
  double v;
  if (in_rect(r, y, x)) {
    const double vn = (double)grad[t * plane + p];
    const double normal = ls_kappa_term(phi, n, n, y, x, dx, dy, b) - ls_normal_term(phi, y, x, vn, 1.0 / dx, 1.0 / dy);
    const double cur = phi(y, x);
    if (STAGE == 1)
      v = cur + dt[t] * normal;
    else
      v = 0.5 * (double)base[t * plane + p] + 0.5 * (cur + dt[t] * normal);
  } else {
    v = STAGE == 1 ? phi(y, x) : (double)base[t * plane + p];
  }
  const float vf = (float)v;
  out[t * plane + p] = vf;
  if (mask_out != nullptr) mask_out[t * plane + p] = vf >= 0.f ? 1.f : 0.f;
}

void launch_lsm_stage(int stage, const float* phi_in, const float* base, const float* grad, const int* rects,
                      const double* dt, int batch, int n, double dx, double dy, double b, float* out, float* mask_out,
                      hipStream_t s) {
  const dim3 grid((n + kStageTX - 1) / kStageTX, (n + kStageTY - 1) / kStageTY, batch);
  if (stage == 1)
    hipLaunchKernelGGL(k_lsm_stage<1>, grid, dim3(kStageTX * kStageTY), 0, s, phi_in, base, grad, rects, dt, n, dx, dy,
                       b, out, mask_out);
  else
    hipLaunchKernelGGL(k_lsm_stage<2>, grid, dim3(kStageTX * kStageTY), 0, s, phi_in, base, grad, rects, dt, n, dx, dy,
                       b, out, mask_out);
}

// ---------------------------------------------------------------- render
// render_levelset_mask (levelset_mask_renderer.cpp:7-33).  Binary: phi >= 0.
// Gray: the float32 phi upscaled by `scale` with pixel-centre bilinear
// interpolation and replicated borders (cv::resize INTER_LINEAR: source
// coordinate f = (d + 0.5) / scale - 0.5 in FP64 rounded to float32, floor
// and fraction, clamped at both ends with a zero fraction), each output row
// interpolated along x first and the two rows then along y, every product and
// sum rounded separately in float32 (no contraction); the subpixels >= 0 of
// each pixel's scale x scale block are counted and divided by scale^2.
namespace {
struct LinearTap {
  int s0, s1;
  float w0, w1;
};
__device__ __forceinline__ LinearTap linear_tap(int d, int n_src, double inv_scale) {
  float f = (float)__dadd_rn(__dmul_rn((double)d + 0.5, inv_scale), -0.5);
  int s = (int)floorf(f);
  f = __fsub_rn(f, (float)s);
  if (s < 0) {
    s = 0;
    f = 0.f;
  }
  if (s >= n_src - 1) {
    s = n_src - 1;
    f = 0.f;
  }
  return {s, s + 1 < n_src ? s + 1 : n_src - 1, __fsub_rn(1.f, f), f};
}
}  // namespace

__global__ void __launch_bounds__(256)
k_render_mask(const float* __restrict__ phi, int rows, int cols, int type, int scale, float* __restrict__ mask) {
  const int t = blockIdx.y;
  const size_t plane = (size_t)rows * cols;
  const float* tile = phi + t * plane;
  const double inv_scale = 1.0 / (double)scale;
  const double count = (double)scale * (double)scale;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < plane; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / cols), x = (int)(p - (size_t)y * cols);
    float v;
    if (type == 0) {
      v = tile[p] >= 0.f ? 1.f : 0.f;
    } else {
      int inside = 0;
      for (int sy = 0; sy < scale; ++sy) {
        const LinearTap ty = linear_tap(y * scale + sy, rows, inv_scale);
        const float* r0 = tile + (size_t)ty.s0 * cols;
        const float* r1 = tile + (size_t)ty.s1 * cols;
        for (int sx = 0; sx < scale; ++sx) {
          const LinearTap tx = linear_tap(x * scale + sx, cols, inv_scale);
          const float h0 = __fadd_rn(__fmul_rn(r0[tx.s0], tx.w0), __fmul_rn(r0[tx.s1], tx.w1));
          const float h1 = __fadd_rn(__fmul_rn(r1[tx.s0], tx.w0), __fmul_rn(r1[tx.s1], tx.w1));
          const float val = __fadd_rn(__fmul_rn(h0, ty.w0), __fmul_rn(h1, ty.w1));
          inside += val >= 0.f ? 1 : 0;
        }
      }
      v = (float)((double)inside / count);
    }
    mask[t * plane + p] = v;
  }
}

void launch_render_mask(const float* phi, int batch, int rows, int cols, int type, int scale, float* mask,
                        hipStream_t s) {
  const size_t plane = (size_t)rows * cols;
  const int blocks = (int)((plane + 255) / 256 < 1024 ? (plane + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_render_mask, dim3(blocks, batch), dim3(256), 0, s, phi, rows, cols, type, scale, mask);
}

// ---------------------------------------------------------------- chamfer SDF
// OpenCV DIST_L2 / mask 5 semantics in 16-bit fixed point: horizontal and
// vertical step 1, diagonal 1.4, knight move 2.1969 (round(w * 65536) of the
// float32 weights); pixels outside the image hold INIT = 2^31 - 1.  The forward
// raster pass per row: t0 = min over the 7 upper neighbours (+ weights),
// u = src ? t0 : 0, then the in-row recurrence tmp[j] = min(u[j], tmp[j-1] + HV)
// which equals the prefix minimum of u[k] + HV (j - k), taken here by a
// block-wide scan.  The backward pass mirrors it from the bottom-right.
namespace {
constexpr long long kHV = 65536, kDiag = 91750, kLong = 143976, kInit = 2147483647LL;
constexpr long long kDistMax = kInit >> 2;
constexpr int kChamferThreads = 256;
constexpr int kChamferMaxCols = 4096;
constexpr int kChamferPer = kChamferMaxCols / kChamferThreads;

// Inclusive block scan (minimum) of one value per thread, Hillis-Steele in LDS.
__device__ long long block_min_scan(long long v, long long* sh, bool reverse) {
  const int t = reverse ? (kChamferThreads - 1 - (int)threadIdx.x) : (int)threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kChamferThreads; off <<= 1) {
    const long long o = t >= off ? sh[t - off] : kInit * 4;
    __syncthreads();
    if (o < v) v = o;
    sh[t] = v;
    __syncthreads();
  }
  const long long r = t > 0 ? sh[t - 1] : kInit * 4;   // exclusive
  __syncthreads();
  return r;
}
}  // namespace

// blockIdx.x = tile, blockIdx.y = 0: distance of outside pixels (src = !inside)
// to the inside, 1: of inside pixels to the outside.  Thread t owns columns
// [t * per, (t + 1) * per), per <= kChamferPer.
__global__ void __launch_bounds__(kChamferThreads)
k_chamfer(const float* __restrict__ phi, int rows, int cols, float level, int inclusive,
          uint32_t* __restrict__ scratch, float* __restrict__ dist) {
  __shared__ uint32_t ring[3][kChamferMaxCols];
  __shared__ long long sh[kChamferThreads];
  const int t = blockIdx.x, which = blockIdx.y;
  const size_t plane = (size_t)rows * cols;
  const float* ph = phi + t * plane;
  uint32_t* tmp = scratch + ((size_t)t * 2 + which) * plane;
  float* out = dist + ((size_t)t * 2 + which) * plane;
  const int per = (cols + kChamferThreads - 1) / kChamferThreads;
  const int j0 = threadIdx.x * per;
  const int j1 = j0 + per < cols ? j0 + per : cols;
  auto ringv = [&](int row, int j) -> long long {
    if (row < 0 || row >= rows || j < 0 || j >= cols) return kInit;
    return (long long)ring[row % 3][j];
  };
  long long w[kChamferPer];
  // forward pass, top to bottom
  for (int i = 0; i < rows; ++i) {
    long long run = kInit * 4;
#pragma unroll
    for (int e = 0; e < kChamferPer; ++e) {
      const int j = j0 + e;
      w[e] = kInit * 4;
      if (j < j1) {
        const float v = ph[(size_t)i * cols + j];
        const bool inside = inclusive ? v >= level : v > level;
        const bool src = which == 0 ? !inside : inside;
        long long t0 = 0;
        if (src) {
          t0 = ringv(i - 2, j - 1) + kLong;
          t0 = min(t0, ringv(i - 2, j + 1) + kLong);
          t0 = min(t0, ringv(i - 1, j - 2) + kLong);
          t0 = min(t0, ringv(i - 1, j - 1) + kDiag);
          t0 = min(t0, ringv(i - 1, j) + kHV);
          t0 = min(t0, ringv(i - 1, j + 1) + kDiag);
          t0 = min(t0, ringv(i - 1, j + 2) + kLong);
        }
        w[e] = t0 - kHV * j;            // prefix min of u[k] - HV k, plus HV j
        run = min(run, w[e]);
      }
    }
    long long acc = block_min_scan(run, sh, false);   // also orders the ring reads before the writes below
#pragma unroll
    for (int e = 0; e < kChamferPer; ++e) {
      const int j = j0 + e;
      if (j < j1) {
        acc = min(acc, w[e]);
        const long long left = kInit + kHV * (j + 1);
        const uint32_t val = (uint32_t)min(acc + kHV * j, left);
        ring[i % 3][j] = val;
        tmp[(size_t)i * cols + j] = val;
      }
    }
    __syncthreads();
  }
  // backward pass, bottom to top (the ring now carries backward rows)
  for (int i = rows - 1; i >= 0; --i) {
    long long run = kInit * 4;
#pragma unroll
    for (int e = 0; e < kChamferPer; ++e) {
      const int j = j0 + e;
      w[e] = kInit * 4;
      if (j < j1) {
        long long t0 = (long long)tmp[(size_t)i * cols + j];
        t0 = min(t0, ringv(i + 2, j + 1) + kLong);
        t0 = min(t0, ringv(i + 2, j - 1) + kLong);
        t0 = min(t0, ringv(i + 1, j + 2) + kLong);
        t0 = min(t0, ringv(i + 1, j + 1) + kDiag);
        t0 = min(t0, ringv(i + 1, j) + kHV);
        t0 = min(t0, ringv(i + 1, j - 1) + kDiag);
        t0 = min(t0, ringv(i + 1, j - 2) + kLong);
        w[e] = t0 + kHV * j;            // suffix min of t0[k] + HV k, minus HV j
        run = min(run, w[e]);
      }
    }
    long long acc = block_min_scan(run, sh, true);
#pragma unroll
    for (int e = kChamferPer - 1; e >= 0; --e) {
      const int j = j0 + e;
      if (j < j1) {
        acc = min(acc, w[e]);
        const long long right = kInit + kHV * (cols - j);
        const long long val = min(acc - kHV * j, right);
        ring[i % 3][j] = (uint32_t)val;
        out[(size_t)i * cols + j] = (float)min(val, kDistMax) * (1.0f / 65536.0f);
      }
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256)
k_sdf_combine(const float* phi, const float* __restrict__ dist, int rows, int cols, double dx,
              const int* __restrict__ rects, float* sdf, float* __restrict__ mask_out) {
  // sdf may alias phi (re-initialisation in place): both at the thread's own pixel only
  const int t = blockIdx.y;
  const size_t plane = (size_t)rows * cols;
  const float* dout = dist + ((size_t)t * 2 + 0) * plane;
  const float* din = dist + ((size_t)t * 2 + 1) * plane;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < plane; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / cols), x = (int)(p - (size_t)y * cols);
    float v;
    if (rects == nullptr || in_rect(rects + 4 * t, y, x)) {
      v = (float)((double)din[p] * dx - (double)dout[p] * dx);
      sdf[t * plane + p] = v;
    } else {
      v = phi[t * plane + p];
    }
    if (mask_out != nullptr) mask_out[t * plane + p] = v >= 0.f ? 1.f : 0.f;
  }
}

size_t chamfer_scratch_bytes(int batch, int rows, int cols) {
  return (size_t)batch * 2 * rows * cols * (sizeof(uint32_t) + sizeof(float));
}

void launch_reinit_sdf(const float* phi, int batch, int rows, int cols, double dx, double dy, float level, bool inclusive,
                       uint32_t* scratch, float* dist, const int* rects, float* sdf, float* mask_out, hipStream_t s) {
  (void)dy;   // the reference scales both distances by dx (level_set_utils.cpp:306-316)
  hipLaunchKernelGGL(k_chamfer, dim3(batch, 2), dim3(kChamferThreads), 0, s, phi, rows, cols, level,
                     inclusive ? 1 : 0, scratch, dist);
  const size_t plane = (size_t)rows * cols;
  const int blocks = (int)((plane + 255) / 256 < 1024 ? (plane + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_sdf_combine, dim3(blocks, batch), dim3(256), 0, s, phi, dist, rows, cols, dx, rects, sdf,
                     mask_out);
}

}  // namespace litho
