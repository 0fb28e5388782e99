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
// The stencils read float32 phi and compute in FP32 (the reference's FP64
// arithmetic on the same float32 values agrees to ~1e-7 relative: first
// differences of neighbouring values are exact, and the WENO weights are
// formed on power-of-two-scaled differences, where the reference's epsilon
// 1e-6 max v^2 + 1e-99 is representable).  Edge padding by 3 (twice, as the
// reference pads the already padded field) reduces to clamping every stencil
// index into the tile.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "levelset.hpp"

namespace litho {

namespace {

__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// WENO5-JS combination of five one-sided differences (level_set_utils.cpp:56-111)
// in FP32.  The nonlinear weights a_i = g_i / (beta_i + eps)^2 with
// eps = 1e-6 max v^2 (+ 1e-99) are invariant under scaling all v by one
// factor, up to the 1e-99 term, which only matters when every v is zero or
// within ~1e-46 of it: the differences are scaled by the power of two that
// brings max |v| into [0.5, 1) (exact), the weights formed there, and the
// candidate stencils combined in the original units.  All-zero differences
// give 0, as every candidate is 0.
__device__ __forceinline__ float weno5(float v1, float v2, float v3, float v4, float v5) {
  const float m = fmaxf(fmaxf(fmaxf(fabsf(v1), fabsf(v2)), fmaxf(fabsf(v3), fabsf(v4))), fabsf(v5));
  if (!(m > 0.f)) return 0.f;
  int e;
  (void)frexpf(m, &e);
  const float sc = ldexpf(1.f, -e);
  const float u1 = v1 * sc, u2 = v2 * sc, u3 = v3 * sc, u4 = v4 * sc, u5 = v5 * sc;
  const float c13 = 1.f / 3.f, c76 = 7.f / 6.f, c116 = 11.f / 6.f, c16 = 1.f / 6.f, c56 = 5.f / 6.f;
  const float cs1 = 13.f / 12.f, cs2 = 0.25f;
  const float d1 = c13 * v1 - c76 * v2 + c116 * v3;
  const float d2 = -c16 * v2 + c56 * v3 + c13 * v4;
  const float d3 = c13 * v3 + c56 * v4 - c16 * v5;
  const float t1 = u1 - 2.f * u2 + u3, t2 = u2 - 2.f * u3 + u4, t3 = u3 - 2.f * u4 + u5;
  const float e1 = u1 - 4.f * u2 + 3.f * u3, e2 = u2 - u4, e3 = 3.f * u3 - 4.f * u4 + u5;
  const float s1 = cs1 * t1 * t1 + cs2 * e1 * e1;
  const float s2 = cs1 * t2 * t2 + cs2 * e2 * e2;
  const float s3 = cs1 * t3 * t3 + cs2 * e3 * e3;
  const float mu = m * sc;                   // max |u| in [0.5, 1): (beta + eps)^2 >= 6e-14
  const float eps = 1e-6f * (mu * mu);
  const float q1 = s1 + eps, q2 = s2 + eps, q3 = s3 + eps;
  const float a1 = 0.1f * __builtin_amdgcn_rcpf(q1 * q1);
  const float a2 = 0.6f * __builtin_amdgcn_rcpf(q2 * q2);
  const float a3 = 0.3f * __builtin_amdgcn_rcpf(q3 * q3);
  return (a1 * d1 + a2 * d2 + a3 * d3) * __builtin_amdgcn_rcpf(a1 + a2 + a3);
}

// phi of one Micro at absolute (yy, xx), clamped into the tile (the
// reference's edge padding), read from global memory.
struct GlobalPhi {
  const float* t;
  int rows, cols;
  __device__ __forceinline__ float operator()(int yy, int xx) const {
    return t[(size_t)clampi(yy, rows - 1) * cols + clampi(xx, cols - 1)];
  }
};

// The same from an LDS copy of rows [y0 - 3, y0 + TY + 3) x columns
// [x0 - 3, x0 + TX + 3) filled with the clamped values: every stencil access
// lies within 3 pixels of its output pixel.
template <int W>
struct LdsPhi {
  const float* s;
  int y0, x0;
  __device__ __forceinline__ float operator()(int yy, int xx) const { return s[(yy - y0 + 3) * W + (xx - x0 + 3)]; }
};

// f(o) = phi at offset o in [-3, 3] along the line through the pixel.
// minus: v_m = (f[j+m-2] - f[j+m-3]) / d, m = 0..4; plus: v_q = (f[j+4-q] - f[j+3-q]) / d.
template <class F>
__device__ __forceinline__ void weno_pair(const F& f, float inv_d, float& minus, float& plus) {
  float f_[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) f_[o] = f(o - 3);    // f[j-3] .. f[j+3]
  float d[6];
#pragma unroll
  for (int o = 0; o < 6; ++o) d[o] = (f_[o + 1] - f_[o]) * inv_d;  // d[o] = f[j+o-2] - f[j+o-3]
  minus = weno5(d[0], d[1], d[2], d[3], d[4]);
  plus = weno5(d[5], d[4], d[3], d[2], d[1]);
}

// Godunov |grad phi| for the sign of vn (level_set_utils.cpp:179-214).
__device__ __forceinline__ float godunov(float vn, float pxm, float pxp, float pym, float pyp) {
  float gx = 0.f, gy = 0.f;
  if (vn > 0.f) {
    const float a = fmaxf(pxm, 0.f), b = fminf(pxp, 0.f), c = fmaxf(pym, 0.f), e = fminf(pyp, 0.f);
    gx = fmaxf(a * a, b * b);
    gy = fmaxf(c * c, e * e);
  } else if (vn < 0.f) {
    const float a = fminf(pxm, 0.f), b = fmaxf(pxp, 0.f), c = fminf(pym, 0.f), e = fmaxf(pyp, 0.f);
    gx = fmaxf(a * a, b * b);
    gy = fmaxf(c * c, e * e);
  }
  return sqrtf(gx + gy);
}

// |grad phi| normal term at (y, x): Vn |grad phi|.
template <class P>
__device__ __forceinline__ float ls_normal_term(const P& phi, int y, int x, float vn, float inv_dx, float inv_dy) {
  if (vn == 0.f) return 0.f;
  float pxm, pxp, pym, pyp;
  weno_pair([&](int o) { return phi(y, x + o); }, inv_dx, pxm, pxp);
  weno_pair([&](int o) { return phi(y + o, x); }, inv_dy, pym, pyp);
  return vn * godunov(vn, pxm, pxp, pym, pyp);
}

// 2 d times _gradient_1d at index j of a line of length n (n >= 3) given a
// getter: the raw central / one-sided combination, whose 1 / (2 d) scale the
// caller applies once at the end.  Differences of neighbouring float32 values
// within a factor of two of each other are exact in FP32, so the nested
// differences below stay exact on smooth fields and only the final products
// round.
// The one-sided ends -3 g0 + 4 g1 - g2 and 3 g[n-1] - 4 g[n-2] + g[n-3] are
// regrouped as combinations of neighbour differences, which keeps them exact
// up to one rounding of the small result.
// INTERIOR: the caller guarantees 0 < j < n - 1 (no edge tests).
template <bool INTERIOR = false, class G>
__device__ __forceinline__ float grad1u(const G& g, int j, int n) {
  if constexpr (!INTERIOR) {
    if (j == 0) return 3.f * (g(1) - g(0)) + (g(1) - g(2));
    if (j == n - 1) return 3.f * (g(n - 1) - g(n - 2)) - (g(n - 2) - g(n - 3));
  }
  return g(j + 1) - g(j - 1);
}

// evolve_kappa at (y, x) (level_set_utils.cpp:257-278).  The one-sided
// differences at the tile edges keep every access inside the tile and
// within 3 pixels of (y, x).  inv_dx2 = 1 / (2 dx), inv_dy2 = 1 / (2 dy).
// INTERIOR: (y, x) lies at least 2 pixels inside the tile, so every
// difference is the central one (the same arithmetic without edge tests).
template <bool INTERIOR = false, class P>
__device__ __forceinline__ float ls_kappa_term(const P& phi, int rows, int cols, int y, int x, float inv_dx2,
                                               float inv_dy2, float b) {
  auto px_at = [&](int yy, int xx) { return grad1u<INTERIOR>([&](int k) { return phi(yy, k); }, xx, cols); };
  auto py_at = [&](int yy, int xx) { return grad1u<INTERIOR>([&](int k) { return phi(k, xx); }, yy, rows); };
  const float px = px_at(y, x) * inv_dx2, py = py_at(y, x) * inv_dy2;
  const float pxy = grad1u<INTERIOR>([&](int k) { return px_at(k, x); }, y, rows) * (inv_dx2 * inv_dy2);
  const float pxx = grad1u<INTERIOR>([&](int k) { return px_at(y, k); }, x, cols) * (inv_dx2 * inv_dx2);
  const float pyy = grad1u<INTERIOR>([&](int k) { return py_at(k, x); }, y, rows) * (inv_dy2 * inv_dy2);
  const float g2 = px * px + py * py;
  const float num = pxx * py * py - 2.f * px * py * pxy + pyy * px * px;
  return b * (g2 > 2.220446049250313e-16f ? num / g2 : 0.f);
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
      delta[t * plane + p] = ls_normal_term(tile, y, x, vn[t * plane + p], (float)(1.0 / dx), (float)(1.0 / dy));
    if (kappa != nullptr)
      kappa[t * plane + p] =
          ls_kappa_term(tile, rows, cols, y, x, (float)(1.0 / (2.0 * dx)), (float)(1.0 / (2.0 * dy)), (float)b);
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

// The FP32 factors of the stencils, formed once on the host in FP64 exactly
// as the terms kernel forms them per thread.
struct StageScales {
  float inv_dx, inv_dy;     // 1 / dx, 1 / dy (normal term)
  float inv_dx2, inv_dy2;   // 1 / (2 dx), 1 / (2 dy) (curvature term)
  float b;                  // curvature weight
};

template <int STAGE>
__global__ void __launch_bounds__(kStageTX * kStageTY)
k_lsm_stage(const float* __restrict__ phi_in, const float* base, const float* __restrict__ grad,
            const int* __restrict__ rects, const double* __restrict__ dt, int n, StageScales sc, float* out,
            float* __restrict__ mask_out) {
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
  // blocks whose pixels all lie 2 or more inside the tile take the
  // curvature term without edge tests (block-uniform)
  const bool interior = x0 >= 2 && y0 >= 2 && x0 + kStageTX + 2 <= n && y0 + kStageTY + 2 <= n;
  float vf;
  if (in_rect(r, y, x)) {
    const float vn = grad[t * plane + p];
    const float kap = interior ? ls_kappa_term<true>(phi, n, n, y, x, sc.inv_dx2, sc.inv_dy2, sc.b)
                               : ls_kappa_term<false>(phi, n, n, y, x, sc.inv_dx2, sc.inv_dy2, sc.b);
    const float normal = kap - ls_normal_term(phi, y, x, vn, sc.inv_dx, sc.inv_dy);
    const float cur = phi(y, x);
    const float step = (float)dt[t];
    if (STAGE == 1)
      vf = fmaf(step, normal, cur);
    else
      vf = 0.5f * base[t * plane + p] + 0.5f * fmaf(step, normal, cur);
  } else {
    vf = STAGE == 1 ? phi(y, x) : base[t * plane + p];
  }
  out[t * plane + p] = vf;
  if (mask_out != nullptr) mask_out[t * plane + p] = vf >= 0.f ? 1.f : 0.f;
}

void launch_lsm_stage(int stage, const float* phi_in, const float* base, const float* grad, const int* rects,
                      const double* dt, int batch, int n, double dx, double dy, double b, float* out, float* mask_out,
                      hipStream_t s) {
  const dim3 grid((n + kStageTX - 1) / kStageTX, (n + kStageTY - 1) / kStageTY, batch);
  const StageScales sc{(float)(1.0 / dx), (float)(1.0 / dy), (float)(1.0 / (2.0 * dx)), (float)(1.0 / (2.0 * dy)),
                       (float)b};
  if (stage == 1)
    hipLaunchKernelGGL(k_lsm_stage<1>, grid, dim3(kStageTX * kStageTY), 0, s, phi_in, base, grad, rects, dt, n, sc,
                       out, mask_out);
  else
    hipLaunchKernelGGL(k_lsm_stage<2>, grid, dim3(kStageTX * kStageTY), 0, s, phi_in, base, grad, rects, dt, n, sc,
                       out, mask_out);
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

// Barrier that orders LDS accesses only: __syncthreads() also waits for the
// wave's outstanding global loads and stores, which would serialise the
// chamfer's next-row loads and its distance stores with every row.  (Every
// global location the chamfer reads back is one the same thread wrote.)
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exclusive block scan (minimum) of one value per thread: forward over the
// threads before t, reverse over the threads after t.  Shuffles within each
// wave, then the four wave totals through LDS (two barriers per row instead
// of the 16 of an LDS Hillis-Steele scan).
__device__ long long block_min_scan(long long v, long long* sh, bool reverse) {
  constexpr long long kBig = kInit * 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int kWaves = kChamferThreads / 64;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const long long o = reverse ? __shfl_down(v, off, 64) : __shfl_up(v, off, 64);
    const bool ok = reverse ? lane + off < 64 : lane >= off;
    if (ok && o < v) v = o;
  }
  if (reverse ? lane == 0 : lane == 63) sh[wave] = v;   // the wave's total
  // exclusive within the wave: the inclusive value of the neighbouring lane
  const long long nb = reverse ? __shfl_down(v, 1, 64) : __shfl_up(v, 1, 64);
  long long r = (reverse ? lane < 63 : lane > 0) ? nb : kBig;
  lds_barrier();
#pragma unroll
  for (int w = 0; w < kWaves; ++w)
    if (reverse ? w > wave : w < wave) r = min(r, sh[w]);
  lds_barrier();
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
  // this thread's phi values of the current row, loaded one row ahead
  float pv[kChamferPer];
#pragma unroll
  for (int e = 0; e < kChamferPer; ++e) pv[e] = j0 + e < j1 ? ph[j0 + e] : 0.f;
  // forward pass, top to bottom
  for (int i = 0; i < rows; ++i) {
    float cur[kChamferPer];
#pragma unroll
    for (int e = 0; e < kChamferPer; ++e) {
      cur[e] = pv[e];
      pv[e] = (i + 1 < rows && j0 + e < j1) ? ph[(size_t)(i + 1) * cols + j0 + e] : 0.f;
    }
    long long run = kInit * 4;
#pragma unroll
    for (int e = 0; e < kChamferPer; ++e) {
      const int j = j0 + e;
      w[e] = kInit * 4;
      if (j < j1) {
        const float v = cur[e];
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
    lds_barrier();
  }
  // backward pass, bottom to top (the ring now carries backward rows); the
  // forward distances of the next row up are loaded one row ahead (this
  // thread wrote them itself in the forward pass)
  uint32_t tv[kChamferPer];
#pragma unroll
  for (int e = 0; e < kChamferPer; ++e) tv[e] = j0 + e < j1 ? tmp[(size_t)(rows - 1) * cols + j0 + e] : 0u;
  for (int i = rows - 1; i >= 0; --i) {
    uint32_t tcur[kChamferPer];
#pragma unroll
    for (int e = 0; e < kChamferPer; ++e) {
      tcur[e] = tv[e];
      tv[e] = (i > 0 && j0 + e < j1) ? tmp[(size_t)(i - 1) * cols + j0 + e] : 0u;
    }
    long long run = kInit * 4;
#pragma unroll
    for (int e = 0; e < kChamferPer; ++e) {
      const int j = j0 + e;
      w[e] = kInit * 4;
      if (j < j1) {
        long long t0 = (long long)tcur[e];
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
    lds_barrier();
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
