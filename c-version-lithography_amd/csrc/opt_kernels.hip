// Device kernels of the CTM optimizers (gradient descent and L-BFGS) and small
// elementwise helpers.  All vectors are FP32; every dot product accumulates in
// FP64 and is reduced in a fixed order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include "opt_kernels.hpp"

namespace litho {

// CTM_Optimizer::initialize_theta (CTM_Optimizer.cpp:66-73)
__global__ void k_init_theta(const float* __restrict__ pattern, float* __restrict__ theta, size_t count) {
  constexpr double kPi = 3.14159265358979323846;
  const float hi = (float)(0.3 * kPi), lo = (float)(0.7 * kPi);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    theta[i] = pattern[i] > 0.f ? hi : lo;
}

void launch_init_theta(const float* pattern, float* theta, size_t count, hipStream_t s) {
  const int blocks = (int)((count + 255) / 256 < 4096 ? (count + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_init_theta, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, pattern, theta, count);
}

// y = coef[b] * x for active tiles (used for d = -g)
__global__ void k_scale(const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ coef,
                        const int* __restrict__ active, size_t n) {
  const int b = blockIdx.y;
  if (active != nullptr && active[b] == 0) return;
  const float c = coef[b];
  const float* xb = x + (size_t)b * n;
  float* yb = y + (size_t)b * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    yb[i] = c * xb[i];
}

void launch_scale(const float* x, float* y, const float* coef, const int* active, size_t n, int batch,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(kOptBlocks, batch), dim3(kOptThreads), 0, s, x, y, coef, active, n);
}

// y += coef[b] * x for active tiles
__global__ void k_axpy(const float* __restrict__ x, float* __restrict__ y, const float* __restrict__ coef,
                       const int* __restrict__ active, size_t n) {
  const int b = blockIdx.y;
  if (active != nullptr && active[b] == 0) return;
  const float c = coef[b];
  const float* xb = x + (size_t)b * n;
  float* yb = y + (size_t)b * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    yb[i] = fmaf(c, xb[i], yb[i]);
}

void launch_axpy(const float* x, float* y, const float* coef, const int* active, size_t n, int batch,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(kOptBlocks, batch), dim3(kOptThreads), 0, s, x, y, coef, active, n);
}

// mask = (1 + cos theta) / 2 (CTM_Optimizer.cpp:82)
__global__ void k_mask_from_theta(const float* __restrict__ theta, float* __restrict__ mask, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    mask[i] = 0.5f * (1.0f + cosf(theta[i]));
}

void launch_mask_from_theta(const float* theta, float* mask, size_t count, hipStream_t s) {
  const size_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_mask_from_theta, dim3(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096), dim3(256), 0, s,
                     theta, mask, count);
}

template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NV], double* __restrict__ out) {
  __shared__ double red[kOptThreads / 64][NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double x = acc[v];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][v] = x;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < NV; v += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < kOptThreads / 64; ++w) s += red[w][v];
    out[v] = s;
  }
}

// L-BFGS pair update.  For tile b: s_new = x - xp, y_new = g - gp written to
// slot new_slot of the history; dot products (FP64) per block:
//   [4j+0] s_j . y_new   [4j+1] y_j . y_new   [4j+2] s_j . g   [4j+3] y_j . g
// for the active slots j < m (slot index slots[b*kMaxHist + j]) and the new pair
// at position j = m (slot new_slot).
__global__ void __launch_bounds__(kOptThreads)
k_lbfgs_pair(const float* __restrict__ x, const float* __restrict__ xp, const float* __restrict__ g,
             const float* __restrict__ gp, float* __restrict__ hs, float* __restrict__ hy,
             const int* __restrict__ meta, size_t n, double* __restrict__ partials) {
  const int b = blockIdx.y;
  const int m = meta[b * kMetaStride + 0];
  if (m < 0) return;
  const int new_slot = meta[b * kMetaStride + 1];
  const int* slots = meta + b * kMetaStride + 2;
  const size_t tb = (size_t)b * n;
  const size_t hist = (size_t)(kMaxHist + 1) * n;  // per-tile history stride
  float* hs_b = hs + (size_t)b * hist;
  float* hy_b = hy + (size_t)b * hist;
  double acc[kPairDots];
#pragma unroll
  for (int v = 0; v < kPairDots; ++v) acc[v] = 0.0;
  const float* sp[kMaxHist];
  const float* yp[kMaxHist];
#pragma unroll
  for (int j = 0; j < kMaxHist; ++j) {
    const int sl = j < m ? slots[j] : 0;
    sp[j] = hs_b + (size_t)sl * n;
    yp[j] = hy_b + (size_t)sl * n;
  }
  float* sn = hs_b + (size_t)new_slot * n;
  float* yn = hy_b + (size_t)new_slot * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float s = x[tb + i] - xp[tb + i];
    const float y = g[tb + i] - gp[tb + i];
    const float gv = g[tb + i];
    sn[i] = s;
    yn[i] = y;
#pragma unroll
    for (int j = 0; j < kMaxHist; ++j) {
      if (j < m) {
        const double sj = sp[j][i], yj = yp[j][i];
        acc[4 * j + 0] += sj * (double)y;
        acc[4 * j + 1] += yj * (double)y;
        acc[4 * j + 2] += sj * (double)gv;
        acc[4 * j + 3] += yj * (double)gv;
      }
    }
    acc[4 * kMaxHist + 0] += (double)s * (double)y;
    acc[4 * kMaxHist + 1] += (double)y * (double)y;
    acc[4 * kMaxHist + 2] += (double)s * (double)gv;
    acc[4 * kMaxHist + 3] += (double)y * (double)gv;
  }
  block_reduce_store<kPairDots>(acc, partials + ((size_t)b * gridDim.x + blockIdx.x) * kPairDots);
}

void launch_lbfgs_pair(const float* x, const float* xp, const float* g, const float* gp, float* hs, float* hy,
                       const int* meta, size_t n, int batch, double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_lbfgs_pair, dim3(kOptBlocks, batch), dim3(kOptThreads), 0, s, x, xp, g, gp, hs, hy, meta, n,
                     partials);
}

// d = cg * g + sum_j cs_j s_j + sum_j cy_j y_j over the m active slots, plus the
// FP64 partial of g . d.  coef layout per tile: [0] cg, [1+j] cs_j, [1+kMaxHist+j] cy_j.
__global__ void __launch_bounds__(kOptThreads)
k_lbfgs_direction(const float* __restrict__ g, const float* __restrict__ hs, const float* __restrict__ hy,
                  const int* __restrict__ meta, const double* __restrict__ coef, size_t n, float* __restrict__ d,
                  double* __restrict__ partials) {
  const int b = blockIdx.y;
  const int m = meta[b * kMetaStride + 0];
  if (m < 0) return;
  const int* slots = meta + b * kMetaStride + 2;
  const double* cb = coef + (size_t)b * kCoefStride;
  const size_t tb = (size_t)b * n;
  const size_t hist = (size_t)(kMaxHist + 1) * n;
  const float* hs_b = hs + (size_t)b * hist;
  const float* hy_b = hy + (size_t)b * hist;
  const float cg = (float)cb[0];
  float cs[kMaxHist], cy[kMaxHist];
  const float* sp[kMaxHist];
  const float* yp[kMaxHist];
#pragma unroll
  for (int j = 0; j < kMaxHist; ++j) {
    const int sl = j < m ? slots[j] : 0;
    cs[j] = j < m ? (float)cb[1 + j] : 0.f;
    cy[j] = j < m ? (float)cb[1 + kMaxHist + j] : 0.f;
    sp[j] = hs_b + (size_t)sl * n;
    yp[j] = hy_b + (size_t)sl * n;
  }
  double acc[1] = {0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gv = g[tb + i];
    float v = cg * gv;
#pragma unroll
    for (int j = 0; j < kMaxHist; ++j)
      if (j < m) v = fmaf(cs[j], sp[j][i], fmaf(cy[j], yp[j][i], v));
    d[tb + i] = v;
    acc[0] += (double)gv * (double)v;
  }
  block_reduce_store<1>(acc, partials + ((size_t)b * gridDim.x + blockIdx.x));
}

void launch_lbfgs_direction(const float* g, const float* hs, const float* hy, const int* meta, const double* coef,
                            size_t n, float* d, int batch, double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_lbfgs_direction, dim3(kOptBlocks, batch), dim3(kOptThreads), 0, s, g, hs, hy, meta, coef, n,
                     d, partials);
}

// out[b*nv + v] = sum_blk partials[(b*nblk + blk)*nv + v], fixed order
__global__ void k_reduce_vec(const double* __restrict__ partials, int nblk, int nv, double* __restrict__ out) {
  const int b = blockIdx.x;
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nblk; ++i) s += partials[((size_t)b * nblk + i) * nv + v];
    out[(size_t)b * nv + v] = s;
  }
}

void launch_reduce_vec(const double* partials, int nblk, int nv, double* out, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_vec, dim3(batch), dim3(64), 0, s, partials, nblk, nv, out);
}

}  // namespace litho
