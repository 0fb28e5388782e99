// Device kernels of the CTM optimizers (gradient descent and L-BFGS) and small
// elementwise helpers.  All vectors are FP32; every dot product accumulates in
// FP64 and is reduced in a fixed order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include "opt_kernels.hpp"

#include <utility>

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
  __shared__ double red[kOptThreads / 64][NV > 0 ? NV : 1];
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

__device__ __forceinline__ double dot4(float4 a, float4 b) {
  return fma((double)a.x, (double)b.x, fma((double)a.y, (double)b.y, fma((double)a.z, (double)b.z, (double)a.w * (double)b.w)));
}
__device__ __forceinline__ float4 sub4(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
// per-component select (a float4 ternary can become a private array indexed by the condition)
__device__ __forceinline__ float4 sel4(bool c, float4 a, float4 b) {
  return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}
// the same on single floats: the vector kernels' element type is float4 when
// the tile's element count is a multiple of 4 (every even N), float otherwise
__device__ __forceinline__ double dot4(float a, float b) { return (double)a * (double)b; }
__device__ __forceinline__ float sub4(float a, float b) { return a - b; }
__device__ __forceinline__ float sel4(bool c, float a, float b) { return c ? a : b; }
__device__ __forceinline__ float4 scale4(float c, float4 g) { return make_float4(c * g.x, c * g.y, c * g.z, c * g.w); }
__device__ __forceinline__ float scale4(float c, float g) { return c * g; }
// v += cs s + cy y per component, fmaf(cs, s, fmaf(cy, y, v))
__device__ __forceinline__ void comb4(float4& v, float cs, float4 s, float cy, float4 y) {
  v.x = fmaf(cs, s.x, fmaf(cy, y.x, v.x));
  v.y = fmaf(cs, s.y, fmaf(cy, y.y, v.y));
  v.z = fmaf(cs, s.z, fmaf(cy, y.z, v.z));
  v.w = fmaf(cs, s.w, fmaf(cy, y.w, v.w));
}
__device__ __forceinline__ void comb4(float& v, float cs, float s, float cy, float y) { v = fmaf(cs, s, fmaf(cy, y, v)); }
template <class V>
constexpr size_t vec_width() { return sizeof(V) / sizeof(float); }

// Whether tile b's held pairs form one chain (start of pair j = end of pair
// j - 1): then each point is loaded once.  A rejected pair between two held
// ones breaks the chain; those tiles load both points of every pair (same
// values, same arithmetic).
__device__ __forceinline__ bool pairs_chained(const int* mt, int m) {
  bool chain = true;
  for (int j = 1; j < m; ++j) chain = chain && mt[kMetaStart + j] == mt[kMetaEnd + j - 1];
  return chain;
}
// Whether the chain of held pairs also ends at the base (m = 0 counts: the
// chain is the base alone).  Then the last chain point is the base, and the
// chained sweeps take x_p and g_p from it instead of loading them again.  A
// rejected newest pair (the base moved without a pair) breaks this.
__device__ __forceinline__ bool chain_ends_at_base(const int* mt, int m) {
  return pairs_chained(mt, m) && (m == 0 || mt[kMetaEnd + m - 1] == mt[kMetaBase]);
}
// Point p of a tile's slots (p < 0: the base, a valid stand-in for the pairs
// beyond m, whose loads are issued but not used).
template <class V>
__device__ __forceinline__ const V* point_ptr(const V* slots, const int* mt, int p, size_t n4) {
  return slots + (size_t)(p < 0 ? mt[kMetaBase] : p) * n4;
}
// Non-temporal loads (the nt bit of global_load).
__device__ __forceinline__ float ld_nt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float4 ld_nt(const float4* p) {
  typedef float nf4 __attribute__((ext_vector_type(4)));
  const nf4 v = __builtin_nontemporal_load(reinterpret_cast<const nf4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
// Load at a byte offset from a wave-uniform base: the offset stays 32 bits (a
// point holds at most kGenericMaxN^2 floats), so every load of a sweep shares
// one offset register and takes its base from scalar registers.
template <bool NT = false, class V>
__device__ __forceinline__ V ldv(const V* base, unsigned byte_off) {
  const V* p = reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + byte_off);
  if constexpr (NT)
    return ld_nt(p);
  else
    return *p;
}
// The pair and direction kernels load the history points non-temporally (the
// nt bit of global_load): each point is read once per kernel, and over the
// 24 streams of 16.8 MB of a 2048^2 tile the nt loads read 5.5-5.8 TB/s
// against 4.9-5.1 TB/s for plain loads.  A 2048^2 L-BFGS solve runs 4-5 %
// faster with nt loads in both kernels; with them in the pair kernel only,
// it was no faster.
constexpr bool kNtPair = true, kNtDir = true;

// L-BFGS pair dots.  For tile b the new pair is s = x - xp, y = g - gp with x,
// g the trial point and xp, gp the base; the held pairs are s_j = X[end_j] -
// X[start_j], y_j = G[end_j] - G[start_j] (FP32 differences of the stored
// points, the values the history would hold).  FP64 dot products per block,
// NV = 4 (MB + 1):
//   [4j+0] s_j . y_new   [4j+1] y_j . y_new   [4j+2] s_j . g   [4j+3] y_j . g
// for j < m <= MB, then the new pair at position 4 MB.  Vectors are read as V:
// float4 when n is a multiple of 4 (N^2 with N even), float otherwise.
//
// Element groups of 32 (group q of block k: elements (4k + q) * 32 + l, then
// strided over the grid) are taken by two waves: waves 0 and 1 the s side
// (points of X) of groups 0-1 and 2-3, waves 2 and 3 the y side (points of
// G) of the same groups, lanes 0-31 the first group of a wave and lanes
// 32-63 the second.  So each lane keeps 2 (MB + 1) FP64 accumulators and the
// point bases of a wave are uniform.  In the chained sweep (the chain ends
// at the base) all the point loads of an element are issued before its
// arithmetic, pairs beyond m load the base and are not accumulated, and the
// last chain point is the base: the s side loads x, g, g_p and the X chain,
// the y side g and the G chain, 26 loads per element group for its 24
// points.  Otherwise each held pair's points are loaded under its own test,
// and both sides load x, x_p, g, g_p.  Nothing is written: the trial point
// becomes the base by its index alone.
template <int MB, class V, int SIDE>
__device__ __forceinline__ void pair_sweep_chain(const V* __restrict__ hv, const V* __restrict__ x4,
                                                 const V* __restrict__ g4, const V* __restrict__ gp4, const int* mt,
                                                 int m, unsigned bo0, unsigned bstride, unsigned bend, size_t n4,
                                                 double (&acc)[2 * (MB + 1)]) {
  // chain points c_0 = start of pair 0, c_{j+1} = end of pair j, c_MB = the base
  const V* pc[MB + 1];
  pc[0] = point_ptr(hv, mt, m > 0 ? mt[kMetaStart] : -1, n4);
#pragma unroll
  for (int j = 0; j < MB; ++j) pc[j + 1] = point_ptr(hv, mt, j < m ? mt[kMetaEnd + j] : -1, n4);
  for (unsigned bo = bo0; bo < bend; bo += bstride) {
    V xv{}, gpv{};
    if constexpr (SIDE == 0) {
      xv = ldv<kNtPair>(x4, bo);
      gpv = ldv<kNtPair>(gp4, bo);
    }
    const V gv = ldv<kNtPair>(g4, bo);
    V cv[MB + 1];
#pragma unroll
    for (int j = 0; j <= MB; ++j) cv[j] = ldv<kNtPair>(pc[j], bo);
    // y = g - g_p, s = x - x_p, x_p (g_p) being the last X (G) chain point
    const V y = SIDE == 0 ? sub4(gv, gpv) : sub4(gv, cv[MB]);
    const V w = SIDE == 0 ? sub4(xv, cv[MB]) : y;
#pragma unroll
    for (int j = 0; j < MB; ++j) {
      if (j < m) {
        const V vj = sub4(cv[j + 1], cv[j]);
        acc[2 * j + 0] += dot4(vj, y);
        acc[2 * j + 1] += dot4(vj, gv);
      }
    }
    acc[2 * MB + 0] += dot4(w, y);
    acc[2 * MB + 1] += dot4(w, gv);
  }
}

template <int MB, class V>
__device__ __forceinline__ void pair_sweep_broken(const V* __restrict__ hv, const V* __restrict__ x4,
                                                  const V* __restrict__ xp4, const V* __restrict__ g4,
                                                  const V* __restrict__ gp4, const int* mt, int m, int h,
                                                  unsigned bo0, unsigned bstride, unsigned bend, size_t n4,
                                                  double (&acc)[2 * (MB + 1)]) {
  for (unsigned bo = bo0; bo < bend; bo += bstride) {
    const V xv = ldv<kNtPair>(x4, bo), xpv = ldv<kNtPair>(xp4, bo), gv = ldv<kNtPair>(g4, bo),
            gpv = ldv<kNtPair>(gp4, bo);
    const V y = sub4(gv, gpv);
    const V w = h ? y : sub4(xv, xpv);
#pragma unroll
    for (int j = 0; j < MB; ++j) {
      if (j < m) {
        const V vj = sub4(ldv<kNtPair>(point_ptr(hv, mt, mt[kMetaEnd + j], n4), bo),
                          ldv<kNtPair>(point_ptr(hv, mt, mt[kMetaStart + j], n4), bo));
        acc[2 * j + 0] += dot4(vj, y);
        acc[2 * j + 1] += dot4(vj, gv);
      }
    }
    acc[2 * MB + 0] += dot4(w, y);
    acc[2 * MB + 1] += dot4(w, gv);
  }
}

template <int MB, class V>
__global__ void __launch_bounds__(kOptThreads) __attribute__((amdgpu_waves_per_eu(3)))
k_lbfgs_pair(const float* __restrict__ X, const float* __restrict__ G, int points, const int* __restrict__ need,
             const int* __restrict__ meta, size_t n, double* __restrict__ partials) {
  constexpr int NV = 4 * (MB + 1);
  constexpr int NA = 2 * (MB + 1);
  constexpr int WAVES = kOptThreads / 64;
  static_assert(WAVES == 4, "two element groups per wave, two waves per side");
  const int b = blockIdx.y;
  if (need[b] == 0) return;
  const int* mt = meta + b * kMetaStride;
  const int m = mt[0];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = wave >> 1;                                   // 0: s side, 1: y side
  const int grp = 2 * (wave & 1) + (lane >> 5);              // element group of this lane
  constexpr size_t VW = vec_width<V>();
  const size_t n4 = n / VW;
  const V* X4 = reinterpret_cast<const V*>(X + (size_t)b * points * n);
  const V* G4 = reinterpret_cast<const V*>(G + (size_t)b * points * n);
  const size_t trial = (size_t)mt[kMetaTrial] * n4, base = (size_t)mt[kMetaBase] * n4;
  double acc[NA];
#pragma unroll
  for (int v = 0; v < NA; ++v) acc[v] = 0.0;
  const unsigned bo0 = (unsigned)((((size_t)blockIdx.x * WAVES + grp) * 32 + (lane & 31)) * sizeof(V));
  const unsigned bstride = (unsigned)((size_t)gridDim.x * WAVES * 32 * sizeof(V));
  const unsigned bend = (unsigned)(n4 * sizeof(V));
  const V* hv = h ? G4 : X4;                                 // wave-uniform
  if (!chain_ends_at_base(mt, m))
    pair_sweep_broken<MB, V>(hv, X4 + trial, X4 + base, G4 + trial, G4 + base, mt, m, h, bo0, bstride, bend, n4,
                             acc);
  else if (h == 0)
    pair_sweep_chain<MB, V, 0>(hv, X4 + trial, G4 + trial, G4 + base, mt, m, bo0, bstride, bend, n4, acc);
  else
    pair_sweep_chain<MB, V, 1>(hv, X4 + trial, G4 + trial, G4 + base, mt, m, bo0, bstride, bend, n4, acc);
  // fixed-order reduction: each group over its 32 lanes, then the groups in order
  __shared__ double red[WAVES][2][NA];
#pragma unroll
  for (int v = 0; v < NA; ++v) {
    double t = acc[v];
    for (int o = 16; o > 0; o >>= 1) t += __shfl_down(t, o, 32);
    if ((lane & 31) == 0) red[grp][h][v] = t;
  }
  __syncthreads();
  double* out = partials + ((size_t)b * gridDim.x + blockIdx.x) * NV;
  for (int v = threadIdx.x; v < NV; v += blockDim.x) {
    // v = 4 j + 2 t + hv: side hv, accumulator 2 j + t (j = MB: the new pair)
    const int jj = v >> 2, t = (v >> 1) & 1, hh = v & 1;
    double sum = 0.0;
    for (int w2 = 0; w2 < WAVES; ++w2) sum += red[w2][hh][2 * jj + t];
    out[v] = sum;
  }
}

// d = cg * g + sum_j cs_j s_j + sum_j cy_j y_j over the m <= MB held pairs,
// g = G[base], plus the FP64 partial of g . d.  coef layout per tile: [0] cg,
// [1+j] cs_j, [1+kMaxHist+j] cy_j.  In the chained sweep (the chain ends at
// the base) all the point loads of an element are issued before its
// arithmetic (pairs beyond m load the base, unused), as in the pair kernel,
// and g is the last G chain point.  The sweep runs from the highest
// addresses down: the pair kernel has just read
// the same points upwards, so the lines it read last, still in the
// last-level cache, are read first.
template <int MB, class V, bool CHAIN>
__device__ __forceinline__ double direction_sweep(const V* __restrict__ X4, const V* __restrict__ G4,
                                                  const V* __restrict__ gb, const int* mt, int m, float cg,
                                                  const float (&cs)[MB > 0 ? MB : 1],
                                                  const float (&cy)[MB > 0 ? MB : 1], V* __restrict__ d4,
                                                  unsigned bo0, unsigned bstride, unsigned bend, size_t n4) {
  double acc = 0.0;
  if constexpr (CHAIN) {
    // chain points c_0 = start of pair 0, c_{j+1} = end of pair j
    const V* cx[MB + 1];
    const V* cgp[MB + 1];
    const int p0 = m > 0 ? mt[kMetaStart] : -1;
    cx[0] = point_ptr(X4, mt, p0, n4);
    cgp[0] = point_ptr(G4, mt, p0, n4);
#pragma unroll
    for (int j = 0; j < MB; ++j) {
      const int p = j < m ? mt[kMetaEnd + j] : -1;
      cx[j + 1] = point_ptr(X4, mt, p, n4);
      cgp[j + 1] = point_ptr(G4, mt, p, n4);
    }
    for (unsigned kk = bo0 < bend ? (bend - bo0 + bstride - 1) / bstride : 0; kk-- > 0;) {
      const unsigned bo = bo0 + kk * bstride;
      V xv[MB + 1], gvv[MB + 1];
#pragma unroll
      for (int j = 0; j <= MB; ++j) {
        xv[j] = ldv<kNtDir>(cx[j], bo);
        gvv[j] = ldv<kNtDir>(cgp[j], bo);
      }
      const V gv = gvv[MB];   // the base gradient, the last chain point
      V v = scale4(cg, gv);
#pragma unroll
      for (int j = 0; j < MB; ++j)
        if (j < m) comb4(v, cs[j], sub4(xv[j + 1], xv[j]), cy[j], sub4(gvv[j + 1], gvv[j]));
      *reinterpret_cast<V*>(reinterpret_cast<char*>(d4) + bo) = v;
      acc += dot4(gv, v);
    }
  } else {
    for (unsigned kk = bo0 < bend ? (bend - bo0 + bstride - 1) / bstride : 0; kk-- > 0;) {
      const unsigned bo = bo0 + kk * bstride;
      const V gv = ldv<kNtDir>(gb, bo);
      V v = scale4(cg, gv);
#pragma unroll
      for (int j = 0; j < MB; ++j) {
        if (j < m) {
          const int sp = mt[kMetaStart + j], ep = mt[kMetaEnd + j];
          const V sj = sub4(ldv<kNtDir>(point_ptr(X4, mt, ep, n4), bo), ldv<kNtDir>(point_ptr(X4, mt, sp, n4), bo));
          const V yj = sub4(ldv<kNtDir>(point_ptr(G4, mt, ep, n4), bo), ldv<kNtDir>(point_ptr(G4, mt, sp, n4), bo));
          comb4(v, cs[j], sj, cy[j], yj);
        }
      }
      *reinterpret_cast<V*>(reinterpret_cast<char*>(d4) + bo) = v;
      acc += dot4(gv, v);
    }
  }
  return acc;
}

template <int MB, class V>
__global__ void __launch_bounds__(kOptThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_lbfgs_direction(const float* __restrict__ X, const float* __restrict__ G, int points, const int* __restrict__ need,
                  const int* __restrict__ meta, const double* __restrict__ coef, size_t n, float* __restrict__ d,
                  double* __restrict__ partials) {
  const int b = blockIdx.y;
  if (need[b] == 0) return;
  const int* mt = meta + b * kMetaStride;
  const int m = mt[0];
  const double* cb = coef + (size_t)b * kCoefStride;
  const size_t n4 = n / vec_width<V>();
  const V* X4 = reinterpret_cast<const V*>(X + (size_t)b * points * n);
  const V* G4 = reinterpret_cast<const V*>(G + (size_t)b * points * n);
  const float cg = (float)cb[0];
  float cs[MB > 0 ? MB : 1], cy[MB > 0 ? MB : 1];
#pragma unroll
  for (int j = 0; j < MB; ++j) {
    cs[j] = j < m ? (float)cb[1 + j] : 0.f;
    cy[j] = j < m ? (float)cb[1 + kMaxHist + j] : 0.f;
  }
  const V* gb = G4 + (size_t)mt[kMetaBase] * n4;
  V* d4 = reinterpret_cast<V*>(d + (size_t)b * n);
  const unsigned bo0 = (unsigned)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) * sizeof(V));
  const unsigned bstride = (unsigned)((size_t)gridDim.x * blockDim.x * sizeof(V));
  const unsigned bend = (unsigned)(n4 * sizeof(V));
  double acc[1];
  acc[0] = chain_ends_at_base(mt, m)
               ? direction_sweep<MB, V, true>(X4, G4, gb, mt, m, cg, cs, cy, d4, bo0, bstride, bend, n4)
               : direction_sweep<MB, V, false>(X4, G4, gb, mt, m, cg, cs, cy, d4, bo0, bstride, bend, n4);
  block_reduce_store<1>(acc, partials + ((size_t)b * gridDim.x + blockIdx.x));
}

// One 64-thread block per tile.  Restates LBFGS++'s BFGSMat::add_correction
// (accept when s.y > eps y.y; the oldest pair leaves once m_max are held) on
// the compact representation, then the Byrd-Nocedal-Schnabel direction
//   d = -gamma g - S u + gamma Y w,  w = R^{-1} p,
//   u = R^{-T}((D + gamma Y'Y) w - gamma q),  gamma = s.y / y.y of the newest pair,
// with R the upper triangle of S'Y, D its diagonal, p = S'g, q = Y'g.  The
// matrices are staged in LDS; the two triangular solves run on one lane with
// register arrays (loops unrolled over kMaxHist).
__global__ void __launch_bounds__(64)
k_lbfgs_update(const int* __restrict__ need, const double* __restrict__ dots, int mb, int m_max,
               int* __restrict__ meta, double* __restrict__ sy, double* __restrict__ yy, double* __restrict__ coef) {
  constexpr int H = kMaxHist;
  const int b = blockIdx.x;
  if (need[b] == 0) return;
  __shared__ double s0[H * H], y0[H * H], s1[H * H], y1[H * H];
  __shared__ double v[kPairDots];
  __shared__ int mt[kMetaStride];
  const int tid = threadIdx.x;
  double* sy_b = sy + (size_t)b * H * H;
  double* yy_b = yy + (size_t)b * H * H;
  for (int e = tid; e < H * H; e += 64) {
    s0[e] = sy_b[e];
    y0[e] = yy_b[e];
  }
  for (int e = tid; e < 4 * (mb + 1); e += 64) v[e] = dots[(size_t)b * kPairDots + e];
  if (tid < kMetaStride) mt[tid] = meta[b * kMetaStride + tid];
  __syncthreads();
  const int m = mt[0];
  const double ys = v[4 * mb + 0], yy_new = v[4 * mb + 1];
  const bool accept = ys > 2.220446049250313e-16 * yy_new;
  const int drop = accept && m == m_max ? 1 : 0;
  const int m2 = accept ? m - drop + 1 : m;
  // the updated S'Y (upper triangle) and Y'Y, chronological order
  for (int e = tid; e < H * H; e += 64) {
    const int i = e / H, j = e - i * H;
    double a = s0[e], c = y0[e];
    if (accept) {
      a = 0.0;
      c = 0.0;
      if (i < m2 - 1 && j < m2 - 1) {
        a = s0[(i + drop) * H + j + drop];
        c = y0[(i + drop) * H + j + drop];
      } else if (i < m2 - 1 && j == m2 - 1) {
        a = v[4 * (i + drop) + 0];
        c = v[4 * (i + drop) + 1];
      } else if (i == m2 - 1 && j < m2 - 1) {
        c = v[4 * (j + drop) + 1];
      } else if (i == m2 - 1 && j == m2 - 1) {
        a = ys;
        c = yy_new;
      }
    }
    s1[e] = a;
    y1[e] = c;
    sy_b[e] = a;
    yy_b[e] = c;
  }
  __syncthreads();
  if (tid == 0) {
    int* mt_g = meta + b * kMetaStride;
    // the held pairs after the update (chronological), the trial point as the
    // new base, and the next trial: the smallest point that neither a held
    // pair nor the base uses (at most 2 m_max + 1 points are in use)
    const int base = mt[kMetaBase], trial = mt[kMetaTrial];
    int st2[H], en2[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
      int a0 = -1, e0 = -1;
      if (accept) {
        if (i < m2 - 1) {
          a0 = mt[kMetaStart + i + drop];
          e0 = mt[kMetaEnd + i + drop];
        } else if (i == m2 - 1) {
          a0 = base;
          e0 = trial;
        }
      } else if (i < m2) {
        a0 = mt[kMetaStart + i];
        e0 = mt[kMetaEnd + i];
      }
      st2[i] = a0;
      en2[i] = e0;
    }
    int next = 0;
    for (bool used = true; used;) {
      used = next == trial;
#pragma unroll
      for (int i = 0; i < H; ++i) used = used || st2[i] == next || en2[i] == next;
      if (used) ++next;
    }
    mt_g[0] = m2;
    mt_g[kMetaBase] = trial;
    mt_g[kMetaTrial] = next;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      mt_g[kMetaStart + i] = st2[i];
      mt_g[kMetaEnd + i] = en2[i];
    }
  }
  // The compact-form coefficients, lane l for index l: the triangular solves
  // run over the lanes in the order of the serial recurrences (each lane
  // holds its row of S'Y and Y'Y and its column of S'Y, every w_i and u_i is
  // broadcast as soon as it is known), so every sum adds the same terms in
  // the same order as a one-lane solve.
  double* cf = coef + (size_t)b * kCoefStride;
  const int l = tid;
  const bool lane = l < H;
  if (m2 == 0) {
    if (l == 0) cf[0] = -1.0;
    if (lane) {
      cf[1 + l] = 0.0;
      cf[1 + kMaxHist + l] = 0.0;
    }
    return;
  }
  double srow[H], scol[H], yrow[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    srow[j] = lane ? s1[l * H + j] : 1.0;
    scol[j] = lane ? s1[j * H + l] : 0.0;
    yrow[j] = lane ? y1[l * H + j] : 0.0;
  }
  const double diag = lane ? s1[l * H + l] : 1.0;
  const int src = accept ? (l < m2 - 1 ? l + drop : mb) : l;
  const double pl = lane && l < m2 ? v[4 * src + 2] : 0.0;
  const double ql = lane && l < m2 ? v[4 * src + 3] : 0.0;
  const double gamma = s1[(m2 - 1) * H + m2 - 1] / y1[(m2 - 1) * H + m2 - 1];
  // w = R^{-1} p, i descending: lane i forms w_i
  double w[H], wl = 0.0;
#pragma unroll
  for (int i = H - 1; i >= 0; --i) {
    w[i] = 0.0;
    if (i < m2) {
      double acc = pl;
#pragma unroll
      for (int j = i + 1; j < H; ++j)
        if (j < m2) acc -= srow[j] * w[j];
      const double wi = __shfl(acc / srow[i], i, 64);
      w[i] = wi;
      wl = l == i ? wi : wl;
    }
  }
  // r_l = (D + gamma Y'Y) w - gamma q at row l
  double rl = 0.0;
  if (lane && l < m2) {
    double acc = diag * wl - gamma * ql;
#pragma unroll
    for (int j = 0; j < H; ++j)
      if (j < m2) acc += gamma * yrow[j] * w[j];
    rl = acc;
  }
  // u = R^{-T} r, i ascending: lane i forms u_i
  double u[H], ul = 0.0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    u[i] = 0.0;
    if (i < m2) {
      double acc = rl;
#pragma unroll
      for (int j = 0; j < i; ++j) acc -= scol[j] * u[j];
      const double ui = __shfl(acc / diag, i, 64);
      u[i] = ui;
      ul = l == i ? ui : ul;
    }
  }
  if (l == 0) cf[0] = -gamma;
  if (lane) {
    cf[1 + l] = l < m2 ? -ul : 0.0;
    cf[1 + kMaxHist + l] = l < m2 ? gamma * wl : 0.0;
  }
}

void launch_lbfgs_update(int batch, const int* need, const double* dots, int mb, int m_max, int* meta, double* sy,
                         double* yy, double* coef, hipStream_t s) {
  hipLaunchKernelGGL(k_lbfgs_update, dim3(batch), dim3(64), 0, s, need, dots, mb, m_max, meta, sy, yy, coef);
}

namespace {

using PairLaunch = void (*)(dim3, hipStream_t, const float*, const float*, int, const int*, const int*, size_t,
                            double*);
using DirLaunch = void (*)(dim3, hipStream_t, const float*, const float*, int, const int*, const int*, const double*,
                           size_t, float*, double*);

template <int MB>
void pair_launch(dim3 grid, hipStream_t s, const float* X, const float* G, int points, const int* need,
                 const int* meta, size_t n, double* partials) {
  if (n % 4 == 0)
    hipLaunchKernelGGL((k_lbfgs_pair<MB, float4>), grid, dim3(kOptThreads), 0, s, X, G, points, need, meta, n,
                       partials);
  else   // odd N: per-float loads (point bases are not 16-byte aligned)
    hipLaunchKernelGGL((k_lbfgs_pair<MB, float>), grid, dim3(kOptThreads), 0, s, X, G, points, need, meta, n,
                       partials);
}
template <int MB>
void dir_launch(dim3 grid, hipStream_t s, const float* X, const float* G, int points, const int* need,
                const int* meta, const double* coef, size_t n, float* d, double* partials) {
  if (n % 4 == 0)
    hipLaunchKernelGGL((k_lbfgs_direction<MB, float4>), grid, dim3(kOptThreads), 0, s, X, G, points, need, meta,
                       coef, n, d, partials);
  else
    hipLaunchKernelGGL((k_lbfgs_direction<MB, float>), grid, dim3(kOptThreads), 0, s, X, G, points, need, meta,
                       coef, n, d, partials);
}
template <int... I>
constexpr PairLaunch pair_table_entry(int m, std::integer_sequence<int, I...>) {
  constexpr PairLaunch t[] = {&pair_launch<I>...};
  return t[m];
}
template <int... I>
constexpr DirLaunch dir_table_entry(int m, std::integer_sequence<int, I...>) {
  constexpr DirLaunch t[] = {&dir_launch<I>...};
  return t[m];
}

}  // namespace

void launch_lbfgs_pair(const float* X, const float* G, int points, const int* need, const int* meta, size_t n,
                       int batch, int mb, double* partials, hipStream_t s) {
  pair_table_entry(mb, std::make_integer_sequence<int, kMaxHist + 1>{})(dim3(kOptBlocks, batch), s, X, G, points,
                                                                        need, meta, n, partials);
}

void launch_lbfgs_direction(const float* X, const float* G, int points, const int* need, const int* meta,
                            const double* coef, size_t n, float* d, int batch, int mb, double* partials,
                            hipStream_t s) {
  dir_table_entry(mb, std::make_integer_sequence<int, kMaxHist + 1>{})(dim3(kOptBlocks, batch), s, X, G, points,
                                                                       need, meta, coef, n, d, partials);
}

__global__ void k_lbfgs_reset(int* __restrict__ meta, int b) {
  if (threadIdx.x != 0) return;
  int* mt = meta + b * kMetaStride;
  mt[0] = 0;
  mt[kMetaBase] = 0;
  mt[kMetaTrial] = 1;
}

void launch_lbfgs_reset(int* meta, int b, hipStream_t s) {
  hipLaunchKernelGGL(k_lbfgs_reset, dim3(1), dim3(64), 0, s, meta, b);
}

__global__ void k_lbfgs_start_meta(const int* __restrict__ flags, int* __restrict__ meta, int batch) {
  const int b = threadIdx.x;
  if (b >= batch || flags[b] == 0) return;
  int* mt = meta + b * kMetaStride;
  const int base = mt[kMetaBase];
  mt[kMetaBase] = mt[kMetaTrial];
  mt[kMetaTrial] = base;
}

// d = -G[base] (the first direction, -1 * g as CTM_Optimizer's steepest start)
__global__ void k_lbfgs_steepest(const float* __restrict__ G, int points, const int* __restrict__ flags,
                                 const int* __restrict__ meta, size_t n, float* __restrict__ d) {
  const int b = blockIdx.y;
  if (flags[b] == 0) return;
  const float* gb = G + ((size_t)b * points + meta[b * kMetaStride + kMetaBase]) * n;
  float* db = d + (size_t)b * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    db[i] = -1.0f * gb[i];
}

void launch_lbfgs_start(const float* G, int points, const int* flags, int* meta, size_t n, int batch, float* d,
                        hipStream_t s) {
  const int nt = batch <= 64 ? 64 : (batch + 63) / 64 * 64;
  hipLaunchKernelGGL(k_lbfgs_start_meta, dim3(1), dim3(nt), 0, s, flags, meta, batch);
  hipLaunchKernelGGL(k_lbfgs_steepest, dim3(kOptBlocks, batch), dim3(kOptThreads), 0, s, G, points, flags,
                     (const int*)meta, n, d);
}

__global__ void k_lbfgs_point_out(const float* __restrict__ X, int points, const int* __restrict__ meta, int b,
                                  int which, size_t n, float* __restrict__ out_theta, float* __restrict__ out_mask) {
  const float* p = X + ((size_t)b * points + meta[b * kMetaStride + which]) * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float t = p[i];
    if (out_theta != nullptr) out_theta[i] = t;
    out_mask[i] = 0.5f * (1.0f + cosf(t));
  }
}

void launch_lbfgs_point_out(const float* X, int points, const int* meta, int b, int which, size_t n,
                            float* out_theta, float* out_mask, hipStream_t s) {
  const size_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_lbfgs_point_out, dim3(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096), dim3(256), 0, s, X,
                     points, meta, b, which, n, out_theta, out_mask);
}

// The yardstick of the L-BFGS vector kernels: a plain kernel that reads R and
// writes W vectors of n4 float4 with their grid (kOptBlocks x kOptThreads)
// and their non-temporal loads, every load of an element issued before the
// sum that uses it.
constexpr int kStreamMax = 32;
struct StreamPtrs {
  const float4* in[kStreamMax];
  float4* out[kStreamMax];
};
template <int R, int W>
__global__ void __launch_bounds__(kOptThreads) k_stream(StreamPtrs p, size_t n4, float* sink) {
  float keep = 0.f;
  for (size_t i = blockIdx.x * (size_t)kOptThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kOptThreads) {
    float4 v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = ld_nt(p.in[j] + i);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      acc.x += v[j].x;
      acc.y += v[j].y;
      acc.z += v[j].z;
      acc.w += v[j].w;
    }
#pragma unroll
    for (int j = 0; j < W; ++j) p.out[j][i] = acc;
    keep += acc.x;
  }
  if (keep == 1234.5f) sink[0] = keep;   // keeps the loads of a read-only pattern
}

// Plain loads of n4 float4 (one word of each): a read larger than the
// last-level cache leaves none of the timed pattern's vectors there, as in a
// solve, where an evaluation runs between the L-BFGS sweeps.
__global__ void __launch_bounds__(kOptThreads) k_cache_evict(const float4* __restrict__ buf, size_t n4, float* sink) {
  float keep = 0.f;
  for (size_t i = blockIdx.x * (size_t)kOptThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kOptThreads)
    keep += buf[i].x;
  if (keep == 1234.5f) sink[0] = keep;
}

void launch_cache_evict(const float* buf, size_t n4, float* sink, hipStream_t s) {
  hipLaunchKernelGGL(k_cache_evict, dim3(4096), dim3(kOptThreads), 0, s, reinterpret_cast<const float4*>(buf), n4,
                     sink);
}

bool launch_stream_pattern(int reads, int writes, const float* const* in, float* const* out, size_t n4, float* sink,
                           hipStream_t s) {
  if (reads < 0 || writes < 0 || reads + writes > kStreamMax) return false;
  StreamPtrs p{};
  for (int j = 0; j < reads; ++j) p.in[j] = reinterpret_cast<const float4*>(in[j]);
  for (int j = 0; j < writes; ++j) p.out[j] = reinterpret_cast<float4*>(out[j]);
  const dim3 grid(kOptBlocks), blk(kOptThreads);
  if (reads == 24 && writes == 0)
    hipLaunchKernelGGL((k_stream<24, 0>), grid, blk, 0, s, p, n4, sink);
  else if (reads == 22 && writes == 1)
    hipLaunchKernelGGL((k_stream<22, 1>), grid, blk, 0, s, p, n4, sink);
  else if (reads == 24 && writes == 4)
    hipLaunchKernelGGL((k_stream<24, 4>), grid, blk, 0, s, p, n4, sink);
  else if (reads == 21 && writes == 1)
    hipLaunchKernelGGL((k_stream<21, 1>), grid, blk, 0, s, p, n4, sink);
  else
    return false;
  return true;
}

// One block per (v, tile): thread t sums blocks t, t + 256, ... in order, then a
// fixed shuffle tree and a fixed sum over the four waves.
__global__ void __launch_bounds__(256)
k_reduce_fixed(const double* __restrict__ partials, int nblk, int nv, double* __restrict__ out, int out_stride,
               const int* __restrict__ active) {
  const int v = blockIdx.x, b = blockIdx.y;
  if (active != nullptr && active[b] == 0) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) s += partials[((size_t)b * nblk + i) * nv + v];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[(size_t)b * out_stride + v] = ((w[0] + w[1]) + w[2]) + w[3];
}

// The two reductions that end an evaluation in one launch: values v < 2 from
// the image partials (PE, penalty), v >= 2 from the gradient partials
// (|g|^2, g . d), each summed as k_reduce_fixed sums it.
// (v = 4, with extra set: one more value per tile from extra's partials
// ([b][extra_blk], one value each) into extra_out[b].)
__global__ void __launch_bounds__(256)
k_reduce_eval(const double* __restrict__ img, int img_blk, const double* __restrict__ grad, int grad_blk,
              double* __restrict__ out, const int* __restrict__ active, const double* __restrict__ extra,
              int extra_blk, double* __restrict__ extra_out) {
  const int v = blockIdx.x, b = blockIdx.y;
  if (active != nullptr && active[b] == 0) return;
  const double* part = v < 2 ? img : (v < 4 ? grad : extra);
  const int nblk = v < 2 ? img_blk : (v < 4 ? grad_blk : extra_blk);
  const int nv = v < 4 ? 2 : 1, vv = v < 4 ? (v & 1) : 0;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) s += part[((size_t)b * nblk + i) * nv + vv];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double r = ((w[0] + w[1]) + w[2]) + w[3];
    if (v < 4)
      out[(size_t)b * 4 + v] = r;
    else
      extra_out[b] = r;
  }
}

void launch_reduce_eval(const double* img, int img_blk, const double* grad, int grad_blk, double* out, int batch,
                        hipStream_t s, const int* active, const double* extra, int extra_blk, double* extra_out) {
  const int nv = extra != nullptr && extra_out != nullptr ? 5 : 4;
  hipLaunchKernelGGL(k_reduce_eval, dim3(nv, batch), dim3(256), 0, s, img, img_blk, grad, grad_blk, out, active, extra,
                     extra_blk, extra_out);
}

void launch_reduce_fixed(const double* partials, int nblk, int nv, double* out, int out_stride, int batch,
                         hipStream_t s, const int* active) {
  hipLaunchKernelGGL(k_reduce_fixed, dim3(nv, batch), dim3(256), 0, s, partials, nblk, nv, out, out_stride, active);
}

}  // namespace litho
