// Full-chip Micro tiling kernels for gfx950 (tiling.hpp).
//
// Restates source/utils/partition.cpp (extract_local_tile :83-95,
// assemble_global :105-113), the loss windows and the serial exchange phase of
// source/utils/micro_gradient_exchange.cpp (:65-93, :253-326) and the theta
// step of source/optimizer/full_chip_ctm.cpp:133-141.  The reference walks the
// neighbours of each Micro serially; here one launch covers every (Micro,
// neighbour) Guard strip at once: the strips of one Micro are disjoint from
// each other and from its Core, so every output pixel has a single writer.
// The Core-side CEPT terms of two neighbours can meet in a Core corner; they
// are added by one launch per neighbour slot, in the reference's order.
#include <hip/hip_runtime.h>

#include "tiling.hpp"

namespace litho {

std::vector<int> tile_table(const litho_tile* tiles, int num_tiles) {
  std::vector<int> t((size_t)num_tiles * kTileStride, 0);
  auto put = [](int* d, const litho_rect& r) {
    d[0] = r.y0;
    d[1] = r.y1;
    d[2] = r.x0;
    d[3] = r.x1;
  };
  for (int i = 0; i < num_tiles; ++i) {
    int* d = t.data() + (size_t)i * kTileStride;
    const litho_tile& m = tiles[i];
    put(d + 0, m.tile_global);
    put(d + 4, m.core_global);
    put(d + 8, m.core_local);
    put(d + 12, m.update_local);
    d[16] = m.num_neighbors;
    for (int j = 0; j < m.num_neighbors && j < 4; ++j) {
      int* e = d + 20 + 20 * j;
      const litho_neighbor& nb = m.neighbors[j];
      e[0] = 1;
      e[1] = nb.id;
      e[2] = nb.direction;
      put(e + 4, nb.my_guard_local);
      put(e + 8, nb.neighbor_overlap_local);
      put(e + 12, nb.my_overlap_local);
      put(e + 16, nb.neighbor_guard_local);
    }
  }
  return t;
}

namespace {
__device__ __forceinline__ int rows_of(const int* r) { return r[1] - r[0]; }
__device__ __forceinline__ int cols_of(const int* r) { return r[3] - r[2]; }
__device__ __forceinline__ int wrap(int g, int n) { return ((g % n) + n) % n; }
}  // namespace

// ---------------------------------------------------------------- partition
__global__ void __launch_bounds__(256)
k_extract(const float* __restrict__ global, int height, int width, const int* __restrict__ table, int n,
          float* __restrict__ local) {
  const int t = blockIdx.y;
  const int* tg = table + (size_t)t * kTileStride;
  const size_t plane = (size_t)n * n;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < plane; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / n), x = (int)(p - (size_t)y * n);
    local[t * plane + p] = global[(size_t)wrap(tg[0] + y, height) * width + wrap(tg[2] + x, width)];
  }
}

void launch_extract(const float* global, int height, int width, const int* table, int num_tiles, int n, float* local,
                    hipStream_t s) {
  const size_t plane = (size_t)n * n;
  const int blocks = (int)((plane + 255) / 256 < 1024 ? (plane + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_extract, dim3(blocks, num_tiles), dim3(256), 0, s, global, height, width, table, n, local);
}

__global__ void __launch_bounds__(256)
k_assemble(const float* __restrict__ local, const int* __restrict__ table, int n, int width,
           float* __restrict__ global) {
  const int t = blockIdx.y;
  const int* cg = table + (size_t)t * kTileStride + 4;
  const int* cl = table + (size_t)t * kTileStride + 8;
  const int w = cols_of(cg);
  const size_t cnt = (size_t)rows_of(cg) * w;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(q / w), x = (int)(q - (size_t)y * w);
    global[(size_t)(cg[0] + y) * width + cg[2] + x] = local[(size_t)t * n * n + (size_t)(cl[0] + y) * n + cl[2] + x];
  }
}

void launch_assemble(const float* local, const int* table, int num_tiles, int n, int width, float* global,
                     int max_core_elems, hipStream_t s) {
  const int blocks = (max_core_elems + 255) / 256 < 1024 ? (max_core_elems + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_assemble, dim3(blocks > 0 ? blocks : 1, num_tiles), dim3(256), 0, s, local, table, n, width,
                     global);
}

// ---------------------------------------------------------------- windows
__global__ void __launch_bounds__(256)
k_windows(const float* __restrict__ targets, const int* __restrict__ table, int n, int core_guard, int loss_type,
          float* __restrict__ pe_win, float* __restrict__ pen_win) {
  const int t = blockIdx.y;
  const int* r = table + (size_t)t * kTileStride + (core_guard ? 12 : 8);
  const size_t plane = (size_t)n * n;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < plane; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / n), x = (int)(p - (size_t)y * n);
    const bool in = y >= r[0] && y < r[1] && x >= r[2] && x < r[3];
    const float tv = targets[t * plane + p];
    pen_win[t * plane + p] = (in && tv < 0.5f) ? 1.f : 0.f;
    pe_win[t * plane + p] = in ? (loss_type == 1 ? (tv > 0.5f ? 1.f : 0.f) : 1.f) : 0.f;
  }
}

void launch_windows(const float* targets, const int* table, int num_tiles, int n, int region_core_guard,
                    int loss_type, float* pe_win, float* pen_win, hipStream_t s) {
  const size_t plane = (size_t)n * n;
  const int blocks = (int)((plane + 255) / 256 < 1024 ? (plane + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_windows, dim3(blocks, num_tiles), dim3(256), 0, s, targets, table, n, region_core_guard,
                     loss_type, pe_win, pen_win);
}

// ---------------------------------------------------------------- exchange
// linear_strength (micro_gradient_exchange.cpp:52-61)
__device__ __forceinline__ double linear_strength(int dir, int row, int col, int rows, int cols) {
  if (dir == 0) return cols == 1 ? 1.0 : 1.0 - (double)col / (cols - 1);
  if (dir == 1) return cols == 1 ? 1.0 : (double)col / (cols - 1);
  if (dir == 2) return rows == 1 ? 1.0 : 1.0 - (double)row / (rows - 1);
  return rows == 1 ? 1.0 : (double)row / (rows - 1);
}

__global__ void __launch_bounds__(256)
k_guard_exchange(const int* __restrict__ table, int n, const float* __restrict__ masks, const float* __restrict__ raw,
                 float* __restrict__ mg, ExchangeParams prm, double* __restrict__ partial) {
  const int t = blockIdx.z, slot = blockIdx.y;
  const int* e = table + (size_t)t * kTileStride + 20 + 20 * slot;
  double acc = 0.0;
  if (e[0] != 0) {
    const int nb = e[1], dir = e[2];
    const int* g = e + 4;
    const int* gs = e + 8;
    const int rows = rows_of(g), cols = cols_of(g);
    const size_t cnt = (rows > 0 && cols > 0) ? (size_t)rows * cols : 0;
    const size_t mine = (size_t)t * n * n, theirs = (size_t)nb * n * n;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
      const int y = (int)(q / cols), x = (int)(q - (size_t)y * cols);
      const size_t pm = mine + (size_t)(g[0] + y) * n + g[2] + x;
      const size_t pn = theirs + (size_t)(gs[0] + y) * n + gs[2] + x;
      double v = (double)raw[pm];
      if (prm.fusion) {
        const double s = prm.linear ? linear_strength(dir, y, x, rows, cols) : 1.0;
        const double nw = prm.w_neighbor * s;
        if (nw != 0.0) v = (prm.w_owner * v + nw * (double)raw[pn]) / (prm.w_owner + nw);
      }
      if (prm.cept) {
        const double diff = (double)masks[pm] - (double)masks[pn];
        acc += diff * diff;
        v += prm.cept_weight * (2.0 * diff);
      }
      mg[pm] = (float)v;
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[((size_t)t * 4 + slot) * gridDim.x + blockIdx.x] = ((w[0] + w[1]) + w[2]) + w[3];
}

void launch_guard_exchange(const int* table, int num_tiles, int n, const float* masks, const float* raw,
                           float* mask_gradient, const ExchangeParams& p, int max_strip, double* partial,
                           hipStream_t s) {
  (void)max_strip;
  hipLaunchKernelGGL(k_guard_exchange, dim3(kExchangeBlocks, 4, num_tiles), dim3(256), 0, s, table, n, masks, raw,
                     mask_gradient, p, partial);
}

__global__ void __launch_bounds__(256)
k_overlap_cept(const int* __restrict__ table, int n, int slot, const float* __restrict__ masks,
               float* __restrict__ mg, double cept_weight) {
  const int t = blockIdx.y;
  const int* e = table + (size_t)t * kTileStride + 20 + 20 * slot;
  if (e[0] == 0) return;
  const int nb = e[1];
  const int* ov = e + 12;
  const int* os = e + 16;
  const int rows = rows_of(ov), cols = cols_of(ov);
  const size_t cnt = (rows > 0 && cols > 0) ? (size_t)rows * cols : 0;
  const size_t mine = (size_t)t * n * n, theirs = (size_t)nb * n * n;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(q / cols), x = (int)(q - (size_t)y * cols);
    const size_t pm = mine + (size_t)(ov[0] + y) * n + ov[2] + x;
    const size_t pn = theirs + (size_t)(os[0] + y) * n + os[2] + x;
    const double odiff = (double)masks[pm] - (double)masks[pn];
    mg[pm] = (float)((double)mg[pm] + cept_weight * (2.0 * odiff));
  }
}

void launch_overlap_cept(const int* table, int num_tiles, int n, int slot, const float* masks, float* mask_gradient,
                         double cept_weight, int max_strip, hipStream_t s) {
  (void)max_strip;
  hipLaunchKernelGGL(k_overlap_cept, dim3(kExchangeBlocks, num_tiles), dim3(256), 0, s, table, n, slot, masks,
                     mask_gradient, cept_weight);
}

// ---------------------------------------------------------------- Core metrics
__global__ void __launch_bounds__(256)
k_core_metrics(const int* __restrict__ table, int n, const float* __restrict__ wafer, const float* __restrict__ targets,
               const float* __restrict__ pe_win, const float* __restrict__ pen_win, double pen_thr,
               double* __restrict__ partial) {
  const int t = blockIdx.y;
  const int* c = table + (size_t)t * kTileStride + 8;
  const int cols = cols_of(c);
  const size_t cnt = (size_t)rows_of(c) * cols;
  const size_t base = (size_t)t * n * n;
  double v[3] = {0.0, 0.0, 0.0};   // PE (loss window), background penalty, PE on the main pattern (target > 0.5)
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(q / cols), x = (int)(q - (size_t)y * cols);
    const size_t p = base + (size_t)(c[0] + y) * n + c[2] + x;
    const double w = wafer[p], tv = targets[p];
    const double d = w - tv;
    v[0] += d * d * (double)pe_win[p];
    const double ex = w - pen_thr > 0.0 ? w - pen_thr : 0.0;
    v[1] += (double)pen_win[p] * (tv < 0.5 ? 1.0 : 0.0) * ex * ex;
    v[2] += tv > 0.5 ? d * d : 0.0;
  }
  __shared__ double sh[kCoreMetrics][4];
  for (int i = 0; i < kCoreMetrics; ++i) {
    double x = v[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if ((threadIdx.x & 63) == 0) sh[i][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < kCoreMetrics) {
    const double* r = sh[threadIdx.x];
    partial[((size_t)t * gridDim.x + blockIdx.x) * kCoreMetrics + threadIdx.x] = ((r[0] + r[1]) + r[2]) + r[3];
  }
}

void launch_core_metrics(const int* table, int num_tiles, int n, const float* wafer, const float* targets,
                         const float* pe_win, const float* pen_win, double penalty_threshold, int max_core_elems,
                         double* partial, hipStream_t s) {
  (void)max_core_elems;
  hipLaunchKernelGGL(k_core_metrics, dim3(kMetricBlocks, num_tiles), dim3(256), 0, s, table, n, wafer, targets,
                     pe_win, pen_win, penalty_threshold, partial);
}

// ---------------------------------------------------------------- CTM theta step
__global__ void __launch_bounds__(256)
k_theta_step(const int* __restrict__ table, int n, const float* __restrict__ grad, double lr, float* __restrict__ theta,
             float* __restrict__ mask) {
  const int t = blockIdx.y;
  const int* u = table + (size_t)t * kTileStride + 12;
  const int cols = cols_of(u);
  const size_t cnt = (size_t)rows_of(u) * cols;
  const size_t base = (size_t)t * n * n;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(q / cols), x = (int)(q - (size_t)y * cols);
    const size_t p = base + (size_t)(u[0] + y) * n + u[2] + x;
    const double th = theta[p];
    const double nt = th - lr * (-0.5 * (double)grad[p] * sin(th));
    const float ntf = (float)nt;
    theta[p] = ntf;
    mask[p] = 0.5f * (1.0f + cosf(ntf));
  }
}

void launch_theta_step(const int* table, int num_tiles, int n, const float* grad, double lr, float* theta,
                       float* mask, int max_update_elems, hipStream_t s) {
  const int blocks = (max_update_elems + 255) / 256 < 1024 ? (max_update_elems + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_theta_step, dim3(blocks > 0 ? blocks : 1, num_tiles), dim3(256), 0, s, table, n, grad, lr,
                     theta, mask);
}

__global__ void k_mask_from_theta_tiles(const float* __restrict__ theta, size_t count, float* __restrict__ mask) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    mask[i] = 0.5f * (1.0f + cosf(theta[i]));
}

// evaluate_mask_continuity (mask_continuity.cpp:8-61): block (blk, slot, t)
// takes a grid-strided share of the Guard strip of neighbour slot `slot` of
// Micro t (my_guard_local against the neighbour's neighbor_overlap_local) and,
// for a right / bottom neighbour, of the Core seam pixel pairs of the
// assembled mask.  partial[((t * 4 + slot) * nblk + blk) * kContinuityValues + v].
__global__ void __launch_bounds__(256)
k_mask_continuity(const int* __restrict__ table, int n, int width, const float* __restrict__ masks,
                  const float* __restrict__ global, const float* __restrict__ target, double* __restrict__ partial) {
  const int t = blockIdx.z, slot = blockIdx.y;
  const int* e = table + (size_t)t * kTileStride + 20 + 20 * slot;
  double v[kContinuityValues];
#pragma unroll
  for (int i = 0; i < kContinuityValues; ++i) v[i] = 0.0;
  if (e[0] != 0) {
    const int nb = e[1], dir = e[2];
    const int* a = e + 4;
    const int* b = e + 8;
    const int rows = rows_of(a), cols = cols_of(a);
    const size_t cnt = (rows > 0 && cols > 0) ? (size_t)rows * cols : 0;
    const size_t mine = (size_t)t * n * n, theirs = (size_t)nb * n * n;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
      const int y = (int)(q / cols), x = (int)(q - (size_t)y * cols);
      const double d = (double)masks[mine + (size_t)(a[0] + y) * n + a[2] + x] -
                       (double)masks[theirs + (size_t)(b[0] + y) * n + b[2] + x];
      v[0] += 1.0;
      v[1] += d * d;
      v[2] += fabs(d);
      v[3] = fmax(v[3], fabs(d));
    }
    if (dir == 1 || dir == 3) {
      const int* core = table + (size_t)t * kTileStride + 4;        // my core_global y0, y1, x0, x1
      const int* ncore = table + (size_t)nb * kTileStride + 4;
      const int len = dir == 1 ? core[1] - core[0] : core[3] - core[2];
      for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < len; q += gridDim.x * blockDim.x) {
        const int by = dir == 1 ? core[0] + q : ncore[0], bx = dir == 1 ? ncore[2] : core[2] + q;
        const int ay = dir == 1 ? by : by - 1, ax = dir == 1 ? bx - 1 : bx;
        const size_t ia = (size_t)ay * width + ax, ib = (size_t)by * width + bx;
        const double jump = (double)global[ib] - (double)global[ia];
        const double ref = (double)target[ib] - (double)target[ia];
        v[4] += 1.0;
        v[5] += fabs(jump);
        v[6] += fabs(jump - ref);
        if (ref == 0.0) {
          v[7] += 1.0;
          v[8] += fabs(jump);
        }
        if ((double)target[ia] > 0.5 && (double)target[ib] > 0.5) {
          v[9] += 1.0;
          v[10] += fabs(jump);
        }
      }
    }
  }
  __shared__ double w[4][kContinuityValues];
#pragma unroll
  for (int i = 0; i < kContinuityValues; ++i) {
    double x = v[i];
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(x, off, 64);
      x = i == 3 ? fmax(x, o) : x + o;
    }
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6][i] = x;
  }
  __syncthreads();
  if (threadIdx.x < kContinuityValues) {
    const int i = threadIdx.x;
    const double r = i == 3 ? fmax(fmax(w[0][i], w[1][i]), fmax(w[2][i], w[3][i])) : ((w[0][i] + w[1][i]) + w[2][i]) + w[3][i];
    partial[(((size_t)t * 4 + slot) * gridDim.x + blockIdx.x) * kContinuityValues + i] = r;
  }
}

void launch_mask_continuity(const int* table, int num_tiles, int n, int width, const float* masks,
                            const float* global, const float* target, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_mask_continuity, dim3(kExchangeBlocks, 4, num_tiles), dim3(256), 0, s, table, n, width, masks,
                     global, target, partial);
}

// FP64 partial sums of (a - b)^2: block `blk` sums elements blk*256 + t,
// strided by the grid, then a fixed shuffle tree and a fixed sum over waves.
__global__ void __launch_bounds__(256) k_sq_diff_partials(const float* __restrict__ a, const float* __restrict__ b,
                                                          size_t n, double* __restrict__ partial) {
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double d = (double)a[i] - (double)b[i];
    acc = fma(d, d, acc);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((w[0] + w[1]) + w[2]) + w[3];
}

void launch_sq_diff_partials(const float* a, const float* b, size_t n, double* partial, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_sq_diff_partials, dim3(blocks), dim3(256), 0, s, a, b, n, partial);
}

void launch_mask_from_theta_tiles(const float* theta, size_t count, float* mask, hipStream_t s) {
  const size_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_mask_from_theta_tiles, dim3(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096), dim3(256), 0, s,
                     theta, count, mask);
}

}  // namespace litho
