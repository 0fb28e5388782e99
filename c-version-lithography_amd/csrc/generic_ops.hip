// SOCS forward / adjoint for tile sizes without compiled transforms.
//
// The reference images any grid size N (imaging.cpp:60-171 runs FFTW on
// N x N; its finite-difference test uses N = 65, test_ctm_optimizer.cpp:164),
// while the FFT engine here is instantiated for a list of sizes.  For every
// other N a plan runs this pipeline: the same stages, buffers and reductions
// as the compiled one (socs_kernels.hip), with each length-N transform
// replaced by a direct DFT restricted to the frequencies that matter.  Only
// the W_c window columns and W_r window rows of the spectra are nonzero or
// needed, and the window frequencies are contiguous, so every sum is
//   X[u] = sum_x x[x] e^{-+2 pi i (f0 + u) x / N}
// with the twiddle index advanced by a constant step modulo N (no
// multiplication or division per term).  Sums accumulate in FP64 from FP32
// data and FP32 twiddles (the table of the plan, computed in FP64).  The
// column stage uses the band-limited correlations of length L (power of two,
// independent of N) unchanged; the column image kernel evaluates the length-N
// inverse of the band-limited autocorrelation and the forward DFT of dL/dI at
// the W_r lags directly.  The work per evaluation is O(N^2 W + K W^2 N)
// against O(N^2 log N) for the compiled transforms: a fallback that keeps any
// tile size working, not the fast path.
#include <hip/hip_runtime.h>

#include "socs_kernels.hpp"

namespace litho {

namespace {

constexpr int kGT = 256;          // threads per block
constexpr int kGBlocks = 64;      // column-kernel blocks per tile (one FP64 partial each)
constexpr int kGImgCols = 2;      // columns per block step of the image kernel (2 n floats of LDS)
constexpr int kGFinCols = 8;      // columns per block step of the final column kernel

__device__ __forceinline__ void acc_cmul(double2& acc, float2 a, float2 w) {
  acc.x += (double)a.x * (double)w.x - (double)a.y * (double)w.y;
  acc.y += (double)a.x * (double)w.y + (double)a.y * (double)w.x;
}
__device__ __forceinline__ void acc_cmulc(double2& acc, float2 a, float2 w) {   // a * conj(w)
  acc.x += (double)a.x * (double)w.x + (double)a.y * (double)w.y;
  acc.y += (double)a.y * (double)w.x - (double)a.x * (double)w.y;
}
__device__ __forceinline__ int step_mod(int idx, int step, int n) {
  idx += step;
  return idx >= n ? idx - n : idx;
}

// G1: rows of the mask -> the W_c window columns of the row spectra,
// rt[b][j][y] = sum_x M[y][x] e^{-2 pi i cf_j x / N}; theta / mask handling as K1.
__global__ void __launch_bounds__(kGT) g_row_fwd(SocsArgs a, int mode, const float* base, const float* __restrict__ dir,
                                                 const float* __restrict__ coef, float* theta_out,
                                                 float* __restrict__ mask_out) {
  extern __shared__ float gm_row[];
  const int n = a.n, y = blockIdx.x, b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const size_t row = (size_t)b * n * n + (size_t)y * n;
  const size_t pt = point_offset(a, b, 1) + (size_t)y * n;                      // trial theta row
  const size_t pb = mode == 1 ? point_offset(a, b, 0) + (size_t)y * n : row;     // origin row (or the mask)
  const float cf = dir != nullptr ? coef[b] : 0.f;
  for (int x = threadIdx.x; x < n; x += kGT) {
    const size_t g = row + x;
    float v = base[pb + x], m;
    if (mode == 1) {
      if (dir != nullptr) v = v + cf * dir[g];
      if (theta_out != nullptr) theta_out[pt + x] = v;
      m = 0.5f * (1.0f + cosf(v));
    } else {
      m = v;
    }
    if (mask_out != nullptr) mask_out[g] = m;
    gm_row[x] = m;
  }
  __syncthreads();
  float2* rt = a.rt + (size_t)b * a.wc * n;
  for (int j = threadIdx.x; j < a.wc; j += kGT) {
    const int v = a.col_freq[j];
    double2 acc = make_double2(0.0, 0.0);
    int idx = 0;
    for (int x = 0; x < n; ++x) {
      const float2 w = a.tw[idx];
      acc.x += (double)gm_row[x] * (double)w.x;
      acc.y += (double)gm_row[x] * (double)w.y;
      idx = step_mod(idx, v, n);
    }
    rt[(size_t)j * n + y] = make_float2((float)acc.x, (float)acc.y);
  }
}

// G2: window columns of the row spectra -> M^[b][i][j] = sum_y rt[b][j][y] e^{-2 pi i rf_i y / N}
__global__ void __launch_bounds__(kGT) g_col_fwd(SocsArgs a) {
  extern __shared__ float2 gm_line[];
  const int n = a.n, j = blockIdx.x, b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const float2* rt = a.rt + (size_t)b * a.wc * n + (size_t)j * n;
  for (int y = threadIdx.x; y < n; y += kGT) gm_line[y] = rt[y];
  __syncthreads();
  float2* mh = a.mhat + (size_t)b * a.wr * a.wc;
  for (int i = threadIdx.x; i < a.wr; i += kGT) {
    const int u = a.row_freq[i];
    double2 acc = make_double2(0.0, 0.0);
    int idx = 0;
    for (int y = 0; y < n; ++y) {
      acc_cmul(acc, gm_line[y], a.tw[idx]);
      idx = step_mod(idx, u, n);
    }
    mh[(size_t)i * a.wc + j] = make_float2((float)acc.x, (float)acc.y);
  }
}

// G3: T_k rows, t[b][q][x] = sum_j H~_q[j] M^[i][j] e^{+2 pi i cf_j x / N}, q = k W_r + i
__global__ void __launch_bounds__(kGT) g_row_inv(SocsArgs a) {
  extern __shared__ float2 gm_line[];
  const int n = a.n, q = blockIdx.x, b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int i = q % a.wr;
  const float2* mh = a.mhat + (size_t)b * a.wr * a.wc + (size_t)i * a.wc;
  const float2* hq = a.h + (size_t)q * a.wc;
  for (int j = threadIdx.x; j < a.wc; j += kGT) {
    const float2 h = hq[j], m = mh[j];
    gm_line[j] = make_float2(h.x * m.x - h.y * m.y, h.x * m.y + h.y * m.x);
  }
  __syncthreads();
  const int f0 = a.col_freq[0];
  float2* t = a.t + (size_t)b * a.k * a.wr * n + (size_t)q * n;
  for (int x = threadIdx.x; x < n; x += kGT) {
    // window columns are consecutive frequencies: the index advances by x per column
    int idx = (int)(((long long)f0 * x) % n);
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < a.wc; ++j) {
      acc_cmulc(acc, gm_line[j], a.tw[idx]);
      idx = step_mod(idx, x, n);
    }
    t[x] = make_float2((float)acc.x, (float)acc.y);
  }
}

// G4: column image from the band-limited autocorrelation and the window
// spectrum of dL/dI.  Per column: I[y] = c[0] + 2 Re sum_{0<d<W_r} c[d] e^{2 pi i d y / N},
// the resist / losses / dL/dI as the compiled image kernels, then
// ghat[m] = sum_y g[y] e^{-2 pi i m y / N} for m < W_r.
__global__ void __launch_bounds__(kGT) g_col_image(SocsArgs a, ImageArgs im) {
  extern __shared__ float g_col[];   // [kGImgCols][n] dL/dI, then [kGImgCols][W_r] c (float2)
  float2* c_col = reinterpret_cast<float2*>(g_col + ((kGImgCols * a.n + 1) & ~1));
  __shared__ double red[2 * kGT / 64];
  const int n = a.n, b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const size_t tile = (size_t)b * n * n;
  const float2* ac = a.acorr + (size_t)b * a.wr * n;
  float2* gh = a.ghat + (size_t)b * a.wr * n;
  const int c = threadIdx.x % kGImgCols, lane = threadIdx.x / kGImgCols;
  constexpr int LANES = kGT / kGImgCols;
  double pe = 0.0, bg = 0.0;
  for (int x0 = blockIdx.x * kGImgCols; x0 < n; x0 += gridDim.x * kGImgCols) {
    const int x = x0 + c;
    for (int e = threadIdx.x; e < kGImgCols * a.wr; e += kGT) {
      const int cc = e % kGImgCols, d = e / kGImgCols;
      c_col[cc * a.wr + d] = x0 + cc < n ? ac[(size_t)d * n + x0 + cc] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2* cx = c_col + c * a.wr;
    for (int y = lane; y < n; y += LANES) {
      float gv = 0.f;
      if (x < n) {
        double s = 0.0;
        int idx = 0;
        for (int d = 1; d < a.wr; ++d) {
          idx = step_mod(idx, y, n);
          const float2 cd = cx[d], w = a.tw[idx];
          s += (double)cd.x * (double)w.x + (double)cd.y * (double)w.y;   // Re(c[d] conj(w)), conj(w) = e^{+i}
        }
        const float iv = (float)((double)cx[0].x + 2.0 * s);
        gv = image_pixel(im, tile + (size_t)y * n + x, iv, pe, bg);
      }
      g_col[c * n + y] = gv;
    }
    __syncthreads();
    if (im.target != nullptr && a.v != nullptr && x < n) {
      for (int m = lane; m < a.wr; m += LANES) {
        double2 acc = make_double2(0.0, 0.0);
        int idx = 0;
        for (int y = 0; y < n; ++y) {
          const float gv = g_col[c * n + y];
          const float2 w = a.tw[idx];
          acc.x += (double)gv * (double)w.x;
          acc.y += (double)gv * (double)w.y;
          idx = step_mod(idx, m, n);
        }
        gh[(size_t)m * n + x] = make_float2((float)acc.x, (float)acc.y);
      }
    }
    __syncthreads();
  }
  for (int off = 32; off > 0; off >>= 1) {
    pe += __shfl_down(pe, off, 64);
    bg += __shfl_down(bg, off, 64);
  }
  const int wave = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0) {
    red[2 * wave] = pe;
    red[2 * wave + 1] = bg;
  }
  __syncthreads();
  if (threadIdx.x == 0 && im.partials != nullptr) {
    double sp = 0.0, sb = 0.0;
    for (int w = 0; w < kGT / 64; ++w) {
      sp += red[2 * w];
      sb += red[2 * w + 1];
    }
    double* out = im.partials + ((size_t)b * gridDim.x + blockIdx.x) * 2;
    out[0] = sp;
    out[1] = sb;
  }
}

// G6: Y_k[q][j] = conj(H~_q[j]) sum_x V[q][x] e^{-2 pi i cf_j x / N}
__global__ void __launch_bounds__(kGT) g_row_adj(SocsArgs a) {
  extern __shared__ float2 gm_line[];
  const int n = a.n, q = blockIdx.x, b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const float2* v = a.v + (size_t)b * a.k * a.wr * n + (size_t)q * n;
  for (int x = threadIdx.x; x < n; x += kGT) gm_line[x] = v[x];
  __syncthreads();
  float2* yk = a.y + (size_t)b * a.k * a.wr * a.wc + (size_t)q * a.wc;
  const float2* hq = a.h + (size_t)q * a.wc;
  for (int j = threadIdx.x; j < a.wc; j += kGT) {
    const int f = a.col_freq[j];
    double2 acc = make_double2(0.0, 0.0);
    int idx = 0;
    for (int x = 0; x < n; ++x) {
      acc_cmul(acc, gm_line[x], a.tw[idx]);
      idx = step_mod(idx, f, n);
    }
    const float2 s = make_float2((float)acc.x, (float)acc.y), h = hq[j];
    yk[j] = make_float2(s.x * h.x + s.y * h.y, s.y * h.x - s.x * h.y);   // s * conj(h)
  }
}

// G7: P[i][x] = sum_j (sum_k Y_k[i][j]) e^{+2 pi i cf_j x / N}, k summed in increasing order
__global__ void __launch_bounds__(kGT) g_final_row_inv(SocsArgs a) {
  extern __shared__ float2 gm_line[];
  const int n = a.n, i = blockIdx.x, b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const float2* yk = a.y + (size_t)b * a.k * a.wr * a.wc + (size_t)i * a.wc;
  const size_t ks = (size_t)a.wr * a.wc;
  for (int j = threadIdx.x; j < a.wc; j += kGT) {
    float2 s = make_float2(0.f, 0.f);
    for (int k = 0; k < a.k; ++k) {
      const float2 u = yk[(size_t)k * ks + j];
      s.x += u.x;
      s.y += u.y;
    }
    gm_line[j] = s;
  }
  __syncthreads();
  const int f0 = a.col_freq[0];
  float2* p = a.p + (size_t)b * a.wr * n + (size_t)i * n;
  for (int x = threadIdx.x; x < n; x += kGT) {
    int idx = (int)(((long long)f0 * x) % n);
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < a.wc; ++j) {
      acc_cmulc(acc, gm_line[j], a.tw[idx]);
      idx = step_mod(idx, x, n);
    }
    p[x] = make_float2((float)acc.x, (float)acc.y);
  }
}

// G8: G_M[y][x] = scale Re sum_i P[i][x] e^{+2 pi i rf_i y / N}, then the theta
// chain and the FP64 norm / direction partials as the compiled final column kernel.
__global__ void __launch_bounds__(kGT) g_final_col_inv(SocsArgs a, GradArgs ga) {
  extern __shared__ float2 g_pc[];   // [kGFinCols][wr]
  __shared__ double red[2 * kGT / 64];
  const int n = a.n, b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const size_t tile = (size_t)b * n * n;
  const size_t pt = point_offset(a, b, 1);   // trial theta and its gradient
  const float2* p = a.p + (size_t)b * a.wr * n;
  const int* rect = ga.rect != nullptr ? ga.rect + 4 * b : nullptr;
  const int c = threadIdx.x % kGFinCols, lane = threadIdx.x / kGFinCols;
  constexpr int LANES = kGT / kGFinCols;
  const int r0 = a.row_freq[0];
  double nrm = 0.0, dot = 0.0;
  for (int x0 = blockIdx.x * kGFinCols; x0 < n; x0 += gridDim.x * kGFinCols) {
    __syncthreads();
    for (int e = threadIdx.x; e < kGFinCols * a.wr; e += kGT) {
      const int cc = e % kGFinCols, i = e / kGFinCols, x = x0 + cc;
      g_pc[cc * a.wr + i] = x < n ? p[(size_t)i * n + x] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int x = x0 + c;
    if (x >= n) continue;
    for (int y = lane; y < n; y += LANES) {
      int idx = (int)(((long long)r0 * y) % n);
      double s = 0.0;
      for (int i = 0; i < a.wr; ++i) {
        const float2 pv = g_pc[c * a.wr + i], w = a.tw[idx];
        s += (double)pv.x * (double)w.x + (double)pv.y * (double)w.y;   // Re(p conj(w))
        idx = step_mod(idx, y, n);
      }
      const size_t lg = (size_t)y * n + x, gi = tile + lg;
      const float gmv = ga.scale * (float)s;
      if (ga.grad_mask != nullptr) ga.grad_mask[gi] = gmv;
      if (ga.theta != nullptr) {
        float gt = -0.5f * gmv * sinf(ga.theta[pt + lg]);
        if (rect != nullptr && (y < rect[0] || y >= rect[1] || x < rect[2] || x >= rect[3])) gt = 0.f;
        if (ga.grad_theta != nullptr) ga.grad_theta[pt + lg] = gt;
        nrm += (double)gt * (double)gt;
        if (ga.dir != nullptr) dot += (double)gt * (double)ga.dir[gi];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    nrm += __shfl_down(nrm, off, 64);
    dot += __shfl_down(dot, off, 64);
  }
  const int wave = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0) {
    red[2 * wave] = nrm;
    red[2 * wave + 1] = dot;
  }
  __syncthreads();
  if (threadIdx.x == 0 && ga.partials != nullptr) {
    double sn = 0.0, sd = 0.0;
    for (int w = 0; w < kGT / 64; ++w) {
      sn += red[2 * w];
      sd += red[2 * w + 1];
    }
    double* out = ga.partials + ((size_t)b * gridDim.x + blockIdx.x) * 2;
    out[0] = sn;
    out[1] = sd;
  }
}

int g_blocks() { return kGBlocks; }

void g_forward_half(const SocsArgs& a, int batch, int mode, const float* base, const float* dir, const float* coef,
                    float* theta_out, float* mask_out, hipStream_t s) {
  const int n = a.n;
  hipLaunchKernelGGL(g_row_fwd, dim3(n, batch), dim3(kGT), (size_t)n * sizeof(float), s, a, mode, base, dir, coef,
                     theta_out, mask_out);
  hipLaunchKernelGGL(g_col_fwd, dim3(a.wc, batch), dim3(kGT), (size_t)n * sizeof(float2), s, a);
  hipLaunchKernelGGL(g_row_inv, dim3(a.k * a.wr, batch), dim3(kGT), (size_t)a.wc * sizeof(float2), s, a);
}

void g_image_adjoint(const SocsArgs& a, const ImageArgs& im, int batch, hipStream_t s, int part) {
  if (part != kColumnAdjoint) {
    launch_corr_any(0, a, batch, a.spart, s);
    const size_t lds = (size_t)((kGImgCols * a.n + 1) & ~1) * sizeof(float) + (size_t)kGImgCols * a.wr * sizeof(float2);
    hipLaunchKernelGGL(g_col_image, dim3(kGBlocks, batch), dim3(kGT), lds, s, a, im);
  }
  if (part != kColumnImage && im.target != nullptr && a.v != nullptr) {
    if (a.adj_mfma)
      launch_corr_adjoint_mfma(a, batch, s);
    else
      launch_corr_any(1, a, batch, nullptr, s);
  }
}

void g_adjoint_tail(const SocsArgs& a, const GradArgs& ga, int batch, hipStream_t s) {
  hipLaunchKernelGGL(g_row_adj, dim3(a.k * a.wr, batch), dim3(kGT), (size_t)a.n * sizeof(float2), s, a);
  hipLaunchKernelGGL(g_final_row_inv, dim3(a.wr, batch), dim3(kGT), (size_t)a.wc * sizeof(float2), s, a);
  hipLaunchKernelGGL(g_final_col_inv, dim3(kGBlocks, batch), dim3(kGT),
                     (size_t)kGFinCols * a.wr * sizeof(float2), s, a, ga);
}

}  // namespace

const SocsOps* generic_socs_ops() {
  static const SocsOps ops = {0, g_blocks, g_blocks, g_forward_half, g_image_adjoint, g_adjoint_tail};
  return &ops;
}

}  // namespace litho
