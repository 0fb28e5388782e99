// SOCS forward imaging, PE loss and adjoint gradient for one or a batch of
// equal-size tiles, hand-written for gfx950.
//
// The reference (imaging.cpp:60-171, gradient.cpp:34-109) runs 3K+1 full N x N
// FFTs per gradient evaluation and materialises K complex fields.  Here every
// transform is pruned to the pupil's W_r x W_c frequency window (all SOCS
// kernels vanish outside it, litho_prepare.cpp:238-256,330) and the fields are
// never stored: they are rebuilt from the K row-transformed spectra T_k
// (K x W_r x N) inside the fused column kernel.  The ifftshift / fftshift
// pairs around the reference's transforms cancel exactly (circular
// convolution commutes with the shift), so the device computes
//   E_k  = IFFT(H_k . FFT(M))
//   I    = sum_k |E~_k|^2,  E~_k = E_k sqrt(lambda_k / sum w)
//   W    = 1 / (1 + exp(-alpha (I - thr)))
//   dLdI = (pe_w (W - T) pe_win + pen [T < .5] max(W - tau, 0) pen_win) W (1 - W)
//   G_M  = 4 alpha Re IFFT( sum_k conj(H~_k) . FFT(E~_k dLdI) )
// with H~_k = H_k sqrt(lambda_k / sum w) / N^2 folded in by the host.
//
// Pipeline per evaluation (b = tile index = blockIdx.y):
//   k_row_fwd        theta/mask rows -> row spectra at the W_c window columns
//   k_col_fwd        window columns -> M^ (W_r x W_c)
//   k_row_inv        (k, window row) -> T_k rows  (H~_k . M^, inverse along x)
//   k_col_image_adj  columns: inverse along y for every k -> I, W, dLdI, PE;
//                    then rebuild E~_k, multiply by dLdI, forward along y ->
//                    V_k rows at the window rows
//   k_row_adj        (k, window row): forward along x -> Y_k = conj(H~_k) V^
//   k_final_row_inv  window rows: sum_k Y_k, inverse along x -> P
//   k_final_col_inv  columns: inverse along y -> G_M, theta chain, norms
//   (per-block FP64 partials are summed by launch_reduce_fixed, opt_kernels.hip)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <string>

#include "fft_columns.hpp"
#include "fft_engine.hpp"
#include "fft_regs.hpp"
#include "fft_rows.hpp"
#include "opt_kernels.hpp"
#include "socs_kernels.hpp"


namespace litho {

// Non-temporal load (the nt bit of global_load) of a value read once and not
// again before HBM holds it anew: the row adjoint's V_k loads.  With them a
// 2048^2 L-BFGS solve runs about 3 % faster; the same policy on the
// correlation kernels' T_k loads, the final column kernel's theta and d
// loads, the final row kernel's Y_k loads or the spectrum partials measured
// slower or no faster.
__device__ __forceinline__ float2 ld_nt2(const float2* p) {
  typedef float nf2 __attribute__((ext_vector_type(2)));
  const nf2 v = __builtin_nontemporal_load(reinterpret_cast<const nf2*>(p));
  return make_float2(v.x, v.y);
}

// Non-temporal store of the trial theta an evaluation writes into its point
// slot, which the L-BFGS sweeps read non-temporally from HBM later: a 2048^2
// L-BFGS solve runs about 1 % faster with it, batched tiles the same.  The
// theta gradient is stored the same way only by evaluations of one tile
// (GradArgs::nt_grad_theta): 0.5-1 % faster at 2048^2, while in batches of
// 640^2-1152^2 tiles it cost 7 %.  On the T_k and V_k stores it measured no
// faster; those stores stay plain.
__device__ __forceinline__ void st_nt1(float* p, float v) { __builtin_nontemporal_store(v, p); }

using fft::cconj;
using fft::cmul;
using fft::cmulc;

// ---------------------------------------------------------------- tunables
template <int N>
struct RowCfg {
  static constexpr int LINES = N >= 1024 ? 4 : (N >= 256 ? 8 : 16);
  static constexpr int NT = N >= 512 ? 256 : (N >= 128 ? 128 : 64);
  static constexpr int SH = 5;                          // swizzle shift: lanes run along one line
  static constexpr int LS = fft::padded_stride<SH>(N);  // line stride (float2), swizzled lines
};

// Pruned row kernels (fft_rows.hpp): LINES lines per block for the kernels
// over all K * W_r spectrum rows, WLINES for the kernels over the W_r or W_c
// window lines, which are too few to fill the chip with wide blocks.
#ifndef LITHO_ROWP_LINES_LARGE
#define LITHO_ROWP_LINES_LARGE 1
#endif
#ifndef LITHO_ROWP_WLINES_LARGE
#define LITHO_ROWP_WLINES_LARGE 1
#endif
#ifndef LITHO_ROWP_LINES_MID
#define LITHO_ROWP_LINES_MID 2
#endif
// about one radix-16 butterfly per thread, in whole waves, 64..256 threads
constexpr int row_threads(int lines, int n) {
  return (lines * n / 16 + 63) / 64 * 64 < 64 ? 64 : ((lines * n / 16 + 63) / 64 * 64 > 256 ? 256 : (lines * n / 16 + 63) / 64 * 64);
}
template <int N>
struct RowPCfg {
  static constexpr int SH = 5;
  static constexpr int LS = fft::padded_stride<SH>(N);
  static constexpr int LINES = N >= 1024 ? LITHO_ROWP_LINES_LARGE : (N >= 256 ? LITHO_ROWP_LINES_MID : 8);
  static constexpr int NT = row_threads(LINES, N);
  static constexpr int WLINES = N >= 1024 ? LITHO_ROWP_WLINES_LARGE : (N >= 256 ? 2 : 4);
  static constexpr int WNT = row_threads(WLINES, N);
  // mask-row kernel K1: 256 threads from N = 256 (2048^2 forward stage
  // 0.050 ms against 0.055 with 128; 640^2 batch 8: 0.040 against 0.044)
  static constexpr int K1NT = N >= 256 ? 256 : NT;
};

// band-limited image kernel with two columns per complex line
// (k_col_image_corr2); LITHO_IMAGE_PAIRS = 0 in the environment selects the
// one-column-per-line kernel (read once per process)
static bool image_pairs() {
  static const bool on = [] {
    const char* v = std::getenv("LITHO_IMAGE_PAIRS");
    return v == nullptr || std::string(v) != "0";
  }();
  return on;
}
// The register-resident image kernels (k_col_image_reg at 1024^2,
// k_col_image_reg2 at 2048^2) where they measured faster; LITHO_IMAGE_REG = 0
// in the environment selects the LDS-resident kernels there too (read once
// per process)
static bool image_reg() {
  static const bool on = [] {
    const char* v = std::getenv("LITHO_IMAGE_REG");
    return v == nullptr || std::string(v) != "0";
  }();
  return on;
}

#ifndef LITHO_COL_NT_LARGE
#define LITHO_COL_NT_LARGE 1024
#endif
#ifndef LITHO_TW_LDS
#define LITHO_TW_LDS 1
#endif
// threads of the column kernels for 256 <= N < 1024; 0 = about eight owned
// elements per thread (LINES * N / 8, rounded up to whole waves)
#ifndef LITHO_COL_NT_MID
#define LITHO_COL_NT_MID 512
#endif

// register budget of the column kernels: 0 = bounded by the block size only,
// 1 = at least the waves per SIMD that the LDS-limited block count provides
#ifndef LITHO_COL_OCC_RULE
#define LITHO_COL_OCC_RULE 0
#endif

constexpr int col_nt_mid(int lines, int n) {
  return LITHO_COL_NT_MID > 0 ? LITHO_COL_NT_MID
                              : ((lines * n / 8 + 63) / 64 * 64 > 1024 ? 1024 : (lines * n / 8 + 63) / 64 * 64);
}

// columns per block of the column kernels for the powers of two from 1024 and
// for 640 <= N < 1024 (measured alone: 8 x 1024 threads best from 1152 up,
// 4 x 512 threads for 640..896, 8 x 512 threads below 640 when several tiles
// share a launch; the sizes from 1024 that are not powers of two take the
// 4 x 512 blocks, see ColCfg)
#ifndef LITHO_COL_LINES_LARGE
#define LITHO_COL_LINES_LARGE 8
#endif
#ifndef LITHO_COL_LINES_MID
#define LITHO_COL_LINES_MID 4
#endif

// largest radix of the factorisation: 11, 13 and 17 run as direct O(R^2)
// butterflies whose registers crowd out the 1024-thread, 8-column blocks
constexpr int max_radix(int n) {
  const fft::Radices f = fft::factorize(n);
  int m = 0;
  for (int i = 0; i < f.count; ++i) m = f.r[i] > m ? f.r[i] : m;
  return m;
}

// The 8-column, 1024-thread blocks (about 80 KB of LDS) are kept for the
// powers of two from 1024, which run alone (the 2048^2 tile).  The other sizes
// from 1024 share the GPU with other streams in the batched and served paths,
// where such a block waits for a CU with that much room free: the 1152^2
// final column kernel took 3-6 ms per batch-32 launch beside the 640 / 896
// streams against 0.24 ms alone, and it ran on 4-column, 512-thread blocks
// instead.
template <int N>
struct ColCfg {
  static constexpr bool BIG_ODD = N >= 1024 && max_radix(N) > 16;
  static constexpr bool WIDE = N >= 1024 && (N & (N - 1)) == 0;
  static constexpr int LINES =
      BIG_ODD ? 4 : (WIDE ? LITHO_COL_LINES_LARGE : (N >= 640 ? LITHO_COL_LINES_MID : (N >= 256 ? 8 : 16)));
  static constexpr int NT = BIG_ODD ? 512 : (WIDE ? LITHO_COL_NT_LARGE : (N >= 256 ? col_nt_mid(LINES, N) : 64));
  // swizzle shift: lanes spread over LINES lines; the four-column blocks of the
  // large odd sizes behave like the row kernels
  static constexpr int SH = BIG_ODD ? 5 : 4;
  static constexpr int LS = fft::padded_stride<SH>(N);  // line stride (float2), swizzled lines
  static constexpr int ELEMS = LINES * N;
  static constexpr int PER = (ELEMS + NT - 1) / NT;     // owned elements per thread
  // middle-stage twiddle tables of the pruned column transforms (always in LDS)
  static constexpr int MID = fft::mid_tables(N).off[fft::mid_tables(N).count];
  static constexpr size_t FIXED_BYTES = (size_t)LINES * LS * 8 + (size_t)(MID > 0 ? MID : 1) * 8 + (size_t)NT / 64 * 16;
  // N-entry twiddle table copied to LDS when it fits beside the lines (160 KB per CU)
  static constexpr bool TW_LDS = LITHO_TW_LDS && FIXED_BYTES + (size_t)N * 8 <= 160 * 1024;
  static constexpr size_t LDS_BYTES = FIXED_BYTES + (TW_LDS ? (size_t)N * 8 : 0);
  // register budget: as many waves per SIMD as the blocks that fit in one CU's LDS bring
  static constexpr int BLOCKS_PER_CU = 160 * 1024 / LDS_BYTES > 0 ? (int)(160 * 1024 / LDS_BYTES) : 1;
  static constexpr int LDS_WAVES = BLOCKS_PER_CU * NT / 256 > 8 ? 8 : (BLOCKS_PER_CU * NT / 256 > 0 ? BLOCKS_PER_CU * NT / 256 : 1);
  static constexpr int MIN_WAVES_PER_EU = LITHO_COL_OCC_RULE == 1 ? LDS_WAVES : 1;
  static constexpr int MAX_WAVES_PER_EU = LITHO_COL_OCC_RULE == 1 ? 8 : (NT >= 512 ? NT / 256 : 1);
};

// Block shape of the band-limited image kernel (one length-N inverse and one
// forward per column, stage tables in LDS), tuned apart from the column
// kernels (litho_bench evalb, image stage): 8 columns x 1024 threads for the
// powers of two from 1024 (2048, b = 1: 0.065 ms against 0.081 with 4 x 512),
// 4 x 512 for the other sizes from 1024 (1152, b = 8: 0.159 against 0.172),
// 8 x 512 from 640 (640, b = 8: 0.073 against 0.079 with 4 x 512).
template <int N>
struct ImgCfg {
  using CC = ColCfg<N>;
  static constexpr bool OWN = !CC::BIG_ODD && N >= 640;
  static constexpr bool POW2 = (N & (N - 1)) == 0;
  static constexpr int LINES = !OWN ? CC::LINES : (N >= 1024 ? (POW2 ? 8 : 4) : 8);
  static constexpr int NT = !OWN ? CC::NT : (N >= 1024 ? (POW2 ? 1024 : 512) : 512);
  static constexpr int SH = CC::SH;
  static constexpr int LS = fft::padded_stride<SH>(N);
  static constexpr int ELEMS = LINES * N;
  static constexpr int PER = (ELEMS + NT - 1) / NT;
  static constexpr int TW = fft::staged_entries(N);
  static constexpr size_t LDS_BYTES = (size_t)LINES * LS * 8 + (size_t)(TW > 0 ? TW : 1) * 8 + (size_t)NT / 64 * 16;
  static constexpr int BLOCKS_PER_CU = 160 * 1024 / LDS_BYTES > 0 ? (int)(160 * 1024 / LDS_BYTES) : 1;
  static constexpr int WAVES = BLOCKS_PER_CU * NT / 256 > 8 ? 8 : (BLOCKS_PER_CU * NT / 256 > 0 ? BLOCKS_PER_CU * NT / 256 : 1);
};

// Twiddle sources of the pruned column kernels: the N-entry table (LDS copy or
// global) and the middle-stage tables in LDS.  The opaque zero offset taken
// per call stops the compiler from hoisting the (loop-invariant) twiddle loads
// out of the K loops, which would keep ~60 twiddle registers live across the
// loop and spill.
template <int N>
struct ColTwiddles {
  float2* lds;
  const float2* glob;
  float2* mid;
  __device__ __forceinline__ fft::ColTw get() const {
    int z = 0;
    asm volatile("" : "+s"(z));
    if constexpr (ColCfg<N>::TW_LDS) return {lds + z, mid + z};
    return {glob + z, mid + z};
  }
};

// ---------------------------------------------------------------- K1
// Rows of the mask -> W_c window columns of the row spectra, two real rows
// packed per complex line.  theta = base + coef[b] * dir (dir optional);
// mode 0: base is the mask itself; mode 1: base is theta, M = (1 + cos theta)/2.
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::NT)
k_row_fwd(SocsArgs a, int mode, const float* base, const float* __restrict__ dir, const float* __restrict__ coef,
          float* theta_out, float* __restrict__ mask_out) {
  using C = RowCfg<N>;
  constexpr int PAIRS = C::LINES;        // complex lines per block, 2 rows each
  __shared__ float2 buf[PAIRS * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const size_t tile = (size_t)b * N * N;
  const size_t pb = point_offset(a, b, 0), pt = point_offset(a, b, 1);   // theta origin and trial
  const int row0 = blockIdx.x * 2 * PAIRS;
  const float cf = (dir != nullptr) ? coef[b] : 0.f;
  for (int e = threadIdx.x; e < PAIRS * N; e += C::NT) {
    const int l = e / N, x = e - l * N;
    float m2[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int y = row0 + 2 * l + h;
      float m = 0.f;
      if (y < N) {
        const size_t lg = (size_t)y * N + x, g = tile + lg;
        float v = base[(mode == 1 ? pb : tile) + lg];
        if (mode == 1) {
          if (dir != nullptr) v = v + cf * dir[g];
          if (theta_out != nullptr) st_nt1(theta_out + pt + lg, v);
          m = 0.5f * (1.0f + cosf(v));
        } else {
          m = v;
        }
        if (mask_out != nullptr) mask_out[g] = m;
      }
      m2[h] = m;
    }
    buf[l * C::LS + fft::swz<C::SH>(x)] = make_float2(m2[0], m2[1]);
  }
  __syncthreads();
  fft::fft_lines<N, PAIRS, C::LS, C::NT, -1, C::SH>(buf, a.tw);
  // unpack X_a[v] = (Z[v] + conj Z[-v]) / 2, X_b[v] = (Z[v] - conj Z[-v]) / (2i)
  float2* rt = a.rt + (size_t)b * a.wc * N;
  for (int e = threadIdx.x; e < PAIRS * a.wc; e += C::NT) {
    const int l = e / a.wc, j = e - l * a.wc;
    const int v = a.col_freq[j];
    const int vn = v == 0 ? 0 : N - v;
    const float2 z = buf[l * C::LS + fft::swz<C::SH>(v)], zc = cconj(buf[l * C::LS + fft::swz<C::SH>(vn)]);
    const float2 xa = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    const float2 xb = make_float2(0.5f * (z.y - zc.y), -0.5f * (z.x - zc.x));
    const int y = row0 + 2 * l;
    if (y < N) rt[(size_t)j * N + y] = xa;
    if (y + 1 < N) rt[(size_t)j * N + y + 1] = xb;
  }
}

// ---------------------------------------------------------------- K2
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::NT)
k_col_fwd(SocsArgs a) {
  using C = RowCfg<N>;
  __shared__ float2 buf[C::LINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const float2* rt = a.rt + (size_t)b * a.wc * N;
  const int j0 = blockIdx.x * C::LINES;
  for (int e = threadIdx.x; e < C::LINES * N; e += C::NT) {
    const int l = e / N, y = e - l * N;
    const int j = j0 + l;
    buf[l * C::LS + fft::swz<C::SH>(y)] = j < a.wc ? rt[(size_t)j * N + y] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft::fft_lines<N, C::LINES, C::LS, C::NT, -1, C::SH>(buf, a.tw);
  float2* mh = a.mhat + (size_t)b * a.wr * a.wc;
  for (int e = threadIdx.x; e < C::LINES * a.wr; e += C::NT) {
    const int l = e / a.wr, i = e - l * a.wr;
    const int j = j0 + l;
    if (j < a.wc) mh[(size_t)i * a.wc + j] = buf[l * C::LS + fft::swz<C::SH>(a.row_freq[i])];
  }
}

// ---------------------------------------------------------------- K3
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::NT)
k_row_inv(SocsArgs a) {
  using C = RowCfg<N>;
  __shared__ float2 buf[C::LINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int rows = a.k * a.wr;
  const int q0 = blockIdx.x * C::LINES;
  for (int e = threadIdx.x; e < C::LINES * N; e += C::NT) {
    const int l = e / N, x = e - l * N;
    buf[l * C::LS + fft::swz<C::SH>(x)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* mh = a.mhat + (size_t)b * a.wr * a.wc;
  for (int e = threadIdx.x; e < C::LINES * a.wc; e += C::NT) {
    const int l = e / a.wc, j = e - l * a.wc;
    const int q = q0 + l;
    if (q < rows) {
      const int i = q % a.wr;
      buf[l * C::LS + fft::swz<C::SH>(a.col_freq[j])] = cmul(a.h[(size_t)q * a.wc + j], mh[(size_t)i * a.wc + j]);
    }
  }
  __syncthreads();
  fft::fft_lines<N, C::LINES, C::LS, C::NT, +1, C::SH>(buf, a.tw);
  float2* t = a.t + (size_t)b * rows * N;
  for (int e = threadIdx.x; e < C::LINES * N; e += C::NT) {
    const int l = e / N, x = e - l * N;
    const int q = q0 + l;
    if (q < rows) t[(size_t)q * N + x] = buf[l * C::LS + fft::swz<C::SH>(x)];
  }
}

// ---------------------------------------------------------------- K4+K5
// Column block [x0, x0 + LINES).  Element ownership: e = tid + s*NT,
// column c = e % LINES, row y = e / LINES (coalesced LINES-wide segments).
template <int N>
__device__ __forceinline__ void load_window_columns(float2* buf, const float2* __restrict__ src, int wr,
                                                    const int* __restrict__ row_freq, int x0) {
  using C = ColCfg<N>;
  for (int e = threadIdx.x; e < wr * C::LINES; e += C::NT) {
    const int i = e / C::LINES, c = e - i * C::LINES;
    const int x = x0 + c;
    if (x < N) buf[c * C::LS + fft::swz<C::SH>(row_freq[i])] = src[(size_t)i * N + x];
  }
}

template <int N>
__global__ void __launch_bounds__(ColCfg<N>::NT) __attribute__((amdgpu_waves_per_eu(ColCfg<N>::MIN_WAVES_PER_EU, ColCfg<N>::MAX_WAVES_PER_EU)))
k_col_image_adj(SocsArgs a, ImageArgs im) {
  using C = ColCfg<N>;
  __shared__ float2 buf[C::LINES * C::LS];
  __shared__ double red[2 * C::NT / 64];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int x0 = blockIdx.x * C::LINES;
  const size_t tile = (size_t)b * N * N;
  const int rows = a.k * a.wr;
  const float2* t = a.t + (size_t)b * rows * N;

  float acc[C::PER];
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    acc[s] = 0.f;
    const int e = threadIdx.x + s * C::NT;
    if (e < C::ELEMS) buf[(e % C::LINES) * C::LS + fft::swz<C::SH>(e / C::LINES)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  // ---- pass 1: aerial image
  for (int k = 0; k < a.k; ++k) {
    load_window_columns<N>(buf, t + (size_t)k * a.wr * N, a.wr, a.row_freq, x0);
    __syncthreads();
    fft::fft_lines<N, C::LINES, C::LS, C::NT, +1, C::SH>(buf, a.tw);
#pragma unroll
    for (int s = 0; s < C::PER; ++s) {
      const int e = threadIdx.x + s * C::NT;
      if (e < C::ELEMS) {
        float2* p = buf + (e % C::LINES) * C::LS + fft::swz<C::SH>(e / C::LINES);
        const float2 z = *p;
        acc[s] = fmaf(z.x, z.x, fmaf(z.y, z.y, acc[s]));
        *p = make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
  }
  // ---- resist, loss, dLdI
  float g[C::PER];
  double pe = 0.0, bg = 0.0;
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    g[s] = 0.f;
    const int e = threadIdx.x + s * C::NT;
    const int c = e % C::LINES, y = e / C::LINES;
    const int x = x0 + c;
    if (e < C::ELEMS && x < N) {
      const size_t gi = tile + (size_t)y * N + x;
      const float iv = acc[s];
      const float w = 1.0f / (1.0f + expf(-im.alpha * (iv - im.thr)));
      if (im.aerial != nullptr) im.aerial[gi] = iv;
      if (im.wafer != nullptr) im.wafer[gi] = w;
      if (im.target != nullptr) {
        const float tv = im.target[gi];
        const float r = w - tv;
        const float pw = im.pe_win != nullptr ? im.pe_win[gi] : 1.0f;
        float drive = im.pe_weight * r * pw;
        pe += (double)r * (double)r * (double)pw;
        if (im.pen_win != nullptr) {
          const float bw = im.pen_win[gi];
          const float ex = fmaxf(w - im.pen_thr, 0.f);
          const float bgm = tv < 0.5f ? 1.f : 0.f;
          bg += (double)bw * (double)bgm * (double)ex * (double)ex;
          drive += im.penalty * bgm * ex * bw;
        }
        g[s] = drive * w * (1.0f - w);
        if (im.extra_dldi != nullptr) g[s] += im.extra_dldi[gi];
        if (im.dldi != nullptr) im.dldi[gi] = g[s];
      }
    }
  }
  // block reduction of the two FP64 partials (fixed order)
  {
    for (int off = 32; off > 0; off >>= 1) {
      pe += __shfl_down(pe, off, 64);
      bg += __shfl_down(bg, off, 64);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
      red[2 * wave] = pe;
      red[2 * wave + 1] = bg;
    }
    __syncthreads();
    if (threadIdx.x == 0 && im.partials != nullptr) {
      double sp = 0.0, sb = 0.0;
      for (int w = 0; w < C::NT / 64; ++w) {
        sp += red[2 * w];
        sb += red[2 * w + 1];
      }
      double* out = im.partials + ((size_t)b * gridDim.x + blockIdx.x) * 2;
      out[0] = sp;
      out[1] = sb;
    }
  }
  if (im.target == nullptr || a.v == nullptr) return;
  // ---- pass 2: rebuild E~_k, multiply by dLdI, forward along y
  float2* v = a.v + (size_t)b * rows * N;
  for (int k = 0; k < a.k; ++k) {
    load_window_columns<N>(buf, t + (size_t)k * a.wr * N, a.wr, a.row_freq, x0);
    __syncthreads();
    fft::fft_lines<N, C::LINES, C::LS, C::NT, +1, C::SH>(buf, a.tw);
#pragma unroll
    for (int s = 0; s < C::PER; ++s) {
      const int e = threadIdx.x + s * C::NT;
      if (e < C::ELEMS) {
        float2* p = buf + (e % C::LINES) * C::LS + fft::swz<C::SH>(e / C::LINES);
        const float2 z = *p;
        *p = make_float2(z.x * g[s], z.y * g[s]);
      }
    }
    __syncthreads();
    fft::fft_lines<N, C::LINES, C::LS, C::NT, -1, C::SH>(buf, a.tw);
    float2* vk = v + (size_t)k * a.wr * N;
    for (int e = threadIdx.x; e < a.wr * C::LINES; e += C::NT) {
      const int i = e / C::LINES, c = e - i * C::LINES;
      const int x = x0 + c;
      if (x < N) vk[(size_t)i * N + x] = buf[c * C::LS + fft::swz<C::SH>(a.row_freq[i])];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < C::PER; ++s) {
      const int e = threadIdx.x + s * C::NT;
      if (e < C::ELEMS) buf[(e % C::LINES) * C::LS + fft::swz<C::SH>(e / C::LINES)] = make_float2(0.f, 0.f);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- K6
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::NT)
k_row_adj(SocsArgs a) {
  using C = RowCfg<N>;
  __shared__ float2 buf[C::LINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int rows = a.k * a.wr;
  const int q0 = blockIdx.x * C::LINES;
  const float2* v = a.v + (size_t)b * rows * N;
  for (int e = threadIdx.x; e < C::LINES * N; e += C::NT) {
    const int l = e / N, x = e - l * N;
    const int q = q0 + l;
    buf[l * C::LS + fft::swz<C::SH>(x)] = q < rows ? v[(size_t)q * N + x] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft::fft_lines<N, C::LINES, C::LS, C::NT, -1, C::SH>(buf, a.tw);
  float2* yk = a.y + (size_t)b * rows * a.wc;
  for (int e = threadIdx.x; e < C::LINES * a.wc; e += C::NT) {
    const int l = e / a.wc, j = e - l * a.wc;
    const int q = q0 + l;
    if (q < rows) yk[(size_t)q * a.wc + j] = cmulc(buf[l * C::LS + fft::swz<C::SH>(a.col_freq[j])], a.h[(size_t)q * a.wc + j]);
  }
}

// ---------------------------------------------------------------- K7
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::NT)
k_final_row_inv(SocsArgs a) {
  using C = RowCfg<N>;
  __shared__ float2 buf[C::LINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int rows = a.k * a.wr;
  const int i0 = blockIdx.x * C::LINES;
  for (int e = threadIdx.x; e < C::LINES * N; e += C::NT) {
    const int l = e / N, x = e - l * N;
    buf[l * C::LS + fft::swz<C::SH>(x)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* yk = a.y + (size_t)b * rows * a.wc;
  for (int e = threadIdx.x; e < C::LINES * a.wc; e += C::NT) {
    const int l = e / a.wc, j = e - l * a.wc;
    const int i = i0 + l;
    if (i < a.wr) {
      // eight k per batch of loads, summed in increasing k
      const float2* src = yk + (size_t)i * a.wc + j;
      const size_t ks = (size_t)a.wr * a.wc;
      float2 s = make_float2(0.f, 0.f);
      int k = 0;
      for (; k + 8 <= a.k; k += 8) {
        float2 w[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) w[t] = src[(size_t)(k + t) * ks];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          s.x += w[t].x;
          s.y += w[t].y;
        }
      }
      for (; k < a.k; ++k) {
        const float2 u = src[(size_t)k * ks];
        s.x += u.x;
        s.y += u.y;
      }
      buf[l * C::LS + fft::swz<C::SH>(a.col_freq[j])] = s;
    }
  }
  __syncthreads();
  fft::fft_lines<N, C::LINES, C::LS, C::NT, +1, C::SH>(buf, a.tw);
  float2* p = a.p + (size_t)b * a.wr * N;
  for (int e = threadIdx.x; e < C::LINES * N; e += C::NT) {
    const int l = e / N, x = e - l * N;
    const int i = i0 + l;
    if (i < a.wr) p[(size_t)i * N + x] = buf[l * C::LS + fft::swz<C::SH>(x)];
  }
}

// ---------------------------------------------------------------- pruned rows
// K1, K2, K3, K6 and K7 over the pruned line transforms of fft_rows.hpp: the
// spectra are nonzero (or needed) only at the window frequencies, so the
// inverse transforms start from two taps per butterfly and the forward
// transforms finish with two outputs per butterfly; first and last stages
// read and write HBM directly.  K1 additionally needs N - v in the computed
// set for every window column v (strict upper bound, checked on the host).
template <int N>
__global__ void __launch_bounds__(RowPCfg<N>::K1NT)
k_row_fwd_pruned(SocsArgs a, int mode, const float* base, const float* __restrict__ dir,
                 const float* __restrict__ coef, float* theta_out, float* __restrict__ mask_out) {
  using C = RowPCfg<N>;
  using RF = fft::RowFFT<N, C::LINES, C::LS, C::K1NT, C::SH>;
  constexpr int T1 = RF::T1;
  __shared__ float2 buf[C::LINES * C::LS];
  __shared__ float2 zc[C::LINES * 2 * T1];   // computed frequencies: u < T1 at u, u >= N - T1 at u - (N - 2 T1)
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const size_t tile = (size_t)b * N * N;
  const size_t pt = point_offset(a, b, 1);                            // trial theta
  const size_t pb = mode == 1 ? point_offset(a, b, 0) : tile;         // theta origin (or the mask)
  const int row0 = blockIdx.x * 2 * C::LINES;
  const float cf = (dir != nullptr) ? coef[b] : 0.f;
  const bool use_dir = mode == 1 && dir != nullptr;
  RF::forward_split(
      buf, a.rtw,
      [&](int l, int x) {   // base and direction of the two rows, loads only
        float4 raw = make_float4(0.f, 0.f, 0.f, 0.f);
        const int y = row0 + 2 * l;
        const size_t lg = (size_t)y * N + x, g = tile + lg;
        if (y < N) {
          raw.x = base[pb + lg];
          if (use_dir) raw.z = dir[g];
        }
        if (y + 1 < N) {
          raw.y = base[pb + lg + N];
          if (use_dir) raw.w = dir[g + N];
        }
        return raw;
      },
      [&](int l, int x, float4 raw) {
        float m2[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int y = row0 + 2 * l + h;
          float m = 0.f;
          if (y < N) {
            const size_t lg = (size_t)y * N + x, g = tile + lg;
            float v = h == 0 ? raw.x : raw.y;
            if (mode == 1) {
              if (dir != nullptr) v = v + cf * (h == 0 ? raw.z : raw.w);
              if (theta_out != nullptr) st_nt1(theta_out + pt + lg, v);
              m = 0.5f * (1.0f + cosf(v));
            } else {
              m = v;
            }
            if (mask_out != nullptr) mask_out[g] = m;
          }
          m2[h] = m;
        }
        return make_float2(m2[0], m2[1]);
      },
      [&](int l, int u, float2 z) { zc[l * 2 * T1 + (u < T1 ? u : u - (N - 2 * T1))] = z; });
  __syncthreads();
  // unpack X_a[v] = (Z[v] + conj Z[-v]) / 2, X_b[v] = (Z[v] - conj Z[-v]) / (2i)
  float2* rt = a.rt + (size_t)b * a.wc * N;
  for (int e = threadIdx.x; e < C::LINES * a.wc; e += C::K1NT) {
    const int l = e / a.wc, j = e - l * a.wc;
    const int v = a.col_freq[j];
    const int vn = v == 0 ? 0 : N - v;
    const float2 z = zc[l * 2 * T1 + (v < T1 ? v : v - (N - 2 * T1))];
    const float2 zn = cconj(zc[l * 2 * T1 + (vn < T1 ? vn : vn - (N - 2 * T1))]);
    const float2 xa = make_float2(0.5f * (z.x + zn.x), 0.5f * (z.y + zn.y));
    const float2 xb = make_float2(0.5f * (z.y - zn.y), -0.5f * (z.x - zn.x));
    const int y = row0 + 2 * l;
    if (y < N) rt[(size_t)j * N + y] = xa;
    if (y + 1 < N) rt[(size_t)j * N + y + 1] = xb;
  }
}

// K2: the W_c row-spectrum columns (stored as lines along y) -> mask spectrum
// window rows.  Needs the window rows to fit the pruning (the column flag).
template <int N>
__global__ void __launch_bounds__(RowPCfg<N>::WNT)
k_col_fwd_pruned(SocsArgs a) {
  using C = RowPCfg<N>;
  using RF = fft::RowFFT<N, C::WLINES, C::LS, C::WNT, C::SH>;
  __shared__ float2 buf[C::WLINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const float2* rt = a.rt + (size_t)b * a.wc * N;
  float2* mh = a.mhat + (size_t)b * a.wr * a.wc;
  const int j0 = blockIdx.x * C::WLINES;
  RF::forward(
      buf, a.rtw,
      [&](int l, int y) { return j0 + l < a.wc ? rt[(size_t)(j0 + l) * N + y] : make_float2(0.f, 0.f); },
      [&](int l, int u, float2 z) {
        const int i = a.row_of_u[u];
        if (j0 + l < a.wc && i >= 0) mh[(size_t)i * a.wc + j0 + l] = z;
      });
}

// K3: T_k rows = inverse row transforms of H~_k x M^ over the window columns.
template <int N>
__global__ void __launch_bounds__(RowPCfg<N>::NT)
k_row_inv_pruned(SocsArgs a) {
  using C = RowPCfg<N>;
  using RF = fft::RowFFT<N, C::LINES, C::LS, C::NT, C::SH>;
  __shared__ float2 buf[C::LINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int rows = a.k * a.wr;
  const int q0 = blockIdx.x * C::LINES;
  const float2* mh = a.mhat + (size_t)b * a.wr * a.wc;
  float2* t = a.t + (size_t)b * rows * N;
  RF::inverse(
      buf, a.rtw,
      [&](int l, int u) {
        const int q = q0 + l, j = a.col_of_u[u];
        if (q >= rows || j < 0) return make_float2(0.f, 0.f);
        return cmul(a.h[(size_t)q * a.wc + j], mh[(size_t)(q % a.wr) * a.wc + j]);
      },
      [&](int l, int x, float2 z) {
        if (q0 + l < rows) t[(size_t)(q0 + l) * N + x] = z;
      });
}

// K6: Y_k = conj(H~_k) x forward row transforms of V_k, window columns only.
template <int N>
__global__ void __launch_bounds__(RowPCfg<N>::NT)
k_row_adj_pruned(SocsArgs a) {
  using C = RowPCfg<N>;
  using RF = fft::RowFFT<N, C::LINES, C::LS, C::NT, C::SH>;
  __shared__ float2 buf[C::LINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int rows = a.k * a.wr;
  const int q0 = blockIdx.x * C::LINES;
  const float2* v = a.v + (size_t)b * rows * N;
  float2* yk = a.y + (size_t)b * rows * a.wc;
  RF::forward(
      buf, a.rtw,
      [&](int l, int x) { return q0 + l < rows ? ld_nt2(v + (size_t)(q0 + l) * N + x) : make_float2(0.f, 0.f); },
      [&](int l, int u, float2 z) {
        const int q = q0 + l, j = a.col_of_u[u];
        if (q < rows && j >= 0) yk[(size_t)q * a.wc + j] = cmulc(z, a.h[(size_t)q * a.wc + j]);
      });
}

// K7: P rows = inverse row transforms of sum_k Y_k (k in increasing order).
template <int N>
__global__ void __launch_bounds__(RowPCfg<N>::WNT)
k_final_row_inv_pruned(SocsArgs a) {
  using C = RowPCfg<N>;
  using RF = fft::RowFFT<N, C::WLINES, C::LS, C::WNT, C::SH>;
  __shared__ float2 buf[C::WLINES * C::LS];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int rows = a.k * a.wr;
  const int i0 = blockIdx.x * C::WLINES;
  const float2* yk = a.y + (size_t)b * rows * a.wc;
  float2* p = a.p + (size_t)b * a.wr * N;
  RF::inverse(
      buf, a.rtw,
      [&](int l, int u) {
        const int i = i0 + l, j = a.col_of_u[u];
        float2 s = make_float2(0.f, 0.f);
        if (i >= a.wr || j < 0) return s;
        // loads eight k at a time (one round trip per batch instead of per k),
        // sums in increasing k as before
        const float2* src = yk + (size_t)i * a.wc + j;
        const size_t ks = (size_t)a.wr * a.wc;
        int k = 0;
        for (; k + 8 <= a.k; k += 8) {
          float2 w[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) w[t] = src[(size_t)(k + t) * ks];
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            s.x += w[t].x;
            s.y += w[t].y;
          }
        }
        for (; k < a.k; ++k) {
          const float2 w = src[(size_t)k * ks];
          s.x += w.x;
          s.y += w.y;
        }
        return s;
      },
      [&](int l, int x, float2 z) {
        if (i0 + l < a.wr) p[(size_t)(i0 + l) * N + x] = z;
      });
}

// ---------------------------------------------------------------- K8
template <int N>
__global__ void __launch_bounds__(ColCfg<N>::NT) __attribute__((amdgpu_waves_per_eu(ColCfg<N>::MIN_WAVES_PER_EU, ColCfg<N>::MAX_WAVES_PER_EU)))
k_final_col_inv(SocsArgs a, GradArgs ga) {
  using C = ColCfg<N>;
  __shared__ float2 buf[C::LINES * C::LS];
  __shared__ double red[2 * C::NT / 64];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int x0 = blockIdx.x * C::LINES;
  const size_t tile = (size_t)b * N * N;
  const size_t pt = point_offset(a, b, 1);   // trial theta and its gradient
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    if (e < C::ELEMS) buf[(e % C::LINES) * C::LS + fft::swz<C::SH>(e / C::LINES)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  load_window_columns<N>(buf, a.p + (size_t)b * a.wr * N, a.wr, a.row_freq, x0);
  __syncthreads();
  fft::fft_lines<N, C::LINES, C::LS, C::NT, +1, C::SH>(buf, a.tw);
  double nrm = 0.0, dot = 0.0;
  const int* rect = ga.rect != nullptr ? ga.rect + 4 * b : nullptr;
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    const int c = e % C::LINES, y = e / C::LINES;
    const int x = x0 + c;
    if (e < C::ELEMS && x < N) {
      const size_t lg = (size_t)y * N + x, gi = tile + lg;
      const float gm = ga.scale * buf[c * C::LS + fft::swz<C::SH>(y)].x;
      if (ga.grad_mask != nullptr) ga.grad_mask[gi] = gm;
      if (ga.theta != nullptr) {
        float gt = -0.5f * gm * sinf(ga.theta[pt + lg]);
        if (rect != nullptr && (y < rect[0] || y >= rect[1] || x < rect[2] || x >= rect[3])) gt = 0.f;
        if (ga.grad_theta != nullptr) {
          if (ga.nt_grad_theta)
            st_nt1(ga.grad_theta + pt + lg, gt);
          else
            ga.grad_theta[pt + lg] = gt;
        }
        nrm += (double)gt * (double)gt;
        if (ga.dir != nullptr) dot += (double)gt * (double)ga.dir[gi];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    nrm += __shfl_down(nrm, off, 64);
    dot += __shfl_down(dot, off, 64);
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) {
    red[2 * wave] = nrm;
    red[2 * wave + 1] = dot;
  }
  __syncthreads();
  if (threadIdx.x == 0 && ga.partials != nullptr) {
    double sn = 0.0, sd = 0.0;
    for (int w = 0; w < C::NT / 64; ++w) {
      sn += red[2 * w];
      sd += red[2 * w + 1];
    }
    double* out = ga.partials + ((size_t)b * gridDim.x + blockIdx.x) * 2;
    out[0] = sn;
    out[1] = sd;
  }
}

// ---------------------------------------------------------------- K4+K5 pruned
// Same contract as k_col_image_adj, with the window-pruned column transforms of
// fft_columns.hpp.  Element ownership follows the last inverse stage:
// column q % LINES, row i + r*TM with q = tid + s*NT, i = q / LINES.
template <int N>
__global__ void __launch_bounds__(ColCfg<N>::NT) __attribute__((amdgpu_waves_per_eu(ColCfg<N>::MIN_WAVES_PER_EU, ColCfg<N>::MAX_WAVES_PER_EU)))
k_col_image_adj_pruned(SocsArgs a, ImageArgs im) {
  using C = ColCfg<N>;
  using CF = fft::ColumnFFT<N, C::LINES, C::LS, C::NT, C::SH>;
  __shared__ float2 buf[C::LINES * C::LS];
  __shared__ double red[2 * C::NT / 64];
  __shared__ float2 tws[C::TW_LDS ? N : 1];
  __shared__ float2 mid[CF::MID_ENTRIES > 0 ? CF::MID_ENTRIES : 1];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int x0 = blockIdx.x * C::LINES;
  const size_t tile = (size_t)b * N * N;
  const int rows = a.k * a.wr;
  const float2* t = a.t + (size_t)b * rows * N;
  if constexpr (C::TW_LDS) {
    for (int e = threadIdx.x; e < N; e += C::NT) tws[e] = a.tw[e];
  }
  CF::fill_mid_tables(mid, a.tw);
  __syncthreads();
  const ColTwiddles<N> twd{tws, a.tw, mid};
  float acc[CF::PERM][CF::RM];
  float2 v[CF::PERM][CF::RM];
#pragma unroll
  for (int s = 0; s < CF::PERM; ++s)
#pragma unroll
    for (int r = 0; r < CF::RM; ++r) acc[s][r] = 0.f;
  // ---- pass 1: aerial image
#pragma unroll 1
  for (int k = 0; k < a.k; ++k) {
    CF::inverse_to_regs(buf, t + (size_t)k * a.wr * N, a.row_of_u, x0, twd.get(), v);
#pragma unroll
    for (int s = 0; s < CF::PERM; ++s)
#pragma unroll
      for (int r = 0; r < CF::RM; ++r) acc[s][r] = fmaf(v[s][r].x, v[s][r].x, fmaf(v[s][r].y, v[s][r].y, acc[s][r]));
    __syncthreads();
  }
  // ---- resist, loss, dLdI
  float g[CF::PERM][CF::RM];
  double pe = 0.0, bg = 0.0;
#pragma unroll
  for (int s = 0; s < CF::PERM; ++s) {
    const int q = threadIdx.x + s * C::NT;
    const int line = q % C::LINES, i = q / C::LINES;
    const int x = x0 + line;
    const bool ok = (CF::BM % C::NT == 0 || q < CF::BM) && x < N;
#pragma unroll
    for (int r = 0; r < CF::RM; ++r) {
      g[s][r] = 0.f;
      if (!ok) continue;
      const int y = i + r * CF::TM;
      const size_t gi = tile + (size_t)y * N + x;
      const float iv = acc[s][r];
      const float w = 1.0f / (1.0f + expf(-im.alpha * (iv - im.thr)));
      if (im.aerial != nullptr) im.aerial[gi] = iv;
      if (im.wafer != nullptr) im.wafer[gi] = w;
      if (im.target != nullptr) {
        const float tv = im.target[gi];
        const float rr = w - tv;
        const float pw = im.pe_win != nullptr ? im.pe_win[gi] : 1.0f;
        float drive = im.pe_weight * rr * pw;
        pe += (double)rr * (double)rr * (double)pw;
        if (im.pen_win != nullptr) {
          const float bw = im.pen_win[gi];
          const float ex = fmaxf(w - im.pen_thr, 0.f);
          const float bgm = tv < 0.5f ? 1.f : 0.f;
          bg += (double)bw * (double)bgm * (double)ex * (double)ex;
          drive += im.penalty * bgm * ex * bw;
        }
        g[s][r] = drive * w * (1.0f - w);
        if (im.extra_dldi != nullptr) g[s][r] += im.extra_dldi[gi];
        if (im.dldi != nullptr) im.dldi[gi] = g[s][r];
      }
    }
  }
  {
    for (int off = 32; off > 0; off >>= 1) {
      pe += __shfl_down(pe, off, 64);
      bg += __shfl_down(bg, off, 64);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
      red[2 * wave] = pe;
      red[2 * wave + 1] = bg;
    }
    __syncthreads();
    if (threadIdx.x == 0 && im.partials != nullptr) {
      double sp = 0.0, sb = 0.0;
      for (int w = 0; w < C::NT / 64; ++w) {
        sp += red[2 * w];
        sb += red[2 * w + 1];
      }
      double* out = im.partials + ((size_t)b * gridDim.x + blockIdx.x) * 2;
      out[0] = sp;
      out[1] = sb;
    }
  }
  if (im.target == nullptr || a.v == nullptr) return;
  // ---- pass 2: rebuild E~_k, multiply by dLdI, forward transform to the window rows
  float2* vv = a.v + (size_t)b * rows * N;
#pragma unroll 1
  for (int k = 0; k < a.k; ++k) {
    CF::inverse_to_regs(buf, t + (size_t)k * a.wr * N, a.row_of_u, x0, twd.get(), v);
#pragma unroll
    for (int s = 0; s < CF::PERM; ++s)
#pragma unroll
      for (int r = 0; r < CF::RM; ++r) v[s][r] = make_float2(v[s][r].x * g[s][r], v[s][r].y * g[s][r]);
    __syncthreads();
    CF::forward_from_regs(buf, v, twd.get(), vv + (size_t)k * a.wr * N, a.row_of_u, x0);
    __syncthreads();
  }
}

// ---------------------------------------------------------------- band-limited columns
// The column stage computed from band-limited spectra.  Along a column x the
// field spectrum T_k[:, x] is non-zero only on the W_r contiguous window
// rows (frequencies nu_0 .. nu_0 + W_r - 1), so
//   I[y] = sum_k |E_k[y]|^2 = sum_{|d| < W_r} c[d] e^{2 pi i d y / N},
//   c[d] = sum_k sum_j T_k[j + d] conj(T_k[j])       (c[-d] = conj c[d]),
// and the adjoint column spectra at the window rows are
//   V_k[i] = sum_y g[y] E_k[y] e^{-2 pi i nu_i y / N} = sum_j ghat[i - j] T_k[j],
//   ghat[m] = sum_y g[y] e^{-2 pi i m y / N}          (g real: ghat[-m] = conj ghat[m]).
// Both are length-W_r correlations, evaluated exactly with FFTs of length
// L >= 2 W_r - 1 (no circular wrap-around reaches the kept lags):
//   K4a: c = IFFT_L(sum_k |FFT_L(t_k)|^2) / L, t_k = T_k[:, x] zero-padded;
//   K4b: I = IDFT_N(c) (one column transform), resist, loss, dL/dI = g,
//        ghat = DFT_N(g) at m in [0, W_r);
//   K5:  V_k = IFFT_L(FFT_L(t_k) . FFT_L(h) / L)[W_r - 1 + i], h[p] = ghat[p - W_r + 1].
// Per column that is 2K + 2 transforms of length L and two of length N,
// against 3K transforms of length N in the fused kernel.
// The length-L transforms run register-resident (fft_regs.hpp): thread
// (column x0 + line, j) owns positions j + r L/8 of its column, loads them
// from the window rows of T_k and stores them to the window rows of V_k
// directly, and exchanges through LDS only between passes.
template <int L>
struct CorrCfg {
  static constexpr int LINES = 4096 / L;                       // columns per block
  using F = fft::RegFFT<L, LINES>;
  static constexpr int NT = F::NT;                             // 512
  static constexpr int Q = F::Q;                               // threads per column
  // waves per SIMD the register budget is sized for: two blocks per CU where
  // the two exchange buffers fit twice in LDS (L >= 256), else one
  static constexpr int WAVES = 2 * F::BUF * 8 * 2 <= 160 * 1024 ? 4 : 2;
};

// Workgroups are dispatched round-robin over the 8 XCDs (linear id mod 8).
// When the grid's x extent is a multiple of 8, block x runs on XCD x mod 8;
// giving XCD c the contiguous column blocks [c nb/8, (c+1) nb/8) keeps the
// 128-byte lines shared by neighbouring column blocks in one L2.
__device__ __forceinline__ int xcd_column_block(int bx, int nb) {
  return nb % 8 == 0 ? (bx % 8) * (nb / 8) + bx / 8 : bx;
}

constexpr int kCorrMaxGroups = 8;   // k groups of the power spectra (partials summed by k_corr_finish)

bool corr_len_supported(int l) { return l == 64 || l == 128 || l == 256 || l == 512 || l == 1024; }

// Loads positions j + r Q, r < 4, of a column's window rows (zero beyond W_r
// or n).  L >= 2 W_r - 1 gives W_r <= L/2 = 4 Q: positions with r >= 4 are
// always zero, and the transforms take them as such (RegFFT HALF).
template <int L>
__device__ __forceinline__ void corr_load_col(float2 (&v)[8], const float2* __restrict__ src, int wr, int n,
                                              bool xin) {
  using C = CorrCfg<L>;
  const int j = threadIdx.x / C::LINES;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int p = j + r * C::Q;
    v[r] = (p < wr && xin) ? src[(size_t)p * n] : make_float2(0.f, 0.f);
  }
}

// K4a (1/2): partial power spectra S_z[m] = sum_{k in group z} |FFT_L(t_k)[m]|^2
// of each column, groups of k over blockIdx.z (blocks enough to fill the chip
// at batch 1); S layout [b][z][m][x].
template <int L>
__global__ void __launch_bounds__(CorrCfg<L>::NT) __attribute__((amdgpu_waves_per_eu(CorrCfg<L>::WAVES)))
k_corr_power(SocsArgs a, float* __restrict__ spart) {
  using C = CorrCfg<L>;
  using F = typename C::F;
  __shared__ float2 lds[2 * F::BUF];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int n = a.n, wr = a.wr;
  const int line = threadIdx.x % C::LINES, j = threadIdx.x / C::LINES;
  const int x = xcd_column_block(blockIdx.x, gridDim.x) * C::LINES + line;
  const bool xin = x < n;
  const int groups = gridDim.z, z = blockIdx.z;
  const int k0 = (int)((long long)a.k * z / groups), k1 = (int)((long long)a.k * (z + 1) / groups);
  float2 tw[F::TWN];
  F::load_twiddles(a.twl, tw, j);
  const float2* t = a.t + (size_t)b * a.k * wr * n + (xin ? x : 0);
  float acc[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) acc[r] = 0.f;
  float2 nx[8], v[8];
  if (k0 < k1) corr_load_col<L>(nx, t + (size_t)k0 * wr * n, wr, n, xin);
  int ex = 0;
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = nx[r];
    if (k + 1 < k1) corr_load_col<L>(nx, t + (size_t)(k + 1) * wr * n, wr, n, xin);
    F::template transform<-1, true>(v, lds, tw, line, j, ex);
#pragma unroll
    for (int r = 0; r < 8; ++r) acc[r] = fmaf(v[r].x, v[r].x, fmaf(v[r].y, v[r].y, acc[r]));
  }
  if (!xin) return;
  float* out = spart + ((size_t)b * groups + z) * L * n + x;
#pragma unroll
  for (int r = 0; r < 8; ++r) out[(size_t)(j + r * C::Q) * n] = acc[r];
}

// K4a (2/2): S = sum of the groups' partials (increasing z), c = IFFT_L(S) / L on d in [0, W_r)
template <int L>
__global__ void __launch_bounds__(CorrCfg<L>::NT)
k_corr_finish(SocsArgs a, const float* __restrict__ spart, int groups) {
  using C = CorrCfg<L>;
  using F = typename C::F;
  __shared__ float2 lds[2 * F::BUF];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int n = a.n, wr = a.wr;
  const int line = threadIdx.x % C::LINES, j = threadIdx.x / C::LINES;
  const int x = xcd_column_block(blockIdx.x, gridDim.x) * C::LINES + line;
  const bool xin = x < n;
  float2 tw[F::TWN];
  F::load_twiddles(a.twl, tw, j);
  const float* sp = spart + (size_t)b * groups * L * n + (xin ? x : 0);
  float2 v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = make_float2(0.f, 0.f);
  // groups <= kCorrMaxGroups: the unrolled loop keeps every partial's load in flight
#pragma unroll
  for (int z = 0; z < kCorrMaxGroups; ++z) {
    if (z < groups && xin) {
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r].x += sp[((size_t)z * L + j + r * C::Q) * n];
    }
  }
  int ex = 0;
  F::template transform<+1>(v, lds, tw, line, j, ex);
  if (!xin) return;
  const float inv = 1.f / (float)L;
  float2* out = a.acorr + (size_t)b * wr * n + x;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int p = j + r * C::Q;
    if (p < wr) out[(size_t)p * n] = make_float2(v[r].x * inv, v[r].y * inv);
  }
}

// K4b: I = IDFT_N(c) per column, resist / losses / dL/dI, ghat = DFT_N(g) on [0, W_r)
template <int N>
__global__ void __launch_bounds__(ImgCfg<N>::NT) __attribute__((amdgpu_waves_per_eu(ImgCfg<N>::WAVES, 8)))
k_col_image_corr(SocsArgs a, ImageArgs im) {
  using C = ImgCfg<N>;
  constexpr int TW = C::TW;
  __shared__ float2 buf[C::LINES * C::LS];
  __shared__ double red[2 * C::NT / 64];
  __shared__ float2 tw[TW > 0 ? TW : 1];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int cb = xcd_column_block(blockIdx.x, gridDim.x);   // XCD-contiguous column blocks
  const int x0 = cb * C::LINES;
  // the inverse-order half of the plan's stage tables (row_stage_twiddles(N))
  for (int e = threadIdx.x; e < TW; e += C::NT) tw[e] = a.rtw[e];
  const size_t tile = (size_t)b * N * N;
  const float2* ac = a.acorr + (size_t)b * a.wr * N;
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    if (e < C::ELEMS) {
      const int line = e % C::LINES, d = e / C::LINES;
      const int x = x0 + line;
      float2 v = make_float2(0.f, 0.f);
      if (x < N) {
        if (d < a.wr)
          v = ac[(size_t)d * N + x];
        else if (N - d < a.wr)
          v = cconj(ac[(size_t)(N - d) * N + x]);
      }
      buf[line * C::LS + fft::swz<C::SH>(d)] = v;
    }
  }
  __syncthreads();
  fft::fft_lines_staged<N, C::LINES, C::LS, C::NT, +1, C::SH>(buf, tw);
  // each thread replaces its own elements of I by (dL/dI, 0), the input of the forward transform
  double pe = 0.0, bg = 0.0;
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    const int line = e % C::LINES, y = e / C::LINES;
    const int x = x0 + line;
    if (e < C::ELEMS) {
      float2* q = buf + line * C::LS + fft::swz<C::SH>(y);
      const float gv = x < N ? image_pixel(im, tile + (size_t)y * N + x, q->x, pe, bg) : 0.f;
      *q = make_float2(gv, 0.f);
    }
  }
  {
    for (int off = 32; off > 0; off >>= 1) {
      pe += __shfl_down(pe, off, 64);
      bg += __shfl_down(bg, off, 64);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
      red[2 * wave] = pe;
      red[2 * wave + 1] = bg;
    }
    __syncthreads();
    if (threadIdx.x == 0 && im.partials != nullptr) {
      double sp = 0.0, sb = 0.0;
      for (int w = 0; w < C::NT / 64; ++w) {
        sp += red[2 * w];
        sb += red[2 * w + 1];
      }
      double* out = im.partials + ((size_t)b * gridDim.x + cb) * 2;
      out[0] = sp;
      out[1] = sb;
    }
  }
  if (im.target == nullptr || a.v == nullptr) return;
  // the reduction's barrier orders the (dL/dI, 0) writes before the transform
  fft::fft_lines_staged<N, C::LINES, C::LS, C::NT, -1, C::SH>(buf, tw);
  float2* gh = a.ghat + (size_t)b * a.wr * N;
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    const int line = e % C::LINES, m = e / C::LINES;
    const int x = x0 + line;
    if (e < C::ELEMS && m < a.wr && x < N) gh[(size_t)m * N + x] = buf[line * C::LS + fft::swz<C::SH>(m)];
  }
}

// K4b with two columns per complex line.  Both transforms of a column are
// real on one side: I = IDFT_N(c) is real (c Hermitian), g is real.  So the
// columns x0 + 2l and x0 + 2l + 1 share line l: Z = c_A + i c_B (c[0] taken
// real) gives IDFT(Z) = I_A + i I_B, the line then holds g_A + i g_B, and
// Y = DFT(g_A + i g_B) splits into ghat_A[m] = (Y[m] + conj Y[-m]) / 2 and
// ghat_B[m] = (Y[m] - conj Y[-m]) / (2i).  Half the transforms and half the
// LDS of k_col_image_corr for the same column block.
// Used from 768 up except for the powers of two from 1024: image stage per
// batch of 32, 0.350 against 0.387 ms at 896^2, 0.492 against 0.527 at
// 1152^2 (batched tiles 1.3 % faster in all).  At 640^2 it is faster per
// batch of 32 (0.215 against 0.228 ms) but the 256-Micro full-chip update
// ran 3 % slower with it (99.1-99.8 against 101.9-102.7 updates/s), so 640
// keeps the one-column kernel.  At 2048^2 this LDS-resident pairing stayed
// latency bound (29-30 us with 4 or 8 columns per block against 28 us for
// k_col_image_corr); the same pairing on register-resident transforms
// (k_col_image_reg2) is the one that runs there.
template <int N>
struct ImgPairCfg {
  using IC = ImgCfg<N>;
  static constexpr bool USE = IC::LINES % 2 == 0 && N >= 768 && !(N >= 1024 && (N & (N - 1)) == 0);
  static constexpr int COLS = IC::LINES;                      // columns per block (the image kernel's blocks)
  static constexpr int PL = COLS / 2;                         // complex lines
  static constexpr int NT = IC::NT / 2 >= 64 ? IC::NT / 2 : 64;
  static constexpr int SH = IC::SH;
  static constexpr int LS = fft::padded_stride<SH>(N);
  static constexpr int ELEMS = PL * N;
  static constexpr int PER = (ELEMS + NT - 1) / NT;
  static constexpr int TW = fft::staged_entries(N);
  static constexpr size_t LDS_BYTES = (size_t)PL * LS * 8 + (size_t)(TW > 0 ? TW : 1) * 8 + (size_t)NT / 64 * 16;
  static constexpr int BLOCKS_PER_CU = 160 * 1024 / LDS_BYTES > 0 ? (int)(160 * 1024 / LDS_BYTES) : 1;
  static constexpr int WAVES = BLOCKS_PER_CU * NT / 256 > 8 ? 8 : (BLOCKS_PER_CU * NT / 256 > 0 ? BLOCKS_PER_CU * NT / 256 : 1);
};

template <int N>
__global__ void __launch_bounds__(ImgPairCfg<N>::NT) __attribute__((amdgpu_waves_per_eu(ImgPairCfg<N>::WAVES, 8)))
k_col_image_corr2(SocsArgs a, ImageArgs im) {
  using C = ImgPairCfg<N>;
  constexpr int TW = C::TW;
  __shared__ float2 buf[C::PL * C::LS];
  __shared__ double red[2 * C::NT / 64];
  __shared__ float2 tw[TW > 0 ? TW : 1];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int cb = xcd_column_block(blockIdx.x, gridDim.x);   // XCD-contiguous column blocks
  const int x0 = cb * C::COLS;
  for (int e = threadIdx.x; e < TW; e += C::NT) tw[e] = a.rtw[e];
  const size_t tile = (size_t)b * N * N;
  const float2* ac = a.acorr + (size_t)b * a.wr * N;
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    if (e < C::ELEMS) {
      const int l = e % C::PL, d = e / C::PL;
      const int xa = x0 + 2 * l, xb = xa + 1;
      float2 ca = make_float2(0.f, 0.f), cbv = make_float2(0.f, 0.f);
      bool conj_half = false;
      int src = -1;
      if (d < a.wr) {
        src = d;
      } else if (N - d < a.wr) {
        src = N - d;
        conj_half = true;
      }
      if (src >= 0) {
        if (xa < N) ca = ac[(size_t)src * N + xa];
        if (xb < N) cbv = ac[(size_t)src * N + xb];
        if (src == 0) ca.y = cbv.y = 0.f;   // c[0] is real: keep I_A and I_B apart
        if (conj_half) {
          ca.y = -ca.y;
          cbv.y = -cbv.y;
        }
      }
      // Z = c_A + i c_B
      buf[l * C::LS + fft::swz<C::SH>(d)] = make_float2(ca.x - cbv.y, ca.y + cbv.x);
    }
  }
  __syncthreads();
  fft::fft_lines_staged<N, C::PL, C::LS, C::NT, +1, C::SH>(buf, tw);
  double pe = 0.0, bg = 0.0;
#pragma unroll
  for (int s = 0; s < C::PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    const int l = e % C::PL, y = e / C::PL;
    const int xa = x0 + 2 * l, xb = xa + 1;
    if (e < C::ELEMS) {
      float2* q = buf + l * C::LS + fft::swz<C::SH>(y);
      const float2 iv = *q;
      const float ga = xa < N ? image_pixel(im, tile + (size_t)y * N + xa, iv.x, pe, bg) : 0.f;
      const float gb = xb < N ? image_pixel(im, tile + (size_t)y * N + xb, iv.y, pe, bg) : 0.f;
      *q = make_float2(ga, gb);
    }
  }
  {
    for (int off = 32; off > 0; off >>= 1) {
      pe += __shfl_down(pe, off, 64);
      bg += __shfl_down(bg, off, 64);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
      red[2 * wave] = pe;
      red[2 * wave + 1] = bg;
    }
    __syncthreads();
    if (threadIdx.x == 0 && im.partials != nullptr) {
      double sp = 0.0, sb = 0.0;
      for (int w = 0; w < C::NT / 64; ++w) {
        sp += red[2 * w];
        sb += red[2 * w + 1];
      }
      double* out = im.partials + ((size_t)b * gridDim.x + cb) * 2;
      out[0] = sp;
      out[1] = sb;
    }
  }
  if (im.target == nullptr || a.v == nullptr) return;
  // the reduction's barrier orders the (g_A, g_B) writes before the transform
  fft::fft_lines_staged<N, C::PL, C::LS, C::NT, -1, C::SH>(buf, tw);
  float2* gh = a.ghat + (size_t)b * a.wr * N;
  for (int e = threadIdx.x; e < C::PL * a.wr; e += C::NT) {
    const int l = e % C::PL, m = e / C::PL;
    const int xa = x0 + 2 * l, xb = xa + 1;
    const float2 y = buf[l * C::LS + fft::swz<C::SH>(m)];
    const float2 yc = cconj(buf[l * C::LS + fft::swz<C::SH>(m == 0 ? 0 : N - m)]);
    if (xa < N) gh[(size_t)m * N + xa] = make_float2(0.5f * (y.x + yc.x), 0.5f * (y.y + yc.y));
    if (xb < N) gh[(size_t)m * N + xb] = make_float2(0.5f * (y.y - yc.y), -0.5f * (y.x - yc.x));
  }
}

// K4b on register-resident transforms: thread j of a column holds the eight
// elements j + r N/8 through both transforms (fft_regs.hpp), so the image,
// the resist and dL/dI are formed in registers between them and only the
// transforms' inner passes exchange through LDS (single-barrier exchanges:
// two per transform at 1024 = 8*4*4*8, three at 2048 = 8*4*8*8, against the
// LDS-resident kernel's three passes of two barriers each).  LINES lines
// per block.  Measured image kernel (kernel trace, K = 24, one tile):
// * 1024^2: k_col_image_reg<1024, 4> 9.1 us, the paired k_col_image_reg2
//   10.3, k_col_image_corr 17.3;
// * 2048^2: k_col_image_reg2<2048, 2> 21.5 us (1 / 4 pairs per block: 24.1 /
//   20.2; batch of 8: 16.0 us per tile against 16.6 / 18.9), the unpaired
//   k_col_image_reg 28.4, k_col_image_corr 26.5 (batch of 8: 26.7 per
//   tile).  Loading the next column group during the current one's
//   transforms spilled registers and took 45-55 us.
template <int N, int LINES>
struct ImgRegCfg {
  using F = fft::RegFFT<N, LINES>;
  static constexpr int NT = F::NT;
  static constexpr int Q = F::Q;
};
// 1: k_col_image_reg (4 columns per block), 2: k_col_image_reg2 (2 column
// pairs per block), 0: the LDS-resident kernels
template <int N>
constexpr int img_reg_kind() { return N == 1024 ? 1 : (N == 2048 ? 2 : 0); }
constexpr int kImgRegLines = 4, kImgRegPairLines = 2;

// c[d] and the target of column x at d, y = j + r N/8 for r < 8.  A group r
// that can hold window entries (a block-uniform test) loads for every lane,
// row 0 standing in where the entry is zero, so that all its loads are in
// flight together.
template <int N, int Q>
__device__ __forceinline__ void img_reg_load(const SocsArgs& a, const ImageArgs& im, int b, int x, int j,
                                             float2 (&v)[8], float (&tv)[8]) {
  const int wr = a.wr;
  const float2* ac = a.acorr + (size_t)b * wr * N + x;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    v[r] = make_float2(0.f, 0.f);
    if (r * Q < wr || N - ((r + 1) * Q - 1) < wr) {
      const int d = j + r * Q;
      const bool lo = d < wr, hi = !lo && N - d < wr;
      const float2 c = ac[(size_t)(lo ? d : (hi ? N - d : 0)) * N];
      v[r] = lo ? c : (hi ? cconj(c) : make_float2(0.f, 0.f));
    }
  }
  const size_t col = (size_t)b * N * N + x;
#pragma unroll
  for (int r = 0; r < 8; ++r) tv[r] = im.target != nullptr ? im.target[col + (size_t)(j + r * Q) * N] : 0.f;
}

template <int N, int LINES>
__global__ void __launch_bounds__((ImgRegCfg<N, LINES>::NT)) __attribute__((amdgpu_waves_per_eu(4)))
k_col_image_reg(SocsArgs a, ImageArgs im) {
  using C = ImgRegCfg<N, LINES>;
  using F = typename C::F;
  static_assert(N % LINES == 0, "whole column blocks");
  __shared__ float2 lds[2 * F::BUF];
  __shared__ double red[2 * C::NT / 64];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int line = threadIdx.x % LINES, j = threadIdx.x / LINES;
  const int cb = xcd_column_block(blockIdx.x, gridDim.x);   // XCD-contiguous column blocks
  const int x = cb * LINES + line;
  float2 tw[F::TWN];
  F::load_twiddles(a.twi, tw, j);
  const size_t tile = (size_t)b * N * N;
  const int wr = a.wr;
  float2 v[8];
  float tv[8];
  img_reg_load<N, C::Q>(a, im, b, x, j, v, tv);
  int ex = 0;
  F::template transform<+1>(v, lds, tw, line, j, ex);
  // the image at rows j + r N/8: resist, losses, dL/dI (real input of the forward transform)
  double pe = 0.0, bg = 0.0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int y = j + r * C::Q;
    const float gv = image_pixel_t(im, tile + (size_t)y * N + x, v[r].x, tv[r], pe, bg);
    v[r] = make_float2(gv, 0.f);
  }
  {
    for (int off = 32; off > 0; off >>= 1) {
      pe += __shfl_down(pe, off, 64);
      bg += __shfl_down(bg, off, 64);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
      red[2 * wave] = pe;
      red[2 * wave + 1] = bg;
    }
    fft::lds_barrier();
    if (threadIdx.x == 0 && im.partials != nullptr) {
      double sp = 0.0, sb = 0.0;
      for (int w = 0; w < C::NT / 64; ++w) {
        sp += red[2 * w];
        sb += red[2 * w + 1];
      }
      double* out = im.partials + ((size_t)b * gridDim.x + cb) * 2;
      out[0] = sp;
      out[1] = sb;
    }
  }
  if (im.target == nullptr || a.v == nullptr) return;
  F::template transform<-1>(v, lds, tw, line, j, ex);
  // ghat[m] = DFT_N(g)[m] on the window rows m < W_r
  float2* gh = a.ghat + (size_t)b * wr * N + x;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int m = j + r * C::Q;
    if (m < wr) gh[(size_t)m * N] = v[r];
  }
}

// k_col_image_reg with two real columns per complex line (the pairing of
// k_col_image_corr2): line l of a block holds Z = c_A + i c_B for columns
// xa = x0 + 2l and xb = xa + 1 (c[0] taken real), so its inverse is
// I_A + i I_B and its forward transform of g_A + i g_B splits into
// ghat_A[m] = (Y[m] + conj Y[-m]) / 2, ghat_B[m] = (Y[m] - conj Y[-m]) / (2i).
// Y[-m] sits with another thread of the line (Q - j, or thread 0), so the
// forward transform's outputs go through LDS once more.  LINES complex lines
// (2 LINES columns) per block.
template <int N, int LINES>
__global__ void __launch_bounds__((ImgRegCfg<N, LINES>::NT)) __attribute__((amdgpu_waves_per_eu(4)))
k_col_image_reg2(SocsArgs a, ImageArgs im) {
  using C = ImgRegCfg<N, LINES>;
  using F = typename C::F;
  static_assert(N % (2 * LINES) == 0, "whole column blocks");
  constexpr int Q = C::Q;
  __shared__ float2 lds[2 * F::BUF];
  __shared__ double red[2 * C::NT / 64];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int line = threadIdx.x % LINES, j = threadIdx.x / LINES;
  const int cb = xcd_column_block(blockIdx.x, gridDim.x);   // XCD-contiguous column blocks
  const int xa = cb * 2 * LINES + 2 * line, xb = xa + 1;
  float2 tw[F::TWN];
  F::load_twiddles(a.twi, tw, j);
  const size_t tile = (size_t)b * N * N;
  const int wr = a.wr;
  // Z[d] = c_A[d] + i c_B[d] at d = j + r Q (loads as img_reg_load, both columns)
  float2 v[8];
  const float2* ac = a.acorr + (size_t)b * wr * N;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    v[r] = make_float2(0.f, 0.f);
    if (r * Q < wr || N - ((r + 1) * Q - 1) < wr) {
      const int d = j + r * Q;
      const bool lo = d < wr, hi = !lo && N - d < wr;
      const size_t src = (size_t)(lo ? d : (hi ? N - d : 0)) * N;
      float2 ca = ac[src + xa], cbv = ac[src + xb];
      if (d == 0) ca.y = cbv.y = 0.f;   // c[0] is real: keep I_A and I_B apart
      if (hi) {
        ca.y = -ca.y;
        cbv.y = -cbv.y;
      }
      if (lo || hi) v[r] = make_float2(ca.x - cbv.y, ca.y + cbv.x);
    }
  }
  float ta[8], tb[8];
  const size_t col = tile + xa;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const size_t g = col + (size_t)(j + r * Q) * N;
    ta[r] = im.target != nullptr ? im.target[g] : 0.f;
    tb[r] = im.target != nullptr ? im.target[g + 1] : 0.f;
  }
  int ex = 0;
  F::template transform<+1>(v, lds, tw, line, j, ex);
  double pe = 0.0, bg = 0.0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const size_t g = col + (size_t)(j + r * Q) * N;
    const float ga = image_pixel_t(im, g, v[r].x, ta[r], pe, bg);
    const float gb = image_pixel_t(im, g + 1, v[r].y, tb[r], pe, bg);
    v[r] = make_float2(ga, gb);
  }
  {
    for (int off = 32; off > 0; off >>= 1) {
      pe += __shfl_down(pe, off, 64);
      bg += __shfl_down(bg, off, 64);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
      red[2 * wave] = pe;
      red[2 * wave + 1] = bg;
    }
    fft::lds_barrier();
    if (threadIdx.x == 0 && im.partials != nullptr) {
      double sp = 0.0, sb = 0.0;
      for (int w = 0; w < C::NT / 64; ++w) {
        sp += red[2 * w];
        sb += red[2 * w + 1];
      }
      double* out = im.partials + ((size_t)b * gridDim.x + cb) * 2;
      out[0] = sp;
      out[1] = sb;
    }
  }
  if (im.target == nullptr || a.v == nullptr) return;
  F::template transform<-1>(v, lds, tw, line, j, ex);
  // Y at positions j + r Q into the next exchange buffer, then each window row
  // m < W_r reads its Y[-m] back
  float2* yb = lds + (ex & 1) * F::BUF + line * F::LS;
#pragma unroll
  for (int r = 0; r < 8; ++r) yb[fft::swz<F::SH>(j + r * Q)] = v[r];
  fft::lds_barrier();
  float2* gh = a.ghat + (size_t)b * wr * N;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int m = j + r * Q;
    if (r * Q < wr && m < wr) {
      const float2 y = v[r];
      const float2 yc = cconj(yb[fft::swz<F::SH>(m == 0 ? 0 : N - m)]);
      gh[(size_t)m * N + xa] = make_float2(0.5f * (y.x + yc.x), 0.5f * (y.y + yc.y));
      gh[(size_t)m * N + xb] = make_float2(0.5f * (y.y - yc.y), -0.5f * (y.x - yc.x));
    }
  }
}

// K5: V_k at the window rows from the band-limited ghat
template <int L>
__global__ void __launch_bounds__(CorrCfg<L>::NT) __attribute__((amdgpu_waves_per_eu(CorrCfg<L>::WAVES)))
k_corr_adjoint(SocsArgs a) {
  using C = CorrCfg<L>;
  using F = typename C::F;
  __shared__ float2 lds[2 * F::BUF];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int n = a.n, wr = a.wr;
  const int line = threadIdx.x % C::LINES, j = threadIdx.x / C::LINES;
  const int x = xcd_column_block(blockIdx.x, gridDim.x) * C::LINES + line;
  const bool xin = x < n;
  const int xs = xin ? x : 0;
  const int groups = gridDim.z, z = blockIdx.z;
  const int k0 = (int)((long long)a.k * z / groups), k1 = (int)((long long)a.k * (z + 1) / groups);
  float2 tw[F::TWN];
  F::load_twiddles(a.twl, tw, j);
  // h[p] = ghat[p - (W_r - 1)] on p in [0, 2 W_r - 2], ghat[-m] = conj ghat[m]
  const float2* gh = a.ghat + (size_t)b * wr * n + xs;
  float2 hh[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int p = j + r * C::Q, m = p - (wr - 1);
    float2 v = make_float2(0.f, 0.f);
    if (xin && p <= 2 * (wr - 1)) v = m >= 0 ? gh[(size_t)m * n] : cconj(gh[(size_t)(-m) * n]);
    hh[r] = v;
  }
  int ex = 0;
  F::template transform<-1>(hh, lds, tw, line, j, ex);
  const float inv = 1.f / (float)L;
#pragma unroll
  for (int r = 0; r < 8; ++r) hh[r] = make_float2(hh[r].x * inv, hh[r].y * inv);
  const float2* t = a.t + (size_t)b * a.k * wr * n + xs;
  float2* vv = a.v + (size_t)b * a.k * wr * n + xs;
  float2 nx[8], v[8];
  if (k0 < k1) corr_load_col<L>(nx, t + (size_t)k0 * wr * n, wr, n, xin);
#pragma unroll 1
  for (int k = k0; k < k1; ++k) {
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = nx[r];
    if (k + 1 < k1) corr_load_col<L>(nx, t + (size_t)(k + 1) * wr * n, wr, n, xin);
    F::template transform<-1, true>(v, lds, tw, line, j, ex);
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = cmul(v[r], hh[r]);
    F::template transform<+1>(v, lds, tw, line, j, ex);
    if (xin) {
      float2* vk = vv + (size_t)k * wr * n;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int i = j + r * C::Q - (wr - 1);
        if (i >= 0 && i < wr) vk[(size_t)i * n] = v[r];
      }
    }
  }
}

// Groups of k per column block (k in [k z / g, k (z+1) / g) for group z).
// More groups give more blocks but repeat the transform of h once per block
// and add partial spectra; g in {1, 2, 4, 8} minimises the modelled time
//   rounds(g) * (ceil(k / g) + 1) + g / 4,
// rounds = blocks * g / resident blocks, rounded up below two rounds (a
// partly filled last round costs a whole one) and fractional above.  Resident
// blocks: two per CU when both exchange buffers fit twice in LDS (L >= 256).
static int corr_groups(int n, int l, int k, int batch) {
  static const int cus = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    return v;
  }();
  const int lines = 4096 / l;
  const double slots = (double)cus * (l >= 256 ? 2 : 1);
  const double blocks = (double)((n + lines - 1) / lines) * batch;
  int best = 1;
  double best_cost = 0.0;
  for (int g = 1; g <= kCorrMaxGroups && g <= k; g *= 2) {
    double rounds = blocks * g / slots;
    if (rounds < 2.0) rounds = std::ceil(rounds);
    const double cost = rounds * ((k + g - 1) / g + 1) + g / 4.0;
    if (g == 1 || cost < best_cost) {
      best = g;
      best_cost = cost;
    }
  }
  return best;
}

// The power spectra are summed over the k groups of a single tile (batch 1)
// whatever the batch, so a tile's image does not depend on the tiles it is
// batched with: the FP32 partial sums depend on the grouping.  The adjoint's
// groups only split the independent V_k and follow the batch.
static int corr_power_groups(int n, int l, int k) {
  // at most 4: each group adds a partial spectrum per column to write and sum
  // (640^2 alone: 0.113 ms per eval with 4 groups against 0.114 with the
  // model's 8; 896^2, batch 32: 1.277 against 1.333 ms)
  const int g = corr_groups(n, l, k, 1);
  return g < 4 ? g : 4;
}

template <int L>
static void launch_corr(int stage, const SocsArgs& a, int batch, float* spart, hipStream_t s) {
  using C = CorrCfg<L>;
  const int cols = (a.n + C::LINES - 1) / C::LINES;
  if (stage == 0) {
    const int groups = corr_power_groups(a.n, L, a.k);
    hipLaunchKernelGGL(k_corr_power<L>, dim3(cols, batch, groups), dim3(C::NT), 0, s, a, spart);
    hipLaunchKernelGGL(k_corr_finish<L>, dim3(cols, batch), dim3(C::NT), 0, s, a, (const float*)spart, groups);
  } else {
    hipLaunchKernelGGL(k_corr_adjoint<L>, dim3(cols, batch, corr_groups(a.n, L, a.k, batch)), dim3(C::NT), 0, s, a);
  }
}

// stage 0: K4a (power spectra and autocorrelation); stage 1: K5
void launch_corr_any(int stage, const SocsArgs& a, int batch, float* spart, hipStream_t s) {
  switch (a.corr_len) {
    case 64: launch_corr<64>(stage, a, batch, spart, s); break;
    case 128: launch_corr<128>(stage, a, batch, spart, s); break;
    case 256: launch_corr<256>(stage, a, batch, spart, s); break;
    case 512: launch_corr<512>(stage, a, batch, spart, s); break;
    case 1024: launch_corr<1024>(stage, a, batch, spart, s); break;
    default: break;
  }
}

size_t corr_power_bytes(int n, int l, int k, int batch) {
  return (size_t)batch * corr_power_groups(n, l, k) * l * n * sizeof(float);
}

// ---------------------------------------------------------------- K8 pruned
template <int N>
__global__ void __launch_bounds__(ColCfg<N>::NT) __attribute__((amdgpu_waves_per_eu(ColCfg<N>::MIN_WAVES_PER_EU, ColCfg<N>::MAX_WAVES_PER_EU)))
k_final_col_inv_pruned(SocsArgs a, GradArgs ga) {
  using C = ColCfg<N>;
  using CF = fft::ColumnFFT<N, C::LINES, C::LS, C::NT, C::SH>;
  __shared__ float2 buf[C::LINES * C::LS];
  __shared__ double red[2 * C::NT / 64];
  __shared__ float2 mid[CF::MID_ENTRIES > 0 ? CF::MID_ENTRIES : 1];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int cb = xcd_column_block(blockIdx.x, gridDim.x);   // XCD-contiguous column blocks
  const int x0 = cb * C::LINES;
  const size_t tile = (size_t)b * N * N;
  const size_t pt = point_offset(a, b, 1);   // trial theta and its gradient
  // theta at this thread's output elements, all loads in flight at once
  // (inside the epilogue, sinf's branches left one round trip per element)
  float th[CF::PERM][CF::RM];
#pragma unroll
  for (int s = 0; s < CF::PERM; ++s) {
    const int q = threadIdx.x + s * C::NT;
    const int x = x0 + q % C::LINES, i = q / C::LINES;
    const bool ok = (CF::BM % C::NT == 0 || q < CF::BM) && x < N;
#pragma unroll
    for (int r = 0; r < CF::RM; ++r) {
      const size_t lg = (size_t)(i + r * CF::TM) * N + x;
      th[s][r] = ok && ga.theta != nullptr ? ga.theta[pt + lg] : 0.f;
    }
  }
  CF::fill_mid_tables(mid, a.tw);
  __syncthreads();
  float2 v[CF::PERM][CF::RM];
  CF::inverse_to_regs(buf, a.p + (size_t)b * a.wr * N, a.row_of_u, x0, fft::ColTw{a.tw, mid}, v);
  double nrm = 0.0, dot = 0.0;
  const int* rect = ga.rect != nullptr ? ga.rect + 4 * b : nullptr;
#pragma unroll
  for (int s = 0; s < CF::PERM; ++s) {
    const int q = threadIdx.x + s * C::NT;
    const int line = q % C::LINES, i = q / C::LINES;
    const int x = x0 + line;
    if (!((CF::BM % C::NT == 0 || q < CF::BM) && x < N)) continue;
#pragma unroll
    for (int r = 0; r < CF::RM; ++r) {
      const int y = i + r * CF::TM;
      const size_t lg = (size_t)y * N + x, gi = tile + lg;
      const float gm = ga.scale * v[s][r].x;
      if (ga.grad_mask != nullptr) ga.grad_mask[gi] = gm;
      if (ga.theta != nullptr) {
        float gt = -0.5f * gm * sinf(th[s][r]);
        if (rect != nullptr && (y < rect[0] || y >= rect[1] || x < rect[2] || x >= rect[3])) gt = 0.f;
        if (ga.grad_theta != nullptr) {
          if (ga.nt_grad_theta)
            st_nt1(ga.grad_theta + pt + lg, gt);
          else
            ga.grad_theta[pt + lg] = gt;
        }
        nrm += (double)gt * (double)gt;
        th[s][r] = gt;
      }
    }
  }
  // G_theta . dir in a second pass without sinf's branches, so that the
  // direction's loads are in flight together (same summation order)
  if (ga.theta != nullptr && ga.dir != nullptr) {
#pragma unroll
    for (int s = 0; s < CF::PERM; ++s) {
      const int q = threadIdx.x + s * C::NT;
      const int x = x0 + q % C::LINES, i = q / C::LINES;
      if (!((CF::BM % C::NT == 0 || q < CF::BM) && x < N)) continue;
#pragma unroll
      for (int r = 0; r < CF::RM; ++r)
        dot += (double)th[s][r] * (double)ga.dir[tile + (size_t)(i + r * CF::TM) * N + x];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    nrm += __shfl_down(nrm, off, 64);
    dot += __shfl_down(dot, off, 64);
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) {
    red[2 * wave] = nrm;
    red[2 * wave + 1] = dot;
  }
  __syncthreads();
  if (threadIdx.x == 0 && ga.partials != nullptr) {
    double sn = 0.0, sd = 0.0;
    for (int w = 0; w < C::NT / 64; ++w) {
      sn += red[2 * w];
      sd += red[2 * w + 1];
    }
    double* out = ga.partials + ((size_t)b * gridDim.x + cb) * 2;
    out[0] = sn;
    out[1] = sd;
  }
}

// ---------------------------------------------------------------- launchers
template <int N>
struct SocsImpl {
  static int col_blocks() { return (N + ColCfg<N>::LINES - 1) / ColCfg<N>::LINES; }
  static int img_cols() {
    if (img_reg_kind<N>() == 1 && image_reg()) return kImgRegLines;
    if (img_reg_kind<N>() == 2 && image_reg()) return 2 * kImgRegPairLines;
    return image_pairs() && ImgPairCfg<N>::USE ? ImgPairCfg<N>::COLS : ImgCfg<N>::LINES;
  }
  static int img_blocks() { return (N + img_cols() - 1) / img_cols(); }

  static void forward_half(const SocsArgs& a, int batch, int mode, const float* base, const float* dir,
                           const float* coef, float* theta_out, float* mask_out, hipStream_t s) {
    using R = RowCfg<N>;
    using P = RowPCfg<N>;
    constexpr bool kPrunable = fft::factorize(N).count >= 2;
    // K1: mask rows -> row spectra at the window columns
    bool done = false;
    if constexpr (kPrunable) {
      if (a.row_pruned) {
        dim3 g1((N + 2 * P::LINES - 1) / (2 * P::LINES), batch);
        hipLaunchKernelGGL(k_row_fwd_pruned<N>, g1, dim3(P::K1NT), 0, s, a, mode, base, dir, coef, theta_out, mask_out);
        done = true;
      }
    }
    if (!done) {
      dim3 g1((N + 2 * R::LINES - 1) / (2 * R::LINES), batch);
      hipLaunchKernelGGL(k_row_fwd<N>, g1, dim3(R::NT), 0, s, a, mode, base, dir, coef, theta_out, mask_out);
    }
    // K2: columns -> mask spectrum window
    done = false;
    if constexpr (kPrunable) {
      if (a.pruned) {
        dim3 g2((a.wc + P::WLINES - 1) / P::WLINES, batch);
        hipLaunchKernelGGL(k_col_fwd_pruned<N>, g2, dim3(P::WNT), 0, s, a);
        done = true;
      }
    }
    if (!done) {
      dim3 g2((a.wc + R::LINES - 1) / R::LINES, batch);
      hipLaunchKernelGGL(k_col_fwd<N>, g2, dim3(R::NT), 0, s, a);
    }
    // K3: T_k rows
    done = false;
    if constexpr (kPrunable) {
      if (a.row_pruned) {
        dim3 g3((a.k * a.wr + P::LINES - 1) / P::LINES, batch);
        hipLaunchKernelGGL(k_row_inv_pruned<N>, g3, dim3(P::NT), 0, s, a);
        done = true;
      }
    }
    if (!done) {
      dim3 g3((a.k * a.wr + R::LINES - 1) / R::LINES, batch);
      hipLaunchKernelGGL(k_row_inv<N>, g3, dim3(R::NT), 0, s, a);
    }
  }

  static void image_adjoint(const SocsArgs& a, const ImageArgs& im, int batch, hipStream_t s, int part) {
    using C = ColCfg<N>;
    dim3 g(col_blocks(), batch);
    if (a.corr_len > 0) {
      if (part != kColumnAdjoint) {
        launch_corr_any(0, a, batch, a.spart, s);
        bool image_done = false;
        if constexpr (img_reg_kind<N>() == 1) {
          if (image_reg()) {
            hipLaunchKernelGGL((k_col_image_reg<N, kImgRegLines>), dim3(img_blocks(), batch),
                               dim3(ImgRegCfg<N, kImgRegLines>::NT), 0, s, a, im);
            image_done = true;
          }
        } else if constexpr (img_reg_kind<N>() == 2) {
          if (image_reg()) {
            hipLaunchKernelGGL((k_col_image_reg2<N, kImgRegPairLines>), dim3(img_blocks(), batch),
                               dim3(ImgRegCfg<N, kImgRegPairLines>::NT), 0, s, a, im);
            image_done = true;
          }
        }
        if (!image_done) {
          if (image_pairs() && ImgPairCfg<N>::USE)
            hipLaunchKernelGGL(k_col_image_corr2<N>, dim3(img_blocks(), batch), dim3(ImgPairCfg<N>::NT), 0, s, a, im);
          else
            hipLaunchKernelGGL(k_col_image_corr<N>, dim3(img_blocks(), batch), dim3(ImgCfg<N>::NT), 0, s, a, im);
        }
      }
      if (part != kColumnImage && im.target != nullptr && a.v != nullptr) {
        if (a.adj_mfma)
          launch_corr_adjoint_mfma(a, batch, s);
        else
          launch_corr_any(1, a, batch, nullptr, s);
      }
      return;
    }
    if (part == kColumnAdjoint) return;
    if constexpr (fft::factorize(N).count >= 2) {
      if (a.pruned) {
        hipLaunchKernelGGL(k_col_image_adj_pruned<N>, g, dim3(C::NT), 0, s, a, im);
        return;
      }
    }
    hipLaunchKernelGGL(k_col_image_adj<N>, g, dim3(C::NT), 0, s, a, im);
  }

  static void adjoint_tail(const SocsArgs& a, const GradArgs& ga, int batch, hipStream_t s) {
    using R = RowCfg<N>;
    using P = RowPCfg<N>;
    using C = ColCfg<N>;
    bool done = false;
    if constexpr (fft::factorize(N).count >= 2) {
      if (a.row_pruned) {
        dim3 g6((a.k * a.wr + P::LINES - 1) / P::LINES, batch);
        hipLaunchKernelGGL(k_row_adj_pruned<N>, g6, dim3(P::NT), 0, s, a);
        dim3 g7((a.wr + P::WLINES - 1) / P::WLINES, batch);
        hipLaunchKernelGGL(k_final_row_inv_pruned<N>, g7, dim3(P::WNT), 0, s, a);
        done = true;
      }
    }
    if (!done) {
      dim3 g6((a.k * a.wr + R::LINES - 1) / R::LINES, batch);
      hipLaunchKernelGGL(k_row_adj<N>, g6, dim3(R::NT), 0, s, a);
      dim3 g7((a.wr + R::LINES - 1) / R::LINES, batch);
      hipLaunchKernelGGL(k_final_row_inv<N>, g7, dim3(R::NT), 0, s, a);
    }
    dim3 g8(col_blocks(), batch);
    if constexpr (fft::factorize(N).count >= 2) {
      if (a.pruned) {
        hipLaunchKernelGGL(k_final_col_inv_pruned<N>, g8, dim3(C::NT), 0, s, a, ga);
        return;
      }
    }
    hipLaunchKernelGGL(k_final_col_inv<N>, g8, dim3(C::NT), 0, s, a, ga);
  }
};

void reduce_pairs(const double* partials, int nblk, double* out, int out_stride, int batch, hipStream_t s,
                  const int* active) {
  launch_reduce_fixed(partials, nblk, 2, out, out_stride, batch, s, active);
}

template <int N>
static const SocsOps* make_ops() {
  static const SocsOps ops = {N, &SocsImpl<N>::col_blocks, &SocsImpl<N>::img_blocks, &SocsImpl<N>::forward_half,
                              &SocsImpl<N>::image_adjoint, &SocsImpl<N>::adjoint_tail};
  return &ops;
}

int first_radix(int n) {
  const fft::Radices f = fft::factorize(n);
  return f.count >= 2 ? f.r[0] : 0;
}

std::vector<float2> corr_twiddles(int l) {
  const fft::RegPlan pl = fft::reg_plan(l);
  if (pl.count < 2) return {};
  std::vector<float2> out((size_t)pl.tab);
  for (int q = 1; q < pl.count; ++q)
    for (int r = 1; r < pl.R[q]; ++r)
      for (int k = 0; k < pl.P[q]; ++k) {
        const double ang = -2.0 * fft::kPi * (double)r * (double)k / ((double)pl.P[q] * (double)pl.R[q]);
        out[(size_t)pl.taboff[q] + (size_t)(r - 1) * pl.P[q] + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
      }
  return out;
}

std::vector<float2> row_stage_twiddles(int n) {
  const fft::Radices f = fft::factorize(n);
  if (f.count < 2) return {};
  const fft::StageTables t = fft::stage_tables(n);
  std::vector<float2> out((size_t)t.off[t.count]);
  for (int st = 0; st < t.count; ++st)
    for (int r = 1; r < t.R[st]; ++r)
      for (int k = 0; k < t.P[st]; ++k) {
        const double ang = -2.0 * fft::kPi * (double)r * (double)k / ((double)t.P[st] * (double)t.R[st]);
        out[(size_t)t.off[st] + (size_t)(r - 1) * t.P[st] + k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
      }
  return out;
}

const SocsOps* socs_ops_for(int n) {
  switch (n) {
#define LITHO_N(v) \
  case v:          \
    return make_ops<v>();
    LITHO_SUPPORTED_SIZES(LITHO_N)
#undef LITHO_N
    default:
      return nullptr;
  }
}

}  // namespace litho
