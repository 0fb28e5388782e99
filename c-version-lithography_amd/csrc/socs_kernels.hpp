// Kernel-side argument blocks and the per-size launcher table of the SOCS
// forward / adjoint pipeline (socs_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace litho {

// Tile sizes with compiled FFT instantiations: the CPU-test sizes, the Micro
// tile sizes (core + 2*halo) and the single-tile sizes 512 / 2048.
#ifndef LITHO_SUPPORTED_SIZES
#define LITHO_SUPPORTED_SIZES(X)                                                            \
  X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) \
  X(320) X(384) X(512) X(640) X(768) X(896) X(1024) X(1152) X(1280) X(1536) X(2048) X(2176)
#endif

// Plan-resident data shared by all kernels of one tile size (per-tile buffers
// are indexed by blockIdx.y with the strides implied by the shapes).
struct SocsArgs {
  int k, wr, wc;                  // kernels, window rows, window columns
  const int* row_freq;            // [wr] DC-corner frequency index of window row i
  const int* col_freq;            // [wc]
  const float2* h;                // [k][wr][wc] scaled kernels H~_k
  const float2* tw;               // [N] twiddles e^{-2 pi i m / N}
  float2* rt;                     // [b][wc][N]   row spectra at window columns
  float2* mhat;                   // [b][wr][wc]  mask spectrum window
  float2* t;                      // [b][k][wr][N] row-inverse spectra
  float2* v;                      // [b][k][wr][N] adjoint column spectra
  float2* y;                      // [b][k][wr][wc]
  float2* p;                      // [b][wr][N]
  const int* active;              // [b] nonzero = evaluate tile b; null = all tiles
  const int* row_of_u;            // [N] window row of column frequency u, or -1
  int pruned;                     // window rows fit the pruned column transforms
  const int* col_of_u;            // [N] window column of row frequency u, or -1
  const float2* rtw;              // stage twiddle tables of the row transforms (row_stage_twiddles)
  int row_pruned;                 // window columns fit the pruned row transforms
  int n;                          // tile size
  int corr_len;                   // L of the band-limited column path (0: fused column kernel)
  const float2* twl;              // pass twiddle tables of the length-L transforms (corr_twiddles(L))
  float2* acorr;                  // [b][wr][n] column autocorrelation c[d], d in [0, wr)
  float2* ghat;                   // [b][wr][n] column spectrum of dL/dI at m in [0, wr)
  float* spart;                   // partial power spectra of the k groups (corr_power_bytes)
  const float2* twi;              // pass twiddle tables of the length-N register transforms (corr_twiddles(n)), or null
  int adj_mfma = 0;               // band-limited adjoint as a Toeplitz GEMM on the matrix cores (corr_mfma.hip)
  // Point slots of the L-BFGS tiles (opt_kernels.hpp).  When set, K1 reads
  // the line-search origin theta from point ring[b * ring_stride] of tile b
  // and writes the trial theta to point ring[b * ring_stride + 1]; the final
  // column kernel reads that trial theta and writes its gradient to the same
  // point of the gradient slots.  Tile b's points start at point
  // b * ring_points.  Null: the theta buffers are tile-major like the others.
  const int* ring = nullptr;
  int ring_stride = 0, ring_points = 0;
};

// Element offset of tile b's origin (which = 0) or trial (which = 1) theta.
__device__ __forceinline__ size_t point_offset(const SocsArgs& a, int b, int which) {
  const size_t n2 = (size_t)a.n * a.n;
  if (a.ring == nullptr) return (size_t)b * n2;
  return ((size_t)b * a.ring_points + a.ring[b * a.ring_stride + which]) * n2;
}

// Lengths of the band-limited column path (powers of two).
bool corr_len_supported(int l);
// The band-limited adjoint V_k = Toeplitz(ghat) T_k on the matrix cores
// (corr_mfma.hip): supported for W_r <= 256 and K <= 24.
bool corr_adjoint_mfma_supported(int wr, int k);
void launch_corr_adjoint_mfma(const SocsArgs& a, int batch, hipStream_t s);
// Bytes of the partial power spectra of a batch (the k groups grow as the batch shrinks).
size_t corr_power_bytes(int n, int l, int k, int batch);

struct ImageArgs {
  float thr, alpha;
  float* aerial;                  // [b][N][N] or null
  float* wafer;                   // [b][N][N] or null
  const float* target;            // null: forward only (no loss, no adjoint)
  const float* pe_win;            // null = all ones
  const float* pen_win;           // null = no background penalty
  float pe_weight, penalty, pen_thr;
  float* dldi;                    // optional debug output
  const float* extra_dldi;        // optional: added to dL/dI (before the adjoint's scale) where target is set
  double* partials;               // [b][image blocks][2]: PE, background penalty
};

struct GradArgs {
  float scale;                    // 4 alpha
  float* grad_mask;               // dPE/dM, optional
  const float* theta;             // optional: enables the theta chain
  float* grad_theta;              // optional
  const float* dir;               // optional: partial dot grad_theta . dir
  const int* rect;                // optional [b][4] = y0,y1,x0,x1 update window
  double* partials;               // [b][col_blocks][2]: |G_theta|^2, G_theta . dir
  int nt_grad_theta = 0;          // 1: grad_theta stored non-temporally (evaluations of one tile)
};

// per-pixel resist, losses and dL/dI shared by the column image kernels
// tv: the target value at gi, loaded by the caller (read only when im.target is set)
__device__ __forceinline__ float image_pixel_t(const ImageArgs& im, size_t gi, float iv, float tv, double& pe,
                                               double& bg) {
  const float w = 1.0f / (1.0f + expf(-im.alpha * (iv - im.thr)));
  if (im.aerial != nullptr) im.aerial[gi] = iv;
  if (im.wafer != nullptr) im.wafer[gi] = w;
  float g = 0.f;
  if (im.target != nullptr) {
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
    g = drive * w * (1.0f - w);
    if (im.extra_dldi != nullptr) g += im.extra_dldi[gi];
    if (im.dldi != nullptr) im.dldi[gi] = g;
  }
  return g;
}
__device__ __forceinline__ float image_pixel(const ImageArgs& im, size_t gi, float iv, double& pe, double& bg) {
  return image_pixel_t(im, gi, iv, im.target != nullptr ? im.target[gi] : 0.f, pe, bg);
}

enum ColumnPart { kColumnAll = 0, kColumnImage = 1, kColumnAdjoint = 2 };

struct SocsOps {
  int n;
  int (*col_blocks)();
  int (*img_blocks)();   // blocks of the band-limited image kernel (its FP64 partials)
  void (*forward_half)(const SocsArgs&, int batch, int mode, const float* base, const float* dir,
                       const float* coef, float* theta_out, float* mask_out, hipStream_t);
  // part: kColumnAll, or kColumnImage / kColumnAdjoint for the two halves of
  // the band-limited path (the fused kernels run whole under kColumnImage)
  void (*image_adjoint)(const SocsArgs&, const ImageArgs&, int batch, hipStream_t, int part);
  void (*adjoint_tail)(const SocsArgs&, const GradArgs&, int batch, hipStream_t);
};

const SocsOps* socs_ops_for(int n);
// Pipeline for tile sizes without compiled transforms (generic_ops.hip): the
// same stages with direct window-restricted DFTs; needs the band-limited
// column path (corr_len > 0) and n <= kGenericMaxN.
const SocsOps* generic_socs_ops();
constexpr int kGenericMaxN = 4096;
// The length-L correlation kernels of the band-limited column path: stage 0
// the power spectra and autocorrelation (K4a), stage 1 the adjoint window
// rows (K5, length-L transforms).
void launch_corr_any(int stage, const SocsArgs& a, int batch, float* spart, hipStream_t s);
// First radix of the column factorisation of n (0 if n has a single stage).
int first_radix(int n);
// Stage twiddle tables of the row transforms of length n (layout of
// fft::stage_tables), computed in FP64 and rounded to FP32; empty if n has a
// single stage.
std::vector<float2> row_stage_twiddles(int n);
// Pass twiddle tables of the register-resident length-L transforms (fft_regs.hpp).
std::vector<float2> corr_twiddles(int l);
void reduce_pairs(const double* partials, int nblk, double* out, int out_stride, int batch, hipStream_t s,
                  const int* active = nullptr);

}  // namespace litho
