// Column-stage adjoint of the band-limited path on the matrix cores: the
// window rows of the adjoint column spectra as a Toeplitz product per column,
//   V_k[i] = sum_{j < W_r} ghat[i - j] T_k[j],   i in [0, W_r),
// ghat[-m] = conj ghat[m] (dL/dI is real), evaluated as one real GEMM per
// column x:
//   [Vr | Vi] (W_r x 2K) = [Gr | Gi] (W_r x 2 W_r) . [[Tr, Ti], [-Ti, Tr]] (2 W_r x 2K),
// G[i][j] = ghat[i - j].  It replaces the length-L transforms of K5 (FFT_L of
// every T_k, the product with FFT_L(h), the inverse) by 8 W_r^2 K real
// multiply-adds per column on v_mfma_f32_16x16x32_bf16, with FP32 operands
// split into bf16 pairs x = x_hi + x_lo and three products per term
// (hi.hi + hi.lo + lo.hi, FP32 accumulation).  Each operand keeps 16
// significant bits, so a product carries a relative error of about 3 * 2^-18,
// inside the adjoint's 1e-4 relative L2 tolerance (the image, which the
// tolerance of 1e-5 relative L-inf covers, is not computed here).
//
// Layout of the work: a block holds CB consecutive columns of one tile, WPC
// waves per column; wave (col, wm) owns the output rows of M-tiles
// [wm MPW, (wm + 1) MPW) of its column over all 2K GEMM columns (NTL tiles of
// 16).  The GEMM's inner dimension runs over window rows j in chunks of 32:
// the chunk of T (CB columns x K kernels x 32 rows) is staged in LDS (rows j
// contiguous per (column, k), so a lane's eight B values are one aligned
// 64-byte run) while the next chunk's global loads are in flight, and each
// chunk is one MFMA k-step of the Re half ([Gr] against [Tr | Ti]) and one of
// the Im half ([Gi] against [-Ti | Tr]).  The Toeplitz operand is read from a
// per-column lag table ghat[m], m from 16 MT - 1 down to -(32 KS - 1), stored
// reversed so a lane's eight consecutive j are eight ascending entries.
//
// MFMA 16x16x32 lane layout (checked on the device): A row = lane % 16,
// B column = lane % 16, A and B inner index 8 (lane / 16) + e for element e
// (any mapping shared by A and B gives the same sum), D row 4 (lane / 16) + e,
// D column lane % 16.
#include <hip/hip_runtime.h>

#include "socs_kernels.hpp"

namespace litho {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaWaves = 4;      // waves per block: two blocks per CU overlap one's loads with the other's stores
constexpr int kChunk = 32;          // window rows per k-step
// staged chunk row stride (bf16): 80 bytes, so a lane's eight values are one
// aligned 16-byte read and the 16 lanes of a quarter-wave (rows 80 bytes
// apart) cover the 64 banks once
constexpr int kChunkStride = 40;

__device__ __forceinline__ unsigned short bf16_bits(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ float bf16_value(unsigned short h) { return __builtin_bit_cast(float, (unsigned int)h << 16); }
// x = hi + lo to 16 significant bits: hi = bf16(x), lo = bf16(x - hi)
__device__ __forceinline__ void split_bf16(float x, unsigned short& hi, unsigned short& lo) {
  hi = bf16_bits(x);
  lo = bf16_bits(x - bf16_value(hi));
}

__device__ __forceinline__ int xcd_block(int bx, int nb) { return nb % 8 == 0 ? (bx % 8) * (nb / 8) + bx / 8 : bx; }

template <int MPW, int WPC, int NTL>
struct MfmaCfg {
  static constexpr int CB = kMfmaWaves / WPC;         // columns per block
  static constexpr int NT = 64 * kMfmaWaves;          // threads
  static constexpr int MT = MPW * WPC;                // M-tiles of a column (W_r <= 16 MT)
  static constexpr int KP = 8 * NTL;                  // kernels per block (2 K <= 16 NTL)
  static constexpr int GL = 32 * ((MT + 1) / 2) + 16 * MT;   // lag-table length >= 16 MT + 32 KS
  // column stride of a staged chunk array: KP rows plus 32 bf16 of padding, so the
  // stash's CB lanes of one row pair (c fastest) land 48 banks apart
  static constexpr int CS = KP * kChunkStride + 32;
  static constexpr int CH = CB * CS;                  // bf16 per staged chunk array
  static constexpr int NP = 16 * NTL + 4;             // staged output row (floats): 4 NP = 16 mod 64 banks
  static constexpr int SS = 16 * NP + 4;              // staged tile stride per (column, wave), 4 banks apart
  static constexpr int STG = CB * WPC * SS;           // staged output tile (floats)
  // the chunk buffer (in bf16 units) also holds the FP32 lag table first and the output staging last
  static constexpr int CHK_MIN = 4 * CH > CB * GL * 4 ? 4 * CH : CB * GL * 4;
  static constexpr int CHK = CHK_MIN > STG * 2 ? CHK_MIN : STG * 2;
};

template <int MPW, int WPC, int NTL>
__global__ void __launch_bounds__(64 * kMfmaWaves) __attribute__((amdgpu_waves_per_eu(2))) k_corr_adjoint_mfma(SocsArgs a) {
  using C = MfmaCfg<MPW, WPC, NTL>;
  // lag tables, pair-packed: lagp[o][c][u] = (G_o[u], G_o[u + 1]) as bf16 x 2 for the
  // operands o = Re hi, Re lo, Im hi, Im lo of column c
  __shared__ unsigned int lagp[4 * C::CB * C::GL];
  // chunk of T as bf16: Tr hi, Tr lo, Ti hi, Ti lo, each [c][k][j] (row stride kChunkStride);
  // before the first chunk it holds the FP32 lag table, after the last the output staging
  __shared__ __attribute__((aligned(16))) unsigned short chk[C::CHK];
  const int b = blockIdx.y;
  if (a.active != nullptr && a.active[b] == 0) return;
  const int n = a.n, W = a.wr, K = a.k;
  const int x0 = xcd_block(blockIdx.x, gridDim.x) * C::CB;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int col = wave % C::CB, wm = wave / C::CB;
  const int nn = lane % 16, q = lane / 16;
  const int KS = (W + kChunk - 1) / kChunk;
  const int mmax = 16 * C::MT - 1;
  const int glen = 16 * C::MT + kChunk * KS;

  // FP32 lag table tab[c][u] = ghat[m] of column x0 + c, m = mmax - u (zero for |m| >= W)
  float2* tab = reinterpret_cast<float2*>(chk);
  const float2* gh = a.ghat + (size_t)b * W * n;
  for (int e = threadIdx.x; e < C::CB * glen; e += C::NT) {
    const int c = e % C::CB, u = e / C::CB;
    const int x = x0 + c, m = mmax - u;
    float2 v = make_float2(0.f, 0.f);
    if (x < n && m > -W && m < W) {
      if (m >= 0) {
        v = gh[(size_t)m * n + x];
      } else {
        const float2 w = gh[(size_t)(-m) * n + x];
        v = make_float2(w.x, -w.y);
      }
    }
    tab[c * C::GL + u] = v;
  }
  __syncthreads();
  for (int c = 0; c < C::CB; ++c) {
    for (int u = threadIdx.x; u < glen; u += C::NT) {
      const float2 g0 = tab[c * C::GL + u];
      const float2 g1 = u + 1 < glen ? tab[c * C::GL + u + 1] : make_float2(0.f, 0.f);
      unsigned short rh0, rl0, ih0, il0, rh1, rl1, ih1, il1;
      split_bf16(g0.x, rh0, rl0);
      split_bf16(g0.y, ih0, il0);
      split_bf16(g1.x, rh1, rl1);
      split_bf16(g1.y, ih1, il1);
      const int o = c * C::GL + u;
      lagp[0 * C::CB * C::GL + o] = rh0 | ((unsigned int)rh1 << 16);
      lagp[1 * C::CB * C::GL + o] = rl0 | ((unsigned int)rl1 << 16);
      lagp[2 * C::CB * C::GL + o] = ih0 | ((unsigned int)ih1 << 16);
      lagp[3 * C::CB * C::GL + o] = il0 | ((unsigned int)il1 << 16);
    }
  }

  // chunk staging: thread e covers rows (j, j + 1) of (c, k), c fastest (CB consecutive columns per row)
  constexpr int PAIRS = C::CB * C::KP * (kChunk / 2);
  constexpr int PER = (PAIRS + C::NT - 1) / C::NT;
  const float2* t = a.t + (size_t)b * K * W * n;
  float2 pre[PER][2];
  // per pair slot: element offset of row 2 jp of (k, x) and the pair's first row
  int poff[PER], prow[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int e = threadIdx.x + s * C::NT;
    const int c = e % C::CB, r = e / C::CB;
    const int jp = r % (kChunk / 2), k = r / (kChunk / 2);
    const bool ok = e < PAIRS && k < K && x0 + c < n;
    poff[s] = ok ? (k * W + 2 * jp) * n + x0 + c : 0;
    prow[s] = ok ? 2 * jp : W;   // rows at or past W load zero
  }
  auto fetch = [&](int j0) {
    const float2* tj = t + (size_t)j0 * n;
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const float2* src = tj + poff[s];
      const int j = j0 + prow[s];
      pre[s][0] = j < W ? src[0] : make_float2(0.f, 0.f);
      pre[s][1] = j + 1 < W ? src[n] : make_float2(0.f, 0.f);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const int e = threadIdx.x + s * C::NT;
      if (e < PAIRS) {
        const int c = e % C::CB, r = e / C::CB;
        const int jp = r % (kChunk / 2), k = r / (kChunk / 2);
        unsigned short rh0, rl0, ih0, il0, rh1, rl1, ih1, il1;
        split_bf16(pre[s][0].x, rh0, rl0);
        split_bf16(pre[s][0].y, ih0, il0);
        split_bf16(pre[s][1].x, rh1, rl1);
        split_bf16(pre[s][1].y, ih1, il1);
        unsigned int* d = reinterpret_cast<unsigned int*>(chk) + (c * C::CS + k * kChunkStride + 2 * jp) / 2;
        d[0 * C::CH / 2] = rh0 | ((unsigned int)rh1 << 16);
        d[1 * C::CH / 2] = rl0 | ((unsigned int)rl1 << 16);
        d[2 * C::CH / 2] = ih0 | ((unsigned int)ih1 << 16);
        d[3 * C::CH / 2] = il0 | ((unsigned int)il1 << 16);
      }
    }
  };

  f32x4 acc[MPW][NTL];
#pragma unroll
  for (int s = 0; s < MPW; ++s)
#pragma unroll
    for (int t2 = 0; t2 < NTL; ++t2) acc[s][t2] = f32x4{0.f, 0.f, 0.f, 0.f};

  // B operands of this lane: GEMM column 16 nt + nn is Vr of kernel g (g < K)
  // or Vi of kernel g - K; the Re half multiplies Tr / Ti, the Im half -Ti / Tr
  int boff[NTL];
  bool bim[NTL], blive[NTL];
#pragma unroll
  for (int nt = 0; nt < NTL; ++nt) {
    const int g = 16 * nt + nn;
    bim[nt] = g >= K;
    blive[nt] = g < 2 * K;
    const int k = !blive[nt] ? 0 : (bim[nt] ? g - K : g);
    boff[nt] = col * C::CS + k * kChunkStride + 8 * q;
  }
  const unsigned int* lag = lagp + col * C::GL + mmax - nn + 8 * q;

  fetch(0);
#pragma unroll 1
  for (int ck = 0; ck < KS; ++ck) {
    __syncthreads();   // the previous chunk's readers are done (the lag tables complete)
    stash();
    __syncthreads();
    if (ck + 1 < KS) fetch((ck + 1) * kChunk);
    bf16x8 brh[NTL], brl[NTL], bih[NTL], bil[NTL];
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt) {
      const unsigned short* base = chk + boff[nt];
      const u32x4 trh = *reinterpret_cast<const u32x4*>(base + 0 * C::CH);
      const u32x4 trl = *reinterpret_cast<const u32x4*>(base + 1 * C::CH);
      const u32x4 tih = *reinterpret_cast<const u32x4*>(base + 2 * C::CH);
      const u32x4 til = *reinterpret_cast<const u32x4*>(base + 3 * C::CH);
      const u32x4 zero = {0u, 0u, 0u, 0u};
      const unsigned int neg = 0x80008000u;
      const u32x4 negm = {neg, neg, neg, neg};
      u32x4 rh = bim[nt] ? tih : trh, rl = bim[nt] ? til : trl;
      u32x4 ih = bim[nt] ? trh : (tih ^ negm), il = bim[nt] ? trl : (til ^ negm);
      if (!blive[nt]) rh = rl = ih = il = zero;
      brh[nt] = __builtin_bit_cast(bf16x8, rh);
      brl[nt] = __builtin_bit_cast(bf16x8, rl);
      bih[nt] = __builtin_bit_cast(bf16x8, ih);
      bil[nt] = __builtin_bit_cast(bf16x8, il);
    }
    // A operands: row i = 16 mt + nn, inner j = 32 ck + 8 q + e -> lag entry u = mmax - i + j
#pragma unroll
    for (int s = 0; s < MPW; ++s) {
      const int mt = wm * MPW + s;
      const unsigned int* gp = lag + kChunk * ck - 16 * mt;
      u32x4 w[4];
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        const unsigned int* go = gp + o * C::CB * C::GL;
        w[o] = u32x4{go[0], go[2], go[4], go[6]};
      }
      const bf16x8 arh = __builtin_bit_cast(bf16x8, w[0]), arl = __builtin_bit_cast(bf16x8, w[1]);
      const bf16x8 aih = __builtin_bit_cast(bf16x8, w[2]), ail = __builtin_bit_cast(bf16x8, w[3]);
#pragma unroll
      for (int nt = 0; nt < NTL; ++nt) {
        f32x4 c4 = acc[s][nt];
        c4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, brh[nt], c4, 0, 0, 0);
        c4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, brl[nt], c4, 0, 0, 0);
        c4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl, brh[nt], c4, 0, 0, 0);
        c4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, bih[nt], c4, 0, 0, 0);
        c4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, bil[nt], c4, 0, 0, 0);
        c4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail, bih[nt], c4, 0, 0, 0);
        acc[s][nt] = c4;
      }
    }
  }

  // write-out, one M-tile slot s at a time through LDS, CB columns contiguous per row
  float* stg = reinterpret_cast<float*>(chk);
  float2* vv = a.v + (size_t)b * K * W * n;
#pragma unroll
  for (int s = 0; s < MPW; ++s) {
    __syncthreads();   // the chunk buffer (or the previous slot) is no longer read
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt)
#pragma unroll
      for (int e = 0; e < 4; ++e) stg[(col * WPC + wm) * C::SS + (4 * q + e) * C::NP + 16 * nt + nn] = acc[s][nt][e];
    __syncthreads();
    // thread -> (column c fastest, row, kernel k0 + kstep i)
    constexpr int KSTEP = C::NT / (16 * C::CB);
    const int c = threadIdx.x % C::CB, row = (threadIdx.x / C::CB) % 16, k0 = threadIdx.x / (16 * C::CB);
    const int x = x0 + c;
#pragma unroll
    for (int w2 = 0; w2 < WPC; ++w2) {
      const int i = (w2 * MPW + s) * 16 + row;
      if (i < W && x < n) {
        const float* d = stg + (c * WPC + w2) * C::SS + row * C::NP;
        float2* dst = vv + (size_t)i * n + x;
        for (int k = k0; k < K; k += KSTEP) dst[(size_t)k * W * n] = make_float2(d[k], d[k + K]);
      }
    }
  }
}

using MfmaLaunch = void (*)(const SocsArgs&, int, hipStream_t);

template <int MPW, int WPC, int NTL>
void mfma_launch(const SocsArgs& a, int batch, hipStream_t s) {
  using C = MfmaCfg<MPW, WPC, NTL>;
  dim3 grid((a.n + C::CB - 1) / C::CB, batch);
  hipLaunchKernelGGL((k_corr_adjoint_mfma<MPW, WPC, NTL>), grid, dim3(C::NT), 0, s, a);
}

template <int NTL>
MfmaLaunch mfma_pick(int mt) {
  // one wave per column up to 8 M-tiles (W_r <= 128), two from 9 to 16 (W_r <= 256)
  switch (mt) {
    case 1: return mfma_launch<1, 1, NTL>;
    case 2: return mfma_launch<2, 1, NTL>;
    case 3: return mfma_launch<3, 1, NTL>;
    case 4: return mfma_launch<4, 1, NTL>;
    case 5: return mfma_launch<5, 1, NTL>;
    case 6: return mfma_launch<6, 1, NTL>;
    case 7: return mfma_launch<7, 1, NTL>;
    case 8: return mfma_launch<8, 1, NTL>;
    case 9: case 10: return mfma_launch<5, 2, NTL>;
    case 11: case 12: return mfma_launch<6, 2, NTL>;
    case 13: case 14: return mfma_launch<7, 2, NTL>;
    case 15: case 16: return mfma_launch<8, 2, NTL>;
    default: return nullptr;
  }
}

MfmaLaunch mfma_select(int wr, int k) {
  const int mt = (wr + 15) / 16;
  switch ((2 * k + 15) / 16) {
    case 1: return mfma_pick<1>(mt);
    case 2: return mfma_pick<2>(mt);
    case 3: return mfma_pick<3>(mt);   // K <= 24 (more kernels would spill the accumulators)
    default: return nullptr;
  }
}

}  // namespace

bool corr_adjoint_mfma_supported(int wr, int k) { return wr >= 1 && k >= 1 && mfma_select(wr, k) != nullptr; }

void launch_corr_adjoint_mfma(const SocsArgs& a, int batch, hipStream_t s) {
  const MfmaLaunch f = mfma_select(a.wr, a.k);
  if (f != nullptr) f(a, batch, s);
}

}  // namespace litho
