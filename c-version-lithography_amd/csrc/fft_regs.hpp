// Register-resident line transforms of power-of-two length L (64 ... 2048)
// for the band-limited column path and its image kernel.  Thread j of a line (0 <= j < L/8) holds
// the eight elements j + r L/8 (r = 0..7) in registers both on entry and on
// exit: the first and last Stockham passes are radix 8, so the first reads
// its butterfly straight from the registers (which the caller fills from
// global memory) and the last leaves its outputs there (which the caller
// stores, or multiplies and feeds to the next transform).  The passes in
// between exchange through LDS.  Every exchange writes one of two buffers,
// alternating, so it needs a single barrier: a buffer is written again only
// after the barrier of the exchange in between, which every thread reaches
// after its reads of that buffer.  The barriers order LDS only (lds_barrier):
// the callers (the power spectra, the autocorrelation and the adjoint of the
// band-limited column path) read only global memory that earlier kernels
// wrote, so the next line's loads and the last line's stores stay in flight
// across the exchanges.
//
// Lanes run across lines (lane = line + LINES * jj), so the global loads and
// stores of a wave cover LINES consecutive columns of a row.  gfx950 serves a
// 64-bit LDS access (each address of a read2 / write2 separately) 16 lanes at
// a time over 16 eight-byte bank slots (measured with SQ_LDS_BANK_CONFLICT on
// probe patterns).  LDS element p of a line lives at p + (p >> 3); the line
// stride LS is congruent to J modulo 16, J = 16 / min(LINES, 16) being the
// number of j values among 16 consecutive lanes, so those lanes fall on
// distinct slots in every pass.
#pragma once
#include "fft_engine.hpp"

namespace litho {
namespace fft {

struct RegPlan {
  int count;       // passes
  int R[4];        // radix of each pass
  int P[4];        // span (product of the earlier radices)
  int twoff[4];    // first per-thread twiddle register of each pass
  int taboff[4];   // first entry of each pass in the twiddle table
  int twn;         // twiddle registers per thread
  int tab;         // table entries
};

// 64 = 8*8, 128 = 8*2*8, 256 = 8*4*8, 512 = 8*8*8, 1024 = 8*4*4*8, 2048 = 8*4*8*8
constexpr RegPlan reg_plan(int l) {
  RegPlan p{};
  int mid[2] = {0, 0}, nm = 0;
  if (l == 128) mid[nm++] = 2;
  else if (l == 256) mid[nm++] = 4;
  else if (l == 512) mid[nm++] = 8;
  else if (l == 1024) { mid[nm++] = 4; mid[nm++] = 4; }
  else if (l == 2048) { mid[nm++] = 4; mid[nm++] = 8; }
  else if (l != 64) { p.count = -1; return p; }
  p.R[p.count++] = 8;
  for (int i = 0; i < nm; ++i) p.R[p.count++] = mid[i];
  p.R[p.count++] = 8;
  int span = 1;
  for (int q = 0; q < p.count; ++q) {
    p.P[q] = span;
    p.twoff[q] = p.twn;
    p.taboff[q] = p.tab;
    if (q > 0) {
      p.twn += (8 / p.R[q]) * (p.R[q] - 1);
      p.tab += (p.R[q] - 1) * span;
    }
    span *= p.R[q];
  }
  return p;
}

constexpr int reg_line_stride(int l, int lines) {
  const int g = lines < 16 ? lines : 16;
  const int j = 16 / g;
  int ls = l + l / 8;
  while (ls % 16 != j % 16) ++ls;
  return ls;
}

template <int L, int LINES>
struct RegFFT {
  static constexpr RegPlan PL = reg_plan(L);
  static_assert(PL.count >= 2, "register transforms cover L = 64 ... 2048");
  static constexpr int Q = L / 8;               // threads per line
  static constexpr int NT = LINES * Q;
  static constexpr int SH = 3;
  static constexpr int LS = reg_line_stride(L, LINES);
  static constexpr int BUF = LINES * LS;        // one of the two exchange buffers (float2)
  static constexpr int TWN = PL.twn > 0 ? PL.twn : 1;

  // Twiddles of this thread's butterflies: pass q >= 1 (radix R, span P),
  // butterfly t (i = j + t Q), factor r >= 1: table entry
  // taboff[q] + (r-1) P + (i mod P) = e^{-2 pi i r (i mod P) / (P R)}.
  __device__ __forceinline__ static void load_twiddles(const float2* __restrict__ tab, float2 (&tw)[TWN], int j) {
    load_tw<1>(tab, tw, j);
  }

  // DIR = -1 forward, +1 unnormalised inverse; ex counts exchanges (buffer ex & 1).
  // HALF: the positions j + r L/8 with r >= 4 are zero on entry (v[4..7] are
  // not read), as for inputs supported on fewer than L/2 positions.
  template <int DIR, bool HALF = false>
  __device__ __forceinline__ static void transform(float2 (&v)[8], float2* lds, const float2 (&tw)[TWN], int line,
                                                   int j, int& ex) {
    if constexpr (HALF) {
      // X[2m] = DFT4(x)[m], X[2m+1] = DFT4(x_r w8^r)[m] for x = (x0, x1, x2, x3, 0, 0, 0, 0)
      float2 a[4] = {v[0], v[1], v[2], v[3]};
      float2 c[4] = {v[0], twiddle_ct<1, 8, DIR>(v[1]), twiddle_ct<2, 8, DIR>(v[2]), twiddle_ct<3, 8, DIR>(v[3])};
      Dft<4, DIR>::run(a);
      Dft<4, DIR>::run(c);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        v[2 * m] = a[m];
        v[2 * m + 1] = c[m];
      }
    } else {
      Dft<8, DIR>::run(v);
    }
    {
      // span 1: butterfly j writes positions 8 j + r
      float2* dst = lds + (ex & 1) * BUF + line * LS + swz<SH>(8 * j);
#pragma unroll
      for (int r = 0; r < 8; ++r) dst[r] = v[r];
    }
    lds_barrier();
    middle<1, DIR>(v, lds, tw, line, j, ex);
    {
      constexpr int q = PL.count - 1;
      const float2* src = lds + (ex & 1) * BUF + line * LS + swz<SH>(j);
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = src[r * Q + ((r * Q) >> SH)];
      ++ex;
#pragma unroll
      for (int r = 1; r < 8; ++r) v[r] = DIR < 0 ? cmul(v[r], tw[PL.twoff[q] + r - 1]) : cmulc(v[r], tw[PL.twoff[q] + r - 1]);
      Dft<8, DIR>::run(v);
    }
  }

 private:
  template <int q>
  __device__ __forceinline__ static void load_tw(const float2* __restrict__ tab, float2 (&tw)[TWN], int j) {
    if constexpr (q < PL.count) {
      constexpr int R = PL.R[q], P = PL.P[q];
#pragma unroll
      for (int t = 0; t < 8 / R; ++t) {
        const int k = (j + t * Q) % P;
#pragma unroll
        for (int r = 1; r < R; ++r) tw[PL.twoff[q] + t * (R - 1) + r - 1] = tab[PL.taboff[q] + (r - 1) * P + k];
      }
      load_tw<q + 1>(tab, tw, j);
    }
  }

  // passes 1 .. count-2: read butterflies i = j + t Q (inputs i + r T) from
  // buffer ex, write outputs (i - k) R + k + r P, k = i mod P, to buffer ex + 1
  template <int q, int DIR>
  __device__ __forceinline__ static void middle(float2 (&v)[8], float2* lds, const float2 (&tw)[TWN], int line,
                                                int j, int& ex) {
    if constexpr (q < PL.count - 1) {
      constexpr int R = PL.R[q], P = PL.P[q], T = L / R, B = 8 / R;
      {
        const float2* src = lds + (ex & 1) * BUF + line * LS;
#pragma unroll
        for (int t = 0; t < B; ++t) {
          const int si = swz<SH>(j + t * Q);
#pragma unroll
          for (int r = 0; r < R; ++r) v[t * R + r] = src[si + r * T + ((r * T) >> SH)];
        }
      }
      ++ex;
      float2* dst = lds + (ex & 1) * BUF + line * LS;
#pragma unroll
      for (int t = 0; t < B; ++t) {
        const int i = j + t * Q, k = i % P;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const float2 w = tw[PL.twoff[q] + t * (R - 1) + r - 1];
          v[t * R + r] = DIR < 0 ? cmul(v[t * R + r], w) : cmulc(v[t * R + r], w);
        }
        Dft<R, DIR>::run(v + t * R);
        const int so = swz<SH>((i - k) * R + k);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[so + r * P + ((r * P) >> SH)] = v[t * R + r];
      }
      lds_barrier();
      middle<q + 1, DIR>(v, lds, tw, line, j, ex);
    }
  }
};

}  // namespace fft
}  // namespace litho
