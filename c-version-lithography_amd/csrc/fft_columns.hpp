// Window-pruned column transforms for the fused image / adjoint kernels.
//
// A block transforms LINES columns of length N.  The inverse transform
// (DIR +1) uses the factorisation (R_1, ..., R_M) of fft_engine.hpp and the
// forward transform (DIR -1) the reversed one (R_M, ..., R_1), so that:
//  * inverse stage 1 (radix R_1 = 16, stride N/16) sees at most two nonzero
//    inputs per butterfly — the pupil window occupies only the first and the
//    last N/16 frequencies — and loads them straight from HBM: no LDS zero fill
//    or scatter;
//  * the last inverse stage keeps its outputs (image rows i + r*N/R_M) in
//    registers, where the caller accumulates |E|^2 or multiplies by dLdI;
//  * forward stage 1 (radix R_M, stride N/R_M) reads exactly those registers;
//  * the last forward stage (radix R_1) computes only the two outputs that land
//    in the window and stores them straight to HBM.
// Butterfly q of a first / last stage maps to column q % LINES and index
// q / LINES, so consecutive lanes touch consecutive columns (coalesced rows).
// Precondition (checked on the host): every window frequency u satisfies
// u < N/R_1 or u >= N - N/R_1.
//
// Twiddles: the first and last stages read the N-entry table at r*i (a few
// distinct i per 32 lanes, so few bank conflicts).  The middle stages read
// per-pass tables in LDS, laid out [r-1][k] so that lanes with consecutive k
// read consecutive bank slots; indexing the N-entry table at r*k*STEP instead
// puts up to 32 lanes on one bank when STEP is a multiple of 32.
#pragma once
#include "fft_engine.hpp"

namespace litho {
namespace fft {

// Layout of the middle-stage tables of a column transform of length n: the
// inverse stages 1..M-2 (radix r[j], span r[0]*...*r[j-1]) then the forward
// stages 1..M-2 (radix r[M-1-j], span r[M-1]*...*r[M-j]).  Stage s of radix R
// and span P holds w(r, k) = e^{-2 pi i r k n/(P R) / n} at off[s] + (r-1)*P + k.
struct MidTables {
  int count;
  int R[24];
  int P[24];
  int off[25];
};
constexpr MidTables mid_tables(int n) {
  const Radices f = factorize(n);
  MidTables t{};
  const int m = f.count;
  int p = f.r[0];
  for (int j = 1; j + 1 < m; ++j) {
    t.R[t.count] = f.r[j];
    t.P[t.count] = p;
    t.off[t.count + 1] = t.off[t.count] + (f.r[j] - 1) * p;
    ++t.count;
    p *= f.r[j];
  }
  p = f.r[m - 1];
  for (int j = 1; j + 1 < m; ++j) {
    t.R[t.count] = f.r[m - 1 - j];
    t.P[t.count] = p;
    t.off[t.count + 1] = t.off[t.count] + (f.r[m - 1 - j] - 1) * p;
    ++t.count;
    p *= f.r[m - 1 - j];
  }
  return t;
}

// Twiddle pointers of one column transform: the N-entry table (LDS or global)
// and the middle-stage tables (LDS).
struct ColTw {
  const float2* full;
  const float2* mid;
};

template <int N, int LINES, int LS, int NT, int SH>
struct ColumnFFT {
  static constexpr Radices F = factorize(N);
  static constexpr int M = F.count;
  static constexpr MidTables MT = mid_tables(N);
  static constexpr int MID_ENTRIES = MT.off[MT.count];   // float2 entries of the middle-stage tables
  static_assert(M >= 2, "pruned column transforms need at least two radix stages");
  static constexpr int R1 = F.r[0];
  static constexpr int RM = F.r[M - 1];
  static constexpr int T1 = N / R1;          // stride of the first inverse stage
  static constexpr int TM = N / RM;          // stride of the last inverse stage
  static constexpr int B1 = LINES * T1;      // radix-R1 butterflies per block
  static constexpr int BM = LINES * TM;      // radix-RM butterflies per block
  static constexpr int PER1 = (B1 + NT - 1) / NT;
  static constexpr int PERM = (BM + NT - 1) / NT;

  // Fill the middle-stage tables from the N-entry table tw (e^{-2 pi i j / N}
  // at j).  Ends without a barrier: the caller barriers before the first use.
  __device__ __forceinline__ static void fill_mid_tables(float2* mid, const float2* __restrict__ tw) {
    if constexpr (MID_ENTRIES > 0)
    for (int e = threadIdx.x; e < MID_ENTRIES; e += NT) {
      int off = 0, p = MT.P[0], step = N / (MT.P[0] * MT.R[0]);
#pragma unroll
      for (int j = 1; j < MT.count; ++j) {
        if (e >= MT.off[j]) {
          off = MT.off[j];
          p = MT.P[j];
          step = N / (MT.P[j] * MT.R[j]);
        }
      }
      const int loc = e - off;
      const int r = 1 + loc / p, k = loc - (loc / p) * p;
      mid[e] = tw[r * k * step];
    }
  }

  // ---- inverse stage 1 from HBM: src = the window rows of one spectrum
  // (row i at src + i*N), row_of_u[u] = window row of frequency u or -1.
  __device__ __forceinline__ static void inv_first_from_global(float2* buf, const float2* __restrict__ src,
                                                               const int* __restrict__ row_of_u, int x0) {
#pragma unroll
    for (int s = 0; s < PER1; ++s) {
      const int q = threadIdx.x + s * NT;
      if (B1 % NT == 0 || q < B1) {
        const int line = q % LINES, i = q / LINES;
        const int x = x0 + line;
        const int ra = row_of_u[i], rb = row_of_u[(R1 - 1) * T1 + i];
        const bool in = x < N;
        const float2 a = (in && ra >= 0) ? src[(size_t)ra * N + x] : make_float2(0.f, 0.f);
        const float2 c = (in && rb >= 0) ? src[(size_t)rb * N + x] : make_float2(0.f, 0.f);
        store_two_tap<0>(buf + line * LS, i * R1, swz<SH>(i * R1), a, c);
      }
    }
  }
  template <int R>
  __device__ __forceinline__ static void store_two_tap(float2* line, int o, int so, float2 a, float2 c) {
    if constexpr (R < R1) {
      // out[r] = a + c * w^{(R1-1) r}, w = e^{+2 pi i / R1}
      line[sw_at<SH, sw_uniform_write<SH, 1, R1>(), 1>(o, so, R)] = cadd(a, twiddle_ct<(R1 - 1) * R, R1, +1>(c));
      store_two_tap<R + 1>(line, o, so, a, c);
    }
  }

  // ---- middle stages (LDS in place)
  template <int IDX, int P>
  __device__ __forceinline__ static void inv_middle(float2* buf, const float2* mid) {
    if constexpr (IDX < M - 1) {
      stockham_pass<N, F.r[IDX], P, LINES, LS, NT, +1, SH>(buf, PassTwiddles<P>{mid + MT.off[IDX - 1]});
      inv_middle<IDX + 1, P * F.r[IDX]>(buf, mid);
    }
  }
  // forward stages use the reversed radix list: stage j has radix F.r[M-1-j]
  template <int J, int P>
  __device__ __forceinline__ static void fwd_middle(float2* buf, const float2* mid) {
    if constexpr (J < M - 1) {
      stockham_pass<N, F.r[M - 1 - J], P, LINES, LS, NT, -1, SH>(buf, PassTwiddles<P>{mid + MT.off[M - 3 + J]});
      fwd_middle<J + 1, P * F.r[M - 1 - J]>(buf, mid);
    }
  }

  // ---- full pruned inverse: spectrum rows -> registers v[s][r], element
  // (column q % LINES, row i + r*TM) with q = tid + s*NT, i = q / LINES.
  // Ends after the LDS reads; the caller must barrier before reusing buf.
  __device__ __forceinline__ static void inverse_to_regs(float2* buf, const float2* __restrict__ src,
                                                         const int* __restrict__ row_of_u, int x0,
                                                         ColTw tw, float2 (&v)[PERM][RM]) {
    inv_first_from_global(buf, src, row_of_u, x0);
    __syncthreads();
    inv_middle<1, R1>(buf, tw.mid);
#pragma unroll
    for (int s = 0; s < PERM; ++s) {
      const int q = threadIdx.x + s * NT;
      if (BM % NT == 0 || q < BM) {
        const int line = q % LINES, i = q / LINES;
        const float2* p = buf + line * LS;
        const int si = swz<SH>(i);
#pragma unroll
        for (int r = 0; r < RM; ++r) v[s][r] = p[sw_at<SH, sw_uniform_read<SH, TM>(), TM>(i, si, r)];
#pragma unroll
        for (int r = 1; r < RM; ++r) v[s][r] = cmulc(v[s][r], tw.full[r * i]);
        Dft<RM, +1>::run(v[s]);
      }
    }
  }

  // ---- forward transform of registers v (same ownership) with pruned output:
  // dst row ra / rb receives frequency u = i / (R1-1)*T1 + i of column x.
  __device__ __forceinline__ static void forward_from_regs(float2* buf, float2 (&v)[PERM][RM], ColTw tw,
                                                           float2* __restrict__ dst,
                                                           const int* __restrict__ row_of_u, int x0) {
#pragma unroll
    for (int s = 0; s < PERM; ++s) {
      const int q = threadIdx.x + s * NT;
      if (BM % NT == 0 || q < BM) {
        const int line = q % LINES, i = q / LINES;
        Dft<RM, -1>::run(v[s]);
        float2* d = buf + line * LS;
        const int so = swz<SH>(i * RM);
#pragma unroll
        for (int r = 0; r < RM; ++r) d[sw_at<SH, sw_uniform_write<SH, 1, RM>(), 1>(i * RM, so, r)] = v[s][r];
      }
    }
    __syncthreads();
    fwd_middle<1, RM>(buf, tw.mid);
#pragma unroll
    for (int s = 0; s < PER1; ++s) {
      const int q = threadIdx.x + s * NT;
      if (B1 % NT == 0 || q < B1) {
        const int line = q % LINES, i = q / LINES;
        const int x = x0 + line;
        float2 w[R1];
        const float2* p = buf + line * LS;
        const int si = swz<SH>(i);
#pragma unroll
        for (int r = 0; r < R1; ++r) w[r] = p[sw_at<SH, sw_uniform_read<SH, T1>(), T1>(i, si, r)];
#pragma unroll
        for (int r = 1; r < R1; ++r) w[r] = cmul(w[r], tw.full[r * i]);
        float2 o0 = w[0], o1 = w[0];
        sum_two_outputs<1>(w, o0, o1);
        const int ra = row_of_u[i], rb = row_of_u[(R1 - 1) * T1 + i];
        if (x < N) {
          if (ra >= 0) dst[(size_t)ra * N + x] = o0;
          if (rb >= 0) dst[(size_t)rb * N + x] = o1;
        }
      }
    }
  }
  // o0 = sum_r w[r]; o1 = sum_r w[r] e^{-2 pi i r (R1-1) / R1}
  template <int R>
  __device__ __forceinline__ static void sum_two_outputs(const float2 (&w)[R1], float2& o0, float2& o1) {
    if constexpr (R < R1) {
      o0 = cadd(o0, w[R]);
      o1 = cadd(o1, twiddle_ct<R * (R1 - 1), R1, -1>(w[R]));
      sum_two_outputs<R + 1>(w, o0, o1);
    }
  }
};

}  // namespace fft
}  // namespace litho
