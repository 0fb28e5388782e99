// Window-pruned line transforms for the row-direction SOCS kernels.
//
// A block transforms LINES lines of length N held in LDS; lanes run along a
// line (butterfly q maps to line q / T and index q % T), so the first and last
// stages touch HBM in coalesced runs.  Like the column transforms, the inverse
// uses the radix order (R_1, ..., R_M) and the forward the reversed order:
//  * inverse stage 1 (radix R_1, span 1) sees at most two nonzero inputs per
//    butterfly — spectra are nonzero only in the pupil window, the first and
//    the last N/R_1 frequencies — which the caller supplies per frequency; no
//    LDS zero fill;
//  * the last inverse stage hands its outputs (line positions i + r*N/R_M) to
//    the caller from registers instead of writing them back to LDS;
//  * forward stage 1 (radix R_M, span 1) takes its inputs from the caller
//    (positions i + r*N/R_M, one call per element);
//  * the last forward stage (radix R_1) computes only the two outputs per
//    butterfly that can lie in the window, frequencies i and N - N/R_1 + i.
// Every stage after the first reads its twiddles from the stage tables of
// stage_tables() (global memory, L1/L2 resident), laid out so consecutive
// lanes read consecutive entries.
// Precondition (checked on the host): every window frequency u satisfies
// u < N/R_1 or u >= N - N/R_1.
#pragma once
#include "fft_engine.hpp"

namespace litho {
namespace fft {

template <int N, int LINES, int LS, int NT, int SH>
struct RowFFT {
  static constexpr Radices F = factorize(N);
  static constexpr int M = F.count;
  static_assert(M >= 2, "pruned row transforms need at least two radix stages");
  static constexpr StageTables ST = stage_tables(N);
  static constexpr int R1 = F.r[0];
  static constexpr int RM = F.r[M - 1];
  static constexpr int T1 = N / R1;
  static constexpr int TM = N / RM;
  static constexpr int B1 = LINES * T1;
  static constexpr int BM = LINES * TM;
  static constexpr int PER1 = (B1 + NT - 1) / NT;
  static constexpr int PERM = (BM + NT - 1) / NT;
  static constexpr int INV_LAST = M - 2;          // stage table of the last inverse stage
  static constexpr int FWD_LAST = 2 * (M - 1) - 1; // stage table of the last forward stage

  template <int IDX, int P>
  __device__ __forceinline__ static void inv_middle(float2* buf, const float2* __restrict__ tabs) {
    if constexpr (IDX < M - 1) {
      stockham_pass<N, F.r[IDX], P, LINES, LS, NT, +1, SH>(buf, PassTwiddles<P>{tabs + ST.off[IDX - 1]});
      inv_middle<IDX + 1, P * F.r[IDX]>(buf, tabs);
    }
  }
  template <int J, int P>
  __device__ __forceinline__ static void fwd_middle(float2* buf, const float2* __restrict__ tabs) {
    if constexpr (J < M - 1) {
      stockham_pass<N, F.r[M - 1 - J], P, LINES, LS, NT, -1, SH>(buf, PassTwiddles<P>{tabs + ST.off[M - 2 + J]});
      fwd_middle<J + 1, P * F.r[M - 1 - J]>(buf, tabs);
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
  template <int R>
  __device__ __forceinline__ static void sum_two_outputs(const float2 (&w)[R1], float2& o0, float2& o1) {
    if constexpr (R < R1) {
      o0 = cadd(o0, w[R]);
      o1 = cadd(o1, twiddle_ct<R * (R1 - 1), R1, -1>(w[R]));
      sum_two_outputs<R + 1>(w, o0, o1);
    }
  }

  // Unnormalised inverse of LINES sparse spectra.  in(line, u) returns the
  // spectrum at frequency u (called for u < N/R_1 and u >= N - N/R_1 only);
  // out(line, x, value) receives every output position x exactly once.  Ends
  // after the last LDS reads: the caller barriers before reusing buf.
  template <class In, class Out>
  __device__ __forceinline__ static void inverse(float2* buf, const float2* __restrict__ tabs, In in, Out out) {
#pragma unroll
    for (int s = 0; s < PER1; ++s) {
      const int q = threadIdx.x + s * NT;
      if (B1 % NT == 0 || q < B1) {
        const int line = q / T1, i = q - line * T1;
        const float2 a = in(line, i);
        const float2 c = in(line, (R1 - 1) * T1 + i);
        store_two_tap<0>(buf + line * LS, i * R1, swz<SH>(i * R1), a, c);
      }
    }
    __syncthreads();
    inv_middle<1, R1>(buf, tabs);
    const float2* lt = tabs + ST.off[INV_LAST];
#pragma unroll
    for (int s = 0; s < PERM; ++s) {
      const int q = threadIdx.x + s * NT;
      if (BM % NT == 0 || q < BM) {
        const int line = q / TM, i = q - line * TM;
        const float2* p = buf + line * LS;
        const int si = swz<SH>(i);
        float2 v[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r) v[r] = p[sw_at<SH, sw_uniform_read<SH, TM>(), TM>(i, si, r)];
#pragma unroll
        for (int r = 1; r < RM; ++r) v[r] = cmulc(v[r], lt[(r - 1) * TM + i]);
        Dft<RM, +1>::run(v);
#pragma unroll
        for (int r = 0; r < RM; ++r) out(line, i + r * TM, v[r]);
      }
    }
  }

  // Forward transform of LINES full lines with pruned output.  in(line, x)
  // is called once per input position; out(line, u, value) receives the
  // frequencies u < N/R_1 and u >= N - N/R_1.  Ends after the last LDS reads.
  template <class In, class Out>
  __device__ __forceinline__ static void forward(float2* buf, const float2* __restrict__ tabs, In in, Out out) {
#pragma unroll
    for (int s = 0; s < PERM; ++s) {
      const int q = threadIdx.x + s * NT;
      if (BM % NT == 0 || q < BM) {
        const int line = q / TM, i = q - line * TM;
        float2 v[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r) v[r] = in(line, i + r * TM);
        first_store(buf, line, i, v);
      }
    }
    forward_tail(buf, tabs, out);
  }

  // forward() with the input split in two: load(line, x) returns the raw
  // values of a position and make(line, x, raw) the transform's input.  All
  // loads of a pass are issued before the first make, so their latency is
  // paid once per pass even when make branches (cosf) or stores.
  template <class Load, class Make, class Out>
  __device__ __forceinline__ static void forward_split(float2* buf, const float2* __restrict__ tabs, Load load,
                                                       Make make, Out out) {
#pragma unroll
    for (int s = 0; s < PERM; ++s) {
      const int q = threadIdx.x + s * NT;
      if (BM % NT == 0 || q < BM) {
        const int line = q / TM, i = q - line * TM;
        decltype(load(0, 0)) raw[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r) raw[r] = load(line, i + r * TM);
        float2 v[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r) v[r] = make(line, i + r * TM, raw[r]);
        first_store(buf, line, i, v);
      }
    }
    forward_tail(buf, tabs, out);
  }

  __device__ __forceinline__ static void first_store(float2* buf, int line, int i, float2 (&v)[RM]) {
    Dft<RM, -1>::run(v);
    float2* d = buf + line * LS;
    const int so = swz<SH>(i * RM);
#pragma unroll
    for (int r = 0; r < RM; ++r) d[sw_at<SH, sw_uniform_write<SH, 1, RM>(), 1>(i * RM, so, r)] = v[r];
  }

  template <class Out>
  __device__ __forceinline__ static void forward_tail(float2* buf, const float2* __restrict__ tabs, Out out) {
    __syncthreads();
    fwd_middle<1, RM>(buf, tabs);
    const float2* lt = tabs + ST.off[FWD_LAST];
#pragma unroll
    for (int s = 0; s < PER1; ++s) {
      const int q = threadIdx.x + s * NT;
      if (B1 % NT == 0 || q < B1) {
        const int line = q / T1, i = q - line * T1;
        const float2* p = buf + line * LS;
        const int si = swz<SH>(i);
        float2 w[R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) w[r] = p[sw_at<SH, sw_uniform_read<SH, T1>(), T1>(i, si, r)];
#pragma unroll
        for (int r = 1; r < R1; ++r) w[r] = cmul(w[r], lt[(r - 1) * T1 + i]);
        float2 o0 = w[0], o1 = w[0];
        sum_two_outputs<1>(w, o0, o1);
        out(line, i, o0);
        out(line, (R1 - 1) * T1 + i, o1);
      }
    }
  }
};

}  // namespace fft
}  // namespace litho
