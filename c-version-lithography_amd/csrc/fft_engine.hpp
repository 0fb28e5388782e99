// Device FFT engine for gfx950: batched complex FFTs of compile-time length N
// held in LDS, mixed radix (16/8/4/2 and odd radices up to 17), Stockham
// autosort passes so the output comes back in natural order without a bit
// reversal.  A workgroup transforms LINES lines of N complex64 values at once;
// every pass stages its butterflies in registers between two barriers, so the
// transform is in place in one LDS buffer.
//
// Conventions: DIR = -1 is the forward transform sum_n x[n] e^{-2 pi i n m/N},
// DIR = +1 the unnormalised inverse.  The twiddle table tw[m] = e^{-2 pi i m/N},
// m in [0, N), is built on the host in FP64 and rounded to FP32.
#pragma once
#include <hip/hip_runtime.h>

namespace litho {
namespace fft {

// ---------------------------------------------------------------- factorisation
struct Radices {
  int count;
  int r[12];
};

constexpr Radices factorize(int n) {
  Radices f{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int a = 0;
  while (n % 2 == 0) { n /= 2; ++a; }
  // power-of-two part: as many radix-16 passes as possible, remainder 8/4/2
  int q = a / 4, rem = a % 4;
  if (rem == 1 && q >= 1) { --q; f.r[f.count++] = 8; f.r[f.count++] = 4; rem = 0; }
  for (int i = 0; i < q; ++i) f.r[f.count++] = 16;
  if (rem == 3) f.r[f.count++] = 8;
  if (rem == 2) f.r[f.count++] = 4;
  if (rem == 1) f.r[f.count++] = 2;
  // odd part: 3 * 3 as one radix-9 pass (two radix-3 layers in registers), then the primes
  while (n % 9 == 0) { n /= 9; f.r[f.count++] = 9; }
  const int odd[6] = {3, 5, 7, 11, 13, 17};
  for (int i = 0; i < 6; ++i)
    while (n % odd[i] == 0) { n /= odd[i]; f.r[f.count++] = odd[i]; }
  if (n != 1) f.count = -1;  // unsupported prime factor
  return f;
}

constexpr bool supported(int n) { return n >= 2 && factorize(n).count > 0; }

// Per-stage twiddle tables of the line transforms (rows and the pruned
// columns): every stage s >= 1 of the inverse order (radix r[s], span
// P = r[0]*...*r[s-1]) at index s-1, then every stage j >= 1 of the forward
// order (radix r[M-1-j], span r[M-1]*...*r[M-j]) at index M-2+j.  Stage
// table t of radix R and span P holds w(r, k) = e^{-2 pi i r k / (P R)} at
// off[t] + (r-1)*P + k, so lanes with consecutive k read consecutive entries.
struct StageTables {
  int count;
  int R[24];
  int P[24];
  int off[25];
};
constexpr StageTables stage_tables(int n) {
  const Radices f = factorize(n);
  StageTables t{};
  const int m = f.count;
  int p = f.r[0];
  for (int j = 1; j < m; ++j) {
    t.R[t.count] = f.r[j];
    t.P[t.count] = p;
    t.off[t.count + 1] = t.off[t.count] + (f.r[j] - 1) * p;
    ++t.count;
    p *= f.r[j];
  }
  p = f.r[m - 1];
  for (int j = 1; j < m; ++j) {
    t.R[t.count] = f.r[m - 1 - j];
    t.P[t.count] = p;
    t.off[t.count + 1] = t.off[t.count] + (f.r[m - 1 - j] - 1) * p;
    ++t.count;
    p *= f.r[m - 1 - j];
  }
  return t;
}

// ------------------------------------------------------- compile-time constants
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_taylor(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 30; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}

// cos / sin of 2 pi m / R, exact at multiples of pi/4
constexpr double ct_cos2pi(int m, int r) {
  m = ((m % r) + r) % r;
  if ((8 * m) % r == 0) {
    const int o = 8 * m / r;  // octant 0..7
    constexpr double h = 0.70710678118654752440084436210484903;
    const double c[8] = {1.0, h, 0.0, -h, -1.0, -h, 0.0, h};
    return c[o];
  }
  double x = 2.0 * kPi * m / r;
  if (x > kPi) x -= 2.0 * kPi;
  return ct_sin_taylor(x + kPi / 2 > kPi ? x + kPi / 2 - 2.0 * kPi : x + kPi / 2);
}

constexpr double ct_sin2pi(int m, int r) {
  m = ((m % r) + r) % r;
  if ((8 * m) % r == 0) {
    const int o = 8 * m / r;
    constexpr double h = 0.70710678118654752440084436210484903;
    const double s[8] = {0.0, h, 1.0, h, 0.0, -h, -1.0, -h};
    return s[o];
  }
  double x = 2.0 * kPi * m / r;
  if (x > kPi) x -= 2.0 * kPi;
  return ct_sin_taylor(x);
}

// ------------------------------------------------------------- complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Barrier that orders the block's LDS accesses only.  __syncthreads() also
// waits for the wave's outstanding global loads and stores, which pins a
// kernel's prefetch of the next line (and its stores of the last) to the first
// exchange of every transform.  For the transforms whose callers never read
// back, within the kernel, global memory that another thread wrote.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// multiply by e^{DIR * 2 pi i m / R} with compile-time constants
template <int M, int R, int DIR>
__device__ __forceinline__ float2 twiddle_ct(float2 a) {
  constexpr int mm = ((M % R) + R) % R;
  if constexpr (mm == 0) {
    return a;
  } else if constexpr (4 * mm == R) {          // angle pi/2: multiply by DIR*i
    return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
  } else if constexpr (2 * mm == R) {          // angle pi
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * mm == 3 * R) {      // angle 3pi/2
    return DIR > 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
  } else {
    constexpr float c = (float)ct_cos2pi(mm, R);
    constexpr float s = (float)(DIR * ct_sin2pi(mm, R));
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// ------------------------------------------------------------- small DFTs
template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const float2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    const float2 t = twiddle_ct<1, 4, DIR>(d13);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, t);
    v[3] = csub(d02, t);
  }
};

// R = A * B by the two-factor (Cooley-Tukey) decomposition:
// X[k1 + A k2] = sum_n2 W_B^{n2 k2} [ W_R^{n2 k1} sum_n1 x[B n1 + n2] W_A^{n1 k1} ]
template <int A, int B, int DIR>
struct DftAB {
  static constexpr int R = A * B;
  __device__ __forceinline__ static void run(float2* v) {
    float2 t[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) t[n2][n1] = v[B * n1 + n2];
      Dft<A, DIR>::run(t[n2]);
    }
    apply_tw<1>(t);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      float2 u[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) u[n2] = t[n2][k1];
      Dft<B, DIR>::run(u);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = u[k2];
    }
  }
  template <int N2>
  __device__ __forceinline__ static void apply_tw(float2 (&t)[B][A]) {
    if constexpr (N2 < B) {
      apply_tw_k<N2, 1>(t);
      apply_tw<N2 + 1>(t);
    }
  }
  template <int N2, int K1>
  __device__ __forceinline__ static void apply_tw_k(float2 (&t)[B][A]) {
    if constexpr (K1 < A) {
      t[N2][K1] = twiddle_ct<N2 * K1, R, DIR>(t[N2][K1]);
      apply_tw_k<N2, K1 + 1>(t);
    }
  }
};

template <int DIR>
struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(float2* v) { DftAB<4, 2, DIR>::run(v); }
};

template <int DIR>
struct Dft<16, DIR> {
  __device__ __forceinline__ static void run(float2* v) { DftAB<4, 4, DIR>::run(v); }
};

template <int DIR>
struct Dft<9, DIR> {
  __device__ __forceinline__ static void run(float2* v) { DftAB<3, 3, DIR>::run(v); }
};

// odd radix: direct O(R^2) with compile-time constants, using the symmetric
// pairing y[q], y[R-q] to halve the multiplies
template <int R, int DIR>
struct Dft {
  static_assert(R % 2 == 1, "generic DFT is for odd radices");
  __device__ __forceinline__ static void run(float2* v) {
    constexpr int H = (R - 1) / 2;
    float2 s[H + 1], d[H + 1];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      s[r] = cadd(v[r], v[R - r]);
      d[r] = csub(v[r], v[R - r]);
    }
    float2 y0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; ++r) y0 = cadd(y0, s[r]);
    float2 out[R];
    out[0] = y0;
    fill<1>(v[0], s, d, out);
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = out[r];
  }
  template <int Q>
  __device__ __forceinline__ static void fill(float2 x0, const float2* s, const float2* d, float2* out) {
    if constexpr (Q <= (R - 1) / 2) {
      float re = x0.x, im = x0.y, pr = 0.f, pi = 0.f;
      accum<Q, 1>(s, d, re, im, pr, pi);
      // y[Q] = A + i*B, y[R-Q] = A - i*B where A = x0 + sum c s, B = DIR * sum sn d
      out[Q] = make_float2(re - pi, im + pr);
      out[R - Q] = make_float2(re + pi, im - pr);
      fill<Q + 1>(x0, s, d, out);
    }
  }
  template <int Q, int P>
  __device__ __forceinline__ static void accum(const float2* s, const float2* d, float& re, float& im,
                                               float& pr, float& pi) {
    if constexpr (P <= (R - 1) / 2) {
      constexpr float c = (float)ct_cos2pi(Q * P, R);
      constexpr float sn = (float)(DIR * ct_sin2pi(Q * P, R));
      re += c * s[P].x;
      im += c * s[P].y;
      pr += sn * d[P].x;
      pi += sn * d[P].y;
      accum<Q, P + 1>(s, d, re, im, pr, pi);
    }
  }
};

// ------------------------------------------------------------------ LDS layout
// gfx950 serves a 64-bit LDS access (each address of a read2 / write2 on its
// own) 16 lanes at a time over 16 eight-byte bank slots; lanes of a group that
// hit the same slot modulo 16 at different addresses serialise (measured with
// SQ_LDS_BANK_CONFLICT on probe patterns).  Element p of a line lives at
// swz<SH>(p) = p + (p >> SH): one pad slot per 2^SH complex values, so strided
// Stockham writes spread over the slots.  Line strides are
// padded_stride<SH>(N) = N + N/2^SH + 4.  The shift is 4 for the column kernels
// (lanes spread across lines) and 5 for the row kernels (lanes along one
// line); row shifts 3, 4 or 6, column shift 5 and pads of 1 or 2 instead of 4
// time the same within 1 % on the 640 ... 2048 tiles, whose transforms are not
// bound by these conflicts.
template <int SH>
__host__ __device__ constexpr int swz(int p) { return p + (p >> SH); }
template <int SH>
__host__ __device__ constexpr int padded_stride(int n) { return n + (n >> SH) + 4; }

// Swizzled address of element base + r*S (r = 0..R-1) given swz(base).  When the
// shifted elements never straddle a 2^SH-slot boundary differently for
// different lanes, the offset is the compile-time constant r*S + (r*S >> SH)
// and folds into the LDS instruction's immediate; otherwise it is computed per
// element.  With B = 2^SH:
//   reads  base + r*T           : uniform when T % B == 0
//   writes (i-k)*R + k + r*P    : uniform when P % B == 0, or B % P == 0 and
//                                 (R*P % B == 0 or B % (R*P) == 0)
// ((i-k)*R is a multiple of R*P; with P | B the B-slot blocks are unions of
// P-blocks, so k never moves an element across a block boundary.)
template <int SH, int S>
__host__ __device__ constexpr bool sw_uniform_read() { return S % (1 << SH) == 0; }
template <int SH, int P, int R>
__host__ __device__ constexpr bool sw_uniform_write() {
  constexpr int B = 1 << SH;
  return P % B == 0 || (B % P == 0 && ((R * P) % B == 0 || B % (R * P) == 0));
}
template <int SH, bool UNIFORM, int S>
__device__ __forceinline__ int sw_at(int base, int sw_base, int r) {
  if constexpr (UNIFORM) return sw_base + r * S + ((r * S) >> SH);
  else return swz<SH>(base + r * S);
}

// Twiddle sources of a Stockham pass: w(r, k, step) = e^{-2 pi i r k step / N}.
struct TableTwiddles {          // the N-entry table (HBM / L2 or an LDS copy)
  const float2* t;
  __device__ __forceinline__ float2 operator()(int r, int k, int step) const { return t[r * k * step]; }
};
template <int P>
struct PassTwiddles {           // one pass's (R-1) x P table: lanes with consecutive k read consecutive slots
  const float2* t;
  __device__ __forceinline__ float2 operator()(int r, int k, int) const { return t[(r - 1) * P + k]; }
};

// ------------------------------------------------------------------ passes
// One Stockham pass of radix R and span P over LINES lines of length N with
// line stride LS (in float2), NT threads and swizzle shift SH.  Input position
// i + r*(N/R); output position (i - k)*R + k + r*P with k = i mod P.
template <int N, int R, int P, int LINES, int LS, int NT, int DIR, int SH, class TW>
__device__ __forceinline__ void stockham_pass(float2* buf, const TW& tw) {
  constexpr int T = N / R;
  constexpr int TOTAL = T * LINES;
  constexpr int PER = (TOTAL + NT - 1) / NT;
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int q = tid + s * NT;
    if (TOTAL % NT == 0 || q < TOTAL) {
      const int line = q / T;
      const int i = q - line * T;
      const float2* src = buf + line * LS;
      const int si = swz<SH>(i);
#pragma unroll
      for (int r = 0; r < R; ++r) v[s][r] = src[sw_at<SH, sw_uniform_read<SH, T>(), T>(i, si, r)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int q = tid + s * NT;
    if (TOTAL % NT == 0 || q < TOTAL) {
      const int line = q / T;
      const int i = q - line * T;
      const int k = i % P;
      if constexpr (P > 1) {
        constexpr int STEP = N / (P * R);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const float2 w = tw(r, k, STEP);
          v[s][r] = DIR < 0 ? cmul(v[s][r], w) : cmulc(v[s][r], w);
        }
      }
      Dft<R, DIR>::run(v[s]);
      float2* dst = buf + line * LS;
      const int o = (i - k) * R + k;
      const int so = swz<SH>(o);
#pragma unroll
      for (int r = 0; r < R; ++r) dst[sw_at<SH, sw_uniform_write<SH, P, R>(), P>(o, so, r)] = v[s][r];
    }
  }
  __syncthreads();
}

template <int N, int IDX, int P, int LINES, int LS, int NT, int DIR, int SH>
__device__ __forceinline__ void run_passes(float2* buf, const float2* __restrict__ tw) {
  constexpr Radices f = factorize(N);
  if constexpr (IDX < f.count) {
    constexpr int R = f.r[IDX];
    stockham_pass<N, R, P, LINES, LS, NT, DIR, SH>(buf, TableTwiddles{tw});
    run_passes<N, IDX + 1, P * R, LINES, LS, NT, DIR, SH>(buf, tw);
  }
}

// The same passes with the twiddles of stage s >= 1 read from the stage
// tables of stage_tables(N) (the inverse-order half: index s - 1), which
// serve both directions (the forward transform applies the conjugates).
template <int N, int IDX, int P, int LINES, int LS, int NT, int DIR, int SH>
__device__ __forceinline__ void run_passes_staged(float2* buf, const float2* tabs) {
  constexpr Radices f = factorize(N);
  constexpr StageTables st = stage_tables(N);
  if constexpr (IDX < f.count) {
    constexpr int R = f.r[IDX];
    if constexpr (IDX == 0)
      stockham_pass<N, R, P, LINES, LS, NT, DIR, SH>(buf, TableTwiddles{tabs});   // P = 1: no twiddles read
    else
      stockham_pass<N, R, P, LINES, LS, NT, DIR, SH>(buf, PassTwiddles<P>{tabs + st.off[IDX - 1]});
    run_passes_staged<N, IDX + 1, P * R, LINES, LS, NT, DIR, SH>(buf, tabs);
  }
}
template <int N, int LINES, int LS, int NT, int DIR, int SH>
__device__ __forceinline__ void fft_lines_staged(float2* buf, const float2* tabs) {
  static_assert(factorize(N).count > 0, "unsupported FFT length");
  run_passes_staged<N, 0, 1, LINES, LS, NT, DIR, SH>(buf, tabs);
}
// Entries of the inverse-order half of stage_tables(n).
constexpr int staged_entries(int n) { return stage_tables(n).off[factorize(n).count - 1]; }

// Transform LINES lines in LDS in place.  The caller must __syncthreads()
// after writing the input; the routine ends with a barrier.
template <int N, int LINES, int LS, int NT, int DIR, int SH>
__device__ __forceinline__ void fft_lines(float2* buf, const float2* __restrict__ tw) {
  static_assert(factorize(N).count > 0, "unsupported FFT length");
  run_passes<N, 0, 1, LINES, LS, NT, DIR, SH>(buf, tw);
}

}  // namespace fft
}  // namespace litho
