// Optimizer kernels (opt_kernels.hip): theta init, L-BFGS pair / direction,
// FP64 fixed-order reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace litho {

constexpr int kOptThreads = 256;
#ifndef LITHO_OPT_BLOCKS
#define LITHO_OPT_BLOCKS 1024
#endif
constexpr int kOptBlocks = LITHO_OPT_BLOCKS;  // grid-stride blocks per tile of the L-BFGS vector kernels
constexpr int kMaxHist = 12;             // largest supported L-BFGS history m
constexpr int kPairDots = 4 * (kMaxHist + 1);
constexpr int kCoefStride = 1 + 2 * kMaxHist;
// L-BFGS state of a tile: its accepted iterates and their gradients as
// points, not as differences.  Tile b owns lbfgs_points(m_max) point slots in
// X and in G (point p of tile b at (b * points + p) * n floats).  Its
// kMetaStride ints of meta (chronological):
//   [0] m, the pairs held
//   [1] base point: the accepted iterate, origin of the line search
//   [2] trial point: the line search writes its trials (and their gradients) there
//   [3 + j], [3 + kMaxHist + j]: start and end point of pair j
// Pair j is s_j = X[end] - X[start], y_j = G[end] - G[start], formed in FP32 by
// the kernels that use it.  Consecutive pairs share a point unless a pair was
// rejected in between, so the kept pairs, the base and the trial use at most
// 2 m_max + 2 points.
constexpr int kMetaStride = 3 + 2 * kMaxHist;
constexpr int kMetaBase = 1, kMetaTrial = 2, kMetaStart = 3, kMetaEnd = 3 + kMaxHist;
__host__ __device__ constexpr int lbfgs_points(int m_max) { return 2 * m_max + 2; }

void launch_init_theta(const float* pattern, float* theta, size_t count, hipStream_t s);
void launch_scale(const float* x, float* y, const float* coef, const int* active, size_t n, int batch,
                  hipStream_t s);
void launch_axpy(const float* x, float* y, const float* coef, const int* active, size_t n, int batch,
                 hipStream_t s);
void launch_mask_from_theta(const float* theta, float* mask, size_t count, hipStream_t s);
// L-BFGS kernels over the tiles b with need[b] != 0 (points: point slots per
// tile, meta as above); mb >= m is the instantiation.  The pair kernel forms
// the new pair (base -> trial) and its FP64 dot partials with the m held
// pairs and the trial's gradient, layout per tile and block (NV = 4 (mb + 1)):
//   [4j+0] s_j . y_new   [4j+1] y_j . y_new   [4j+2] s_j . g   [4j+3] y_j . g
// for j < m, then the new pair at position 4 mb.  It writes no vectors.
void launch_lbfgs_pair(const float* X, const float* G, int points, const int* need, const int* meta, size_t n,
                       int batch, int mb, double* partials, hipStream_t s);
// History update and compact-form coefficients from the pair kernel's reduced
// dots (dots[b * kPairDots + ...], layout of the pair kernel at mb): accepts
// the new pair when s.y > eps y.y, drops the oldest pair beyond m_max, keeps
// the m x m matrices S'Y (upper triangle) and Y'Y in sy / yy
// ([b][kMaxHist][kMaxHist] doubles) and writes coef ([b][kCoefStride]).  The
// trial becomes the base (the accepted iterate, whether or not its pair is
// kept) and the next trial a point that no kept pair uses.
void launch_lbfgs_update(int batch, const int* need, const double* dots, int mb, int m_max, int* meta, double* sy,
                         double* yy, double* coef, hipStream_t s);
// d = cg G[base] + sum_j cs_j s_j + cy_j y_j over the m <= mb held pairs, and
// the FP64 partials of G[base] . d ([b][kOptBlocks]).
void launch_lbfgs_direction(const float* X, const float* G, int points, const int* need, const int* meta,
                            const double* coef, size_t n, float* d, int batch, int mb, double* partials, hipStream_t s);
// A new tile b: m = 0, base point 0 (which the caller fills with theta0), trial point 1.
void launch_lbfgs_reset(int* meta, int b, hipStream_t s);
// Tiles with flags[b] != 0 after their first evaluation (theta0 and its
// gradient in the trial point): the trial becomes the base, the old base the
// trial, and d = -G[base] (steepest descent, the first direction).
void launch_lbfgs_start(const float* G, int points, const int* flags, int* meta, size_t n, int batch, float* d,
                        hipStream_t s);
// Point `which` (kMetaBase or kMetaTrial) of tile b: copied to out_theta
// (optional) and rendered to out_mask = (1 + cos theta) / 2.
void launch_lbfgs_point_out(const float* X, int points, const int* meta, int b, int which, size_t n,
                            float* out_theta, float* out_mask, hipStream_t s);
// A plain kernel reading `reads` and writing `writes` vectors of n4 float4
// with the L-BFGS kernels' grid and load policy (the rate their access
// pattern reaches);
// (reads, writes) one of (24, 0), (22, 1), (24, 4), (21, 1), else false.
// Reads n4 float4 of buf (one word each, plain loads) to evict the
// last-level cache before a timed launch.
void launch_cache_evict(const float* buf, size_t n4, float* sink, hipStream_t s);
bool launch_stream_pattern(int reads, int writes, const float* const* in, float* const* out, size_t n4, float* sink,
                           hipStream_t s);
// out[b * out_stride + v] = sum over blk of partials[(b * nblk + blk) * nv + v] for
// v < nv, summed in a fixed tree order (deterministic, independent of scheduling).
// With active set, only the tiles with active[b] != 0 are written.
void launch_reduce_fixed(const double* partials, int nblk, int nv, double* out, int out_stride, int batch,
                         hipStream_t s, const int* active = nullptr);
// An evaluation's four reductions, out[b * 4 + v]: v = 0, 1 from the image
// partials ([b][img_blk][2]), v = 2, 3 from the gradient partials
// ([b][grad_blk][2]), each in the order of launch_reduce_fixed.  With extra
// set, also extra_out[b] = the sum of extra's partials ([b][extra_blk], one
// value each) in the same order.
void launch_reduce_eval(const double* img, int img_blk, const double* grad, int grad_blk, double* out, int batch,
                        hipStream_t s, const int* active = nullptr, const double* extra = nullptr,
                        int extra_blk = 0, double* extra_out = nullptr);

}  // namespace litho
