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
constexpr int kMetaStride = 2 + kMaxHist;   // [m, new_slot, slots[kMaxHist]]
constexpr int kCoefStride = 1 + 2 * kMaxHist;
// Floats between consecutive L-BFGS history slots of one tile, and per tile
// (kMaxHist + 1 slots: m pairs plus the spare that receives the next one).
__host__ __device__ constexpr size_t hist_slot_stride(size_t n) { return n; }
__host__ __device__ constexpr size_t hist_tile_stride(size_t n) { return (size_t)(kMaxHist + 1) * hist_slot_stride(n); }

void launch_init_theta(const float* pattern, float* theta, size_t count, hipStream_t s);
void launch_scale(const float* x, float* y, const float* coef, const int* active, size_t n, int batch,
                  hipStream_t s);
void launch_axpy(const float* x, float* y, const float* coef, const int* active, size_t n, int batch,
                 hipStream_t s);
void launch_mask_from_theta(const float* theta, float* mask, size_t count, hipStream_t s);
// L-BFGS vector kernels over the tiles b with need[b] != 0.  meta holds the
// device-resident history of each tile ([m, spare slot, slots[kMaxHist]] per
// kMetaStride ints, chronological); mb >= m is the instantiation (the pair
// kernel writes the new pair's dots at position 4 mb).  When shift is set the
// pair kernel also stores xp = x and gp = g after forming the new pair.
// Vectors are read as float4: n must be a multiple of 4.
void launch_lbfgs_pair(const float* x, float* xp, const float* g, float* gp, float* hs, float* hy, const int* need,
                       const int* meta, size_t n, int batch, int mb, bool shift, double* partials, hipStream_t s);
// History update and compact-form coefficients from the pair kernel's reduced
// dots (dots[b * kPairDots + ...], layout of the pair kernel at mb): accepts
// the new pair when s.y > eps y.y, drops the oldest pair beyond m_max, keeps
// the m x m matrices S'Y (upper triangle) and Y'Y in sy / yy
// ([b][kMaxHist][kMaxHist] doubles) and writes coef ([b][kCoefStride]).
void launch_lbfgs_update(int batch, const int* need, const double* dots, int mb, int m_max, int* meta, double* sy,
                         double* yy, double* coef, hipStream_t s);
void launch_lbfgs_direction(const float* g, const float* hs, const float* hy, const int* need, const int* meta,
                            const double* coef, size_t n, float* d, int batch, int mb, double* partials, hipStream_t s);
// out[b * out_stride + v] = sum over blk of partials[(b * nblk + blk) * nv + v] for
// v < nv, summed in a fixed tree order (deterministic, independent of scheduling).
void launch_reduce_fixed(const double* partials, int nblk, int nv, double* out, int out_stride, int batch,
                         hipStream_t s);

}  // namespace litho
