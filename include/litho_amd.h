/*
 * litho_amd.h — C ABI of the MI355X-native inverse-lithography hot path.
 *
 * Drop-in replacement for the FFTW/OpenMP path of wyhcoder/C-version-Lithography
 * ("Litho_cpp").  Every entry point names the reference interface it replaces.
 * Plain C: opaque handles, plain pointers and sizes.  Unless stated otherwise,
 * image buffers are DEVICE pointers to row-major N x N float32 arrays
 * (row = y, column = x, element (y, x) at y*N + x); a batch of B tiles is B
 * such arrays back to back.  Scalars that the reference keeps in FP64 (losses,
 * norms, optimizer scalars) are double.
 *
 * Status codes: 0 on success, otherwise one of LITHO_E*; litho_last_error()
 * returns the calling thread's message.  Argument checks mirror the
 * std::invalid_argument checks of imaging.cpp:68-107, gradient.cpp:47-54 and
 * CTM_Optimizer.cpp:33-43.
 *
 * Threading: one host thread per context at a time (the reference's "one caller
 * per Imaging/Gradient object").  Plans are bound to the context that made them.
 */
#ifndef LITHO_AMD_H
#define LITHO_AMD_H

#include <stddef.h>

/* The library is built with hidden symbol visibility: exactly these entry
 * points are exported. */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define LITHO_OK 0
#define LITHO_EINVAL 1        /* std::invalid_argument in the reference */
#define LITHO_ENOMEM 2        /* allocation failure / budget exceeded */
#define LITHO_EDEVICE 3       /* HIP runtime error */
#define LITHO_EUNSUPPORTED 4  /* tile size without a compiled FFT */
#define LITHO_ERUNTIME 5      /* std::runtime_error in the reference */
#define LITHO_ETIMEDOUT 6     /* litho_server_wait: the request is not finished yet */

typedef struct litho_ctx litho_ctx;
typedef struct litho_plan litho_plan;
typedef struct litho_ep_set litho_ep_set;

/* Message of the last failing call on this thread ("" if none). */
const char* litho_last_error(void);

/* 1 if tile size n has a compiled device FFT, else 0.  Plans of other sizes
 * (any N up to 4096 whose window satisfies 2 W_r - 1 <= min(N, 1023)) run the
 * direct-DFT pipeline instead. */
int litho_supported_size(int n);

/* Device context: one HIP device and one stream (default priority). */
int litho_ctx_create(int device, litho_ctx** out);
/* Fails (LITHO_EINVAL, nothing destroyed) while plans made on the context exist. */
int litho_ctx_destroy(litho_ctx* ctx);
int litho_ctx_synchronize(litho_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work on it. */
void* litho_ctx_stream(litho_ctx* ctx);
/* Optimizer loops replay each repeated evaluation from a captured hipGraph
 * (default on); 0 launches the same kernels one by one.  Results are identical. */
int litho_ctx_set_graphs(litho_ctx* ctx, int enable);
/* Evaluation graphs the optimizer pools of this context have built so far
 * (no reference counterpart): captured and replayed, failed to capture and
 * launched kernel by kernel instead, and launched directly because graphs are
 * off.  Any pointer may be null. */
int litho_ctx_graph_counts(litho_ctx* ctx, long long* captured, long long* failed, long long* direct);
/* HIP-event timing of the L-BFGS pair kernel inside the L-BFGS solves run on
 * this context (litho_ctm_optimize, litho_batch_ctm, the server; no reference
 * counterpart): with enable != 0 each round's pair-kernel launch over a full
 * history (m pairs held: the steady-state instantiation) is bracketed by two
 * events on the context stream and its duration accumulated.  The call
 * resets the totals; litho_ctx_pair_kernel_time reads them (nullable). */
int litho_ctx_time_pair_kernel(litho_ctx* ctx, int enable);
int litho_ctx_pair_kernel_time(litho_ctx* ctx, long long* launches, double* total_ms);

/* ---------------------------------------------------------------- optics
 * Compact SOCS frequency window of the reference optics
 * (Grid grid.cpp:4-48, Pupil pupil.cpp:14-46, Source source.cpp:114-158,
 *  LithoPrepare::_extract_source_points litho_prepare.cpp:65-102 and the
 *  compact window litho_prepare.cpp:238-256).  Host only, FP64. */
int litho_socs_window(int n, double pixel_nm, double wavelength_nm, double na, double sigma_in,
                      double sigma_out, int* win_rows, int* win_cols, int* win_y0, int* win_x0,
                      int* num_source_points, double* weight_sum);

/* One-time SOCS kernel precompute, LithoPrepare(grid, pupil, source, use_socs=true,
 * compact_socs=true) with K as a parameter (litho_prepare.cpp:16-54,225-385).
 * FP64 on the device: A (window pixels x source points) is built by a kernel,
 * G = A^H A by rocBLAS, the top-K eigenpairs of G by rocSOLVER, U_k = A v_k / s_k.
 * kernels_out: host, k*win_rows*win_cols complex128 (re, im interleaved),
 * element (k, r, c) at centred frequency index (win_y0 + r, win_x0 + c).
 * weights_out: host, k values lambda_k = s_k^2.  singular_values_out (nullable):
 * host, num_source_points values, descending. */
int litho_socs_precompute(litho_ctx* ctx, int n, double pixel_nm, double wavelength_nm, double na,
                          double sigma_in, double sigma_out, int k, double* kernels_out,
                          double* weights_out, double* singular_values_out);

/* Relative eigen-gap at the truncation, (lambda_K - lambda_{K+1}) / lambda_K, of
 * the last litho_socs_precompute[_zernike] call on this thread (1 when K is the
 * number of source points; 0 before any call).  The reference keeps the top K
 * singular vectors without a check (litho_prepare.cpp:298-330); here a K that
 * splits a degenerate eigenvalue (gap below 1e-9, as the symmetric pairs of an
 * annular source give) fails with LITHO_EINVAL, since its kernel set would
 * depend on the solver's basis inside the eigenspace. */
double litho_socs_last_eigen_gap(void);

/* The same with pupil aberrations, Pupil (pupil.cpp:5-46) with
 * ZernikeGenerator::wavefront (zernike.cpp:11-53): P = [rho^2 <= 1]
 * exp(i 2 pi sum_j c_j Z_j) with fringe Zernike Z_1..Z_11 on the normalised
 * pupil coordinates; num_terms (index, coefficient-in-waves) pairs, summed in
 * the given order, zero coefficients skipped. */
int litho_socs_precompute_zernike(litho_ctx* ctx, int n, double pixel_nm, double wavelength_nm, double na,
                                  double sigma_in, double sigma_out, const int* zernike_index,
                                  const double* zernike_coeff, int num_terms, int k, double* kernels_out,
                                  double* weights_out, double* singular_values_out);

/* ---------------------------------------------------------------- plan
 * Replaces ImagingCache (litho_prepare.h:18-47) plus the FFTW plans of Imaging
 * (imaging.cpp:20-51) and Gradient (gradient.cpp:14-25).  kernels/weights are
 * host arrays as produced by litho_socs_precompute (or any cache: Abbe kernels
 * with weights = source weights).  weight_sum = sum of source weights.
 * batch_capacity = maximum number of equal-size tiles per call. */
int litho_plan_create(litho_ctx* ctx, int n, int k, int win_rows, int win_cols, int win_y0, int win_x0,
                      const double* kernels, const double* weights, double weight_sum,
                      int batch_capacity, litho_plan** out);
int litho_plan_destroy(litho_plan* plan);
/* Device bytes held by the plan (kernels, tables and workspace). */
size_t litho_plan_device_bytes(const litho_plan* plan);
/* Transforms the plan runs (no reference counterpart: the reference has one
 * FFTW path).  direct_dft: 1 for the direct-DFT pipeline of tile sizes without
 * compiled transforms, 0 for the compiled FFTs; corr_len: length L of the
 * band-limited column correlations, 0 for the length-N column transforms;
 * adjoint_mfma: 1 when the band-limited adjoint runs as a Toeplitz GEMM on the
 * matrix cores, 0 for the length-L transforms.  Pointers nullable. */
int litho_plan_column_path(const litho_plan* plan, int* direct_dft, int* corr_len, int* adjoint_mfma);

/* Imaging::compute (imaging.cpp:60-171): aerial = sum_k w_k |E_k|^2 / sum w,
 * wafer = sigmoid(alpha (aerial - threshold)).  aerial / wafer nullable. */
int litho_socs_forward(litho_plan* plan, int batch, const float* mask, float* aerial, float* wafer,
                       double threshold, double alpha);

/* Imaging::compute followed by Gradient::pe_gradient / pe_gradient_with_penalty
 * (gradient.cpp:121-196) and Loss::pe_loss / background_penalty_loss
 * (loss.cpp:16-64):
 *   dLdI = (pe_weight (W-T) pe_window + penalty [T<0.5] max(W-penalty_threshold,0)
 *           penalty_window) W (1-W),  grad_mask = 4 alpha sum_k ... (adjoint).
 * pe_window nullable (= ones); penalty_window nullable (= no penalty term).
 * pe_loss[b] = sum (W-T)^2 pe_window, bg_loss[b] = sum penalty_window [T<.5]
 * max(W-thr,0)^2 — host arrays of length batch (FP64).  aerial, wafer, grad_mask,
 * pe_loss, bg_loss nullable. */
int litho_pe_gradient(litho_plan* plan, int batch, const float* mask, const float* target,
                      const float* pe_window, const float* penalty_window, double pe_weight,
                      double penalty, double penalty_threshold, double threshold, double alpha,
                      float* aerial, float* wafer, float* grad_mask, double* pe_loss, double* bg_loss);

/* One CTM grad-eval, CTM_Optimizer::_evaluate_theta_objective
 * (CTM_Optimizer.cpp:175-187): mask = (1+cos theta)/2, PE and
 * grad_theta = -0.5 grad_mask sin(theta).  pe / grad_norm2 host arrays of length
 * batch (nullable); mask, grad_theta device (nullable). */
int litho_ctm_eval(litho_plan* plan, int batch, const float* theta, const float* target, double threshold,
                   double alpha, float* mask, float* grad_theta, double* pe, double* grad_norm2);

/* CTM_Optimizer::optimize (CTM_Optimizer.cpp:104-227) with the optimizer state on
 * the device.  method: 0 = gradient_descent (fixed step, grad-norm stop, returns
 * the best-PE mask), 1 = lbfgs (LBFGS++-style: two-loop-equivalent compact
 * update, strong-Wolfe line search; returns the final mask). */
typedef struct {
  int method;
  int iterations;
  double learning_rate;
  int lbfgs_history;
  double gradient_tolerance;
  double threshold;
  double alpha;
} litho_ctm_options;

typedef struct {
  int evaluations;       /* objective evaluations incl. line-search trials */
  int iterations;        /* optimizer iterations completed */
  double final_pe;       /* L-BFGS: PE at the returned point; GD: best PE */
  double final_grad_norm;
  int status;            /* 0 converged (norm <= tol), 1 max iterations, 2 line search failed */
} litho_ctm_info;

/* target, theta0 (nullable: initialize_theta(target)), out_mask, out_theta
 * (nullable) are device N x N.  history (host, nullable) receives one PE per
 * evaluation, at most history_capacity. */
int litho_ctm_optimize(litho_plan* plan, const float* target, const float* theta0,
                       const litho_ctm_options* options, float* out_mask, float* out_theta,
                       double* history, int history_capacity, litho_ctm_info* info);

/* Batched CTM over independent tiles of several sizes: the reference's one
 * CTM_Optimizer per tile (demo_CTM_parallel_benchmark.cpp, OpenMP over tiles).
 * Tiles of one size share plans[plan_index]; they are solved `chunk` at a time,
 * each chunk one batched optimizer run whose tiles keep their own trajectory
 * (identical to solving each tile alone).  chunk = 0 chooses the chunk per
 * tile size by timing batched evaluations (cached in the plan).  Plans that
 * share a context run one after another on its stream; plans on distinct
 * contexts run concurrently (one host thread per context).  info (host,
 * num_tiles, nullable); chunk_used (host, num_plans, nullable) receives the
 * chunk each plan ran with. */
typedef struct {
  int plan_index;
  const float* target;   /* device N x N */
  const float* theta0;   /* device N x N, nullable: initialize_theta(target) */
  float* out_mask;       /* device N x N */
  float* out_theta;      /* device N x N, nullable */
} litho_batch_tile;

int litho_batch_ctm(litho_plan* const* plans, int num_plans, const litho_batch_tile* tiles, int num_tiles,
                    const litho_ctm_options* options, int chunk, litho_ctm_info* info, int* chunk_used);

/* Served CTM with continuous batching (no counterpart in the reference, which
 * solves one tile per optimizer object): a worker thread keeps one optimizer
 * pool per plan (tile size) with plan capacity slots, admits queued requests in
 * arrival order whenever a slot is free and the request state in flight stays
 * within memory_budget_bytes, runs batched rounds of the busy pools and
 * releases each request when its own optimizer finishes.  A request's
 * outputs equal those of solving it alone.  Plans that share a context are
 * served by one worker thread on its stream; each further context gets a
 * worker of its own (their rounds overlap on the GPU); the memory budget is
 * shared.  The workers own the plans' contexts until litho_server_destroy (no
 * other calls on them meanwhile).  Request buffers (device N x N) must stay
 * valid until the request is waited for. */
typedef struct litho_server litho_server;

typedef struct {
  litho_ctm_options ctm;
  double memory_budget_bytes;
} litho_server_options;

typedef struct {
  litho_ctm_info info;
  double queued_ms;      /* submit -> admission */
  double service_ms;     /* admission -> completion */
} litho_server_result;

int litho_server_create(litho_plan* const* plans, int num_plans, const litho_server_options* options,
                        litho_server** out);
/* Finishes every admitted and queued request, then stops the worker. */
int litho_server_destroy(litho_server* server);
int litho_server_submit(litho_server* server, int plan_index, const float* target, const float* theta0,
                        float* out_mask, float* out_theta, long long* ticket);
/* Waits up to timeout_s seconds (negative: no limit).  LITHO_OK: done, result
 * filled and the ticket released; LITHO_ETIMEDOUT: not finished yet. */
int litho_server_wait(litho_server* server, long long ticket, double timeout_s, litho_server_result* result);
/* Device bytes one request of this plan counts against the budget. */
size_t litho_server_request_bytes(const litho_server* server, int plan_index);

/* Timing of one CTM grad-eval by HIP events on the context stream (no reference
 * counterpart; the reference times with std::chrono, demo_CTM.cpp:72-98).
 * stage_ms[5] = mean over `repeats` of: forward half (row / column transforms
 * of the mask), column image (power spectra, autocorrelation and image kernel
 * of the band-limited path, or the whole fused column kernel), column adjoint
 * (the correlation adjoint k_corr_adjoint alone; 0 on the fused path),
 * adjoint tail (row / column transforms back to the mask), whole evaluation. */
int litho_profile_ctm_eval(litho_plan* plan, int batch, const float* theta, const float* target, double threshold,
                           double alpha, int repeats, double* stage_ms);

/* Timing of the L-BFGS pair kernel alone (no reference counterpart; the
 * reference's LBFGS++ forms s, y and its dot products on the host): one tile
 * of n floats (a multiple of 4) with a full history of m chained pairs over
 * its point slots (2 m + 4 vectors read, none written).  Each timed launch
 * follows a 1 GiB read that evicts the last-level cache, so the points
 * stream cold, as in a solve (the call allocates that 1 GiB).  *ms = mean
 * over `repeats` launches, by HIP events around each launch on the context
 * stream. */
int litho_profile_lbfgs_pair(litho_ctx* ctx, size_t n, int m, int repeats, double* ms);

/* The yardstick of that kernel (no reference counterpart): a plain kernel
 * that reads `reads` and writes `writes` vectors of n floats (a multiple of
 * 4) with the L-BFGS vector kernels' grid and non-temporal loads, all loads
 * of an element in flight before its sum.  (reads, writes) is one of (24, 0) (the pair kernel's
 * pattern at m = 10), (22, 1) (the direction kernel's), (24, 4), (21, 1);
 * anything else is LITHO_EINVAL.  Each timed launch follows a 1 GiB read
 * that evicts the last-level cache, so the vectors stream cold, as in a
 * solve (the call allocates that 1 GiB).  *ms = mean over `repeats`
 * launches, by HIP events around each launch on the context stream. */
int litho_profile_stream(litho_ctx* ctx, size_t n, int reads, int writes, int repeats, double* ms);

/* The L-BFGS history and inverse-Hessian product of CTM_Optimizer::optimize_lbfgs
 * (CTM_Optimizer.cpp:194-212), i.e. LBFGS++'s BFGSMat::add_correction and
 * apply_Hv, on the device kernels that litho_ctm_optimize runs: for the
 * iterates (x_j, g_j), j < num_points, the pairs s = x_j - x_{j-1},
 * y = g_j - g_{j-1} are added oldest first (kept when s.y > eps y.y; beyond m
 * the oldest leaves), then d = -H g_last with H0 = (s'y / y'y) I of the newest
 * kept pair (d = -g_last when none is kept).  x_points, g_points: device
 * num_points x n floats (n a multiple of 4); d_out: device n floats.
 * gd_out (host, nullable): g_last . d in FP64; kept_out (host, nullable): pairs held. */
int litho_lbfgs_direction(litho_ctx* ctx, size_t n, int m, int num_points, const float* x_points,
                          const float* g_points, float* d_out, double* gd_out, int* kept_out);

/* ---------------------------------------------------------------- EPE
 * Edge points of one tile size (EpsResult: eps rows (y, x) in fractional
 * pixel coordinates, weight_epe), with their bilinear stencils
 * (make_bilinear_stencil, epe_utils.cpp:38-64).  points_yx: host count x 2
 * (y, x); weights: host count values (points with weight <= 0 inactive).
 * The set uses the plan; destroy it before the plan. */
int litho_ep_set_create(litho_plan* plan, int count, const double* points_yx, const double* weights,
                        litho_ep_set** out);
int litho_ep_set_destroy(litho_ep_set* set);

typedef struct {
  double pe_loss;           /* sum (W - T)^2 pe_window; 0 without a target */
  double epe_loss;          /* sum w ((I_ep - threshold) / |grad I|_ep)^2, pixel^2 */
  double epe_abs_sum_nm;    /* sum |EPE| pixel_size_nm over the active points */
  int active_count;
} litho_epe_result;

/* Loss::evaluate_sparse_epe (loss.cpp:125-163) and Gradient::epe_gradient
 * (gradient.cpp:198-227) of one tile at the edge points, on the device:
 * I_ep and |grad I|_ep (central differences, one-sided at the edges, clamped
 * below at 1e-6 per pixel and per sample) sampled bilinearly,
 * dL/dI_ep = 2 w (I_ep - threshold) / |grad I|_ep^2 scattered back, adjoint
 * with factor 2.  With a target the result is added, weighted, to the PE
 * gradient of litho_pe_gradient (pe_window nullable = ones), the PE_EPE
 * combination of micro_gradient_exchange.cpp:208-221:
 *   grad_mask = pe_weight * dPE/dM + epe_weight * dEPE/dM.
 * target null: epe_weight * dEPE/dM alone.  mask, target, pe_window,
 * grad_mask, aerial, wafer: device N x N (outputs nullable).  result: host. */
int litho_epe_gradient(litho_ep_set* set, const float* mask, const float* target, const float* pe_window,
                       double pe_weight, double epe_weight, double threshold, double alpha, double pixel_size_nm,
                       float* grad_mask, float* aerial, float* wafer, litho_epe_result* result);

/* CTM_Optimizer::initialize_theta (CTM_Optimizer.cpp:66-73): 0.3 pi where
 * pattern > 0, else 0.7 pi.  Device arrays of n*n floats. */
int litho_initialize_theta(litho_ctx* ctx, const float* pattern, float* theta, size_t count);

/* ---------------------------------------------------------------- level set
 * Fields are DEVICE float32 arrays of `batch` row-major rows x cols tiles
 * (rows, cols >= 3; cols <= 4096); stencil arithmetic is FP64. */

/* LevelSetUtils::evolve_normal_WENO_godunov (level_set_utils.cpp:150-225) and
 * evolve_kappa (level_set_utils.cpp:257-278): delta = Vn |grad phi| with WENO5-JS
 * one-sided derivatives and Godunov upwinding (edge padding 3), kappa =
 * curvature_weight (phi_xx phi_y^2 - 2 phi_x phi_y phi_xy + phi_yy phi_x^2) /
 * |grad phi|^2 (0 where |grad phi|^2 <= DBL_EPSILON).  H1 = H2 = |Vn| are not
 * returned.  delta or kappa may be null (vn may then be null too). */
int litho_ls_terms(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, const float* vn, double dx,
                   double dy, double curvature_weight, float* delta, float* kappa);

/* LevelSetUtils::get_dt_normal_kappa (level_set_utils.cpp:280-288), host:
 * cfl / (max_h1/dx + max_h2/dy + 2 b (1/dx^2 + 1/dy^2)), or 1e-3 when that
 * denominator is below 1e-9. */
double litho_ls_dt(double cfl, double dx, double dy, double max_h1, double max_h2, double curvature_weight);

/* LevelSetUtils::reinit_sdf (level_set_utils.cpp:290-319): inside = phi > 0,
 * sdf = (dist_in - dist_out) dx with OpenCV distanceTransform(DIST_L2, 5)
 * semantics (5x5 chamfer, 16-bit fixed point).  sdf may alias phi. */
int litho_ls_reinit_sdf(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, double dx, double dy,
                        float* sdf);

/* litho_ls_reinit_sdf without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its scratch (chamfer row buffers and distance planes, 16 bytes per pixel) is
 * a static 64 MiB device array per GPU, held until the process ends: calls
 * with batch * rows * cols above 4 Mi pixels fail with LITHO_EINVAL.  Calls
 * that overlap on one GPU (any contexts) share that array and race.  Results
 * equal litho_ls_reinit_sdf's bit for bit. */
int litho_ls_reinit_sdf_noalloc(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, double dx,
                                double dy, float* sdf);

typedef struct {
  double dx, dy;             /* FullChipLSMOptions dx / dy (full_chip_lsm.h:8-66) */
  double cfl;                /* CFL number */
  double curvature_weight;   /* b */
  int mask_type;             /* render of the masks: 0 Binary, 1 Gray (full_chip_lsm.h:14) */
  int render_scale;          /* Gray: upscale per direction (full_chip_lsm.h:15, default 4) */
} litho_lsm_options;

/* render_levelset_mask(phi, type, scale) (levelset_mask_renderer.cpp:7-33) of
 * `batch` rows x cols tiles: type 0 Binary (phi >= 0); type 1 Gray: float32
 * phi upscaled by `scale` with cv::resize INTER_LINEAR semantics (pixel-centre
 * bilinear, replicated borders), subpixels >= 0 counted per pixel and divided
 * by scale^2.  phi, mask_out: device batch x rows x cols. */
int litho_render_levelset_mask(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, int type, int scale,
                               float* mask_out);

/* One RK2 level-set update of `batch` n x n Micros, the loop body of
 * FullChipLSMOptimizer::optimize (full_chip_lsm.cpp:208-242):
 *   Vn = mask_gradient inside the tile's update window (Core + Guard), 0 outside;
 *   N(phi) = kappa(phi) - Vn |grad phi|;  dt from stage 1;
 *   phi1 = phi + dt N(phi) (window);  phi' = phi/2 + (phi1 + dt N(phi1))/2 (window);
 *   with reinit != 0 the window is re-initialised to reinit_sdf((phi' >= 0) - 0.5).
 * update_rects: HOST array of batch x {y0, y1, x0, x1} (half-open, tile-local).
 * phi_out may alias phi.  mask_out (nullable): render of options->mask_type.
 * dt_out (HOST, nullable): the step of each tile; passing it waits for the step. */
int litho_lsm_step(litho_ctx* ctx, int batch, int n, const float* phi, const float* mask_gradient,
                   const int* update_rects, const litho_lsm_options* options, int reinit, float* phi_out,
                   float* mask_out, double* dt_out);

/* ---------------------------------------------------------------- full chip
 * Micro tiling of a height x width chip into grid_rows x grid_cols Micros of
 * tile_size x tile_size pixels (Core + Halo, periodic halo), Partition::build
 * (partition.cpp:21-80).  Rectangles are half-open [y0, y1) x [x0, x1). */
typedef struct {
  int y0, y1, x0, x1;
} litho_rect;

typedef struct {
  int id;                               /* neighbour Micro id */
  int direction;                        /* 0 left, 1 right, 2 top, 3 bottom */
  litho_rect guard_global;              /* my Guard strip inside the neighbour's Core, global */
  litho_rect my_guard_local;            /* the same strip in my local coordinates */
  litho_rect neighbor_overlap_local;    /* the same strip in the neighbour's local coordinates */
  litho_rect overlap_global;            /* my Core strip inside the neighbour's Guard, global */
  litho_rect my_overlap_local;          /* that strip in my local coordinates */
  litho_rect neighbor_guard_local;      /* that strip in the neighbour's local coordinates */
} litho_neighbor;                       /* MicroNeighbor (partition.h:26-45) */

typedef struct {
  int id, row, col, halo, guard;
  litho_rect core_global, tile_global, core_local, update_local;   /* update = Core + Guard */
  int num_neighbors;
  litho_neighbor neighbors[4];          /* in left, right, top, bottom order, existing ones only */
} litho_tile;                           /* MicroTile (partition.h:47-65) */

/* Partition::build.  tiles: caller array of grid_rows * grid_cols entries, id =
 * row * grid_cols + col.  tile_size = max core side + 2 halo, rounded up to even. */
int litho_partition_build(int height, int width, int grid_rows, int grid_cols, int halo, int guard,
                          int* tile_size, litho_tile* tiles);

/* Partition::extract_local_tile for every Micro (partition.cpp:83-95): periodic
 * gather of the device height x width array into num_tiles device tiles. */
int litho_partition_extract(litho_ctx* ctx, const float* global, int height, int width, const litho_tile* tiles,
                            int num_tiles, int tile_size, float* local_out);

/* Partition::assemble_global (partition.cpp:105-113): Core scatter of device
 * tiles into the device height x width array (pixels outside every Core untouched). */
int litho_partition_assemble(litho_ctx* ctx, const float* local, const litho_tile* tiles, int num_tiles,
                             int tile_size, int height, int width, float* global_out);

/* MicroGradientOptions (micro_gradient_exchange.h:19-40), PE and PEPenalty data
 * losses (the EPE modes are not part of this library). */
typedef struct {
  int loss_type;                 /* 0 PE (window = loss region), 1 PEPenalty (PE on target > .5, penalty on < .5) */
  int loss_region;               /* 0 Core, 1 CoreGuard */
  double pe_weight;
  double penalty_weight;
  double penalty_threshold;
  int collect_core_metrics;
  int enable_cept;
  int cept_guard_only;
  double cept_weight;
  int enable_grad_fusion;
  int fusion_mode;               /* 0 Average, 1 Linear */
  double fusion_weight_owner;
  double fusion_weight_neighbor;
  double threshold;              /* resist threshold */
  double alpha;                  /* resist steepness */
} litho_exchange_options;

typedef struct {                 /* MicroGradientData scalars (micro_gradient_exchange.h:98-112) */
  double loss, data_loss, pe_loss, penalty_loss, bg_penalty_raw, cept_loss;
  /* collect_core_metrics only, CoreLossMetrics (micro_gradient_exchange.h:64-73): */
  double core_pe_loss, core_bg_penalty_raw;
  double core_pe_penalty_loss;   /* pe_weight * PE on the Core's main pattern + penalty_weight * core_bg_penalty_raw */
  double core_data_loss;         /* the selected loss on the Core alone, without CEPT */
} litho_tile_loss;

/* MicroGradientExchange::compute(local_masks, ...) (micro_gradient_exchange.cpp:134-334)
 * on device-resident Micros: batched imaging + windowed PE / penalty adjoint of
 * every Micro through the plan (plan tile size = tile_size; Micros are processed
 * in chunks of the plan's batch capacity), then the four-neighbour exchange:
 * Guard fusion (Average / Linear) of the raw gradients and CEPT, with
 * mask_gradient = fused gradient + cept_weight * CEPT gradient.
 * local_masks, local_targets: device [num_tiles][tile_size^2].  mask_gradient:
 * device output; raw_gradient, wafer (device) and per_tile (host, num_tiles) are
 * nullable.  totals (host, nullable): {loss, data_loss, cept_loss}. */
int litho_tile_exchange(litho_plan* plan, const litho_tile* tiles, int num_tiles, const float* local_masks,
                        const float* local_targets, const litho_exchange_options* options, float* mask_gradient,
                        float* raw_gradient, float* wafer, litho_tile_loss* per_tile, double* totals);

/* MicroImaging::compute(global_mask, threshold, alpha) (mirco_imaging.cpp:9-27):
 * each Micro's tile is gathered from the global mask with periodic indices,
 * imaged (Imaging::compute, batches of the plan capacity), and every Core is
 * written into the global aerial / wafer images.  global_mask, global_*:
 * device height x width; local_* (nullable): device [num_tiles][n][n]. */
int litho_micro_imaging(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                        const float* global_mask, double threshold, double alpha, float* global_aerial,
                        float* global_wafer, float* local_aerial, float* local_wafer);

/* compare_micro_imaging(params, meta, cache, target, optimized_local_masks)
 * (full_chip_imaging_compare.cpp:19-71), the post-run evaluation: the target's
 * Micros (periodic extracts of global_target) and the optimized Micros as
 * given (device [num_tiles][n][n], their own Guard and Halo kept) are imaged
 * through the plan, each set's Cores assembled into global height x width
 * aerial / wafer images (zero outside every Core, as assemble_global), and
 * target_pe / optimized_pe = Loss::pe_loss(wafer, global_target) in FP64
 * (host, nullable).  optimized_mask: the Core-assembled optimized Micros.
 * Every output is nullable; optimized_local_aerial is [num_tiles][n][n]. */
int litho_compare_micro_imaging(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                                const float* global_target, const float* optimized_local_masks, double threshold,
                                double alpha, float* optimized_mask, float* target_aerial, float* target_wafer,
                                float* optimized_aerial, float* optimized_wafer, float* optimized_local_aerial,
                                double* target_pe, double* optimized_pe);

/* MaskContinuityMetrics (mask_continuity.h:9-19). */
typedef struct {
  long long overlap_samples;       /* directed Guard -> neighbour Core pixel pairs */
  double overlap_sse, overlap_rmse, overlap_mae, overlap_max_abs;
  long long seam_samples, flat_seam_samples, foreground_seam_samples;
  double seam_jump_mae, seam_reference_error_mae, flat_seam_jump_mae, foreground_seam_jump_mae;
} litho_mask_continuity;

/* evaluate_mask_continuity(meta, local_masks, target) (mask_continuity.cpp:8-61):
 * for every Micro and neighbour, its Guard strip of the local masks against the
 * same pixels in the neighbour's local mask; for every right / bottom
 * neighbour, the pixel pairs across the Core seam of the assembled mask, with
 * the target's own change across the seam separated out.  local_masks: device
 * [num_tiles][tile_size^2]; target: device height x width; out: host.  FP64
 * sums in a fixed order. */
int litho_evaluate_mask_continuity(litho_ctx* ctx, const litho_tile* tiles, int num_tiles, int tile_size, int height,
                                   int width, const float* local_masks, const float* target,
                                   litho_mask_continuity* out);

/* One state of a full-chip run: FullChipCTMIteration (full_chip_ctm.h:19-29)
 * and FullChipLSMIteration (full_chip_lsm.h), the columns of the demos'
 * history.csv.  Contributions are weighted (0 for a term that is off); raw
 * values are summed over the Micros' loss regions; core_* sum the Micros'
 * CoreLossMetrics (collected by the LSM, else 0).  The EPE columns stay 0:
 * the exchange runs the PE / PEPenalty losses. */
typedef struct {
  int iteration;
  double loss, data_loss, cept_loss;
  double min_dt, max_dt;                                  /* LSM: the update's step range; CTM: 0 */
  double pe_contribution, penalty_contribution, epe_contribution, cept_contribution;
  double pe_loss_raw, bg_penalty_raw, epe_loss_raw, epe_abs_sum_nm;
  int epe_point_count;
  double core_pe, core_bg_penalty_raw, core_pe_penalty, core_epe_loss, core_epe_abs_sum_nm;
  int core_ep_count;
  double core_epe_mean, core_data_loss, epe_mean;
} litho_fullchip_row;

/* FullChipCTMOptimizer::optimize(initial_pattern, global_target, options)
 * (full_chip_ctm.cpp:25-152): theta per Micro from initialize_theta of the
 * pattern, fixed-step gradient descent on theta inside Core + Guard (Halo frozen),
 * CEPT forced guard-only; the best total loss over iterations 0..n wins.
 * initial_pattern, global_target: device height x width.  global_mask_out:
 * device height x width (best Micros' Cores).  local_masks_out (device,
 * nullable): best Micros, [num_tiles][tile_size^2].  history (host, nullable):
 * iterations + 1 rows. */
int litho_fullchip_ctm(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                       const float* initial_pattern, const float* global_target, const litho_exchange_options* options,
                       int iterations, double learning_rate, float* global_mask_out, float* local_masks_out,
                       litho_fullchip_row* history, int* best_iteration, double* best_loss);

/* FullChipLSMOptimizer::optimize(initial_binary, global_target, options)
 * (full_chip_lsm.cpp:39-255) with the renderer of lsm->mask_type: phi per Micro = reinit_sdf(mask - 0.5),
 * loss region forced to Core + Guard, CEPT guard-only, Core metrics collected;
 * each iteration one RK2 update (litho_lsm_step) of every Micro with re-init
 * every reinit_interval updates (0: never), then the exchange.
 * global_mask_out / global_phi_out: device height x width (nullable).
 * history (host, nullable): iterations + 1 rows. */
int litho_fullchip_lsm(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                       const float* initial_binary, const float* global_target, const litho_exchange_options* options,
                       const litho_lsm_options* lsm, int iterations, int reinit_interval, float* global_mask_out,
                       float* global_phi_out, litho_fullchip_row* history, int* best_iteration, double* best_loss);

/* ---------------------------------------------------------------- CTM -> binary
 * Imbinarize::binarize_advanced(gray_mask, target_pattern, peak_min_intensity,
 * quantile_val, area_threshold, min_spacing_pixels) (imbinarize.cpp:84-190),
 * the hand-off from the CTM gray mask to the level-set stage
 * (demo_full_chip_ctm.cpp:109 calls it with 0.125, 0.1, 5, 5):
 *   threshold = quantile_val-quantile (linear) of the gray values at local
 *               maxima (>= their in-image 3 x 3 neighbourhood) above
 *               peak_min_intensity;
 *   initial   = gray >= threshold;  filtered = initial without 8-connected
 *               components of area < area_threshold;
 *   sraf_initial = filtered and not (target > 0.5);  keep_out = (target > 0.5)
 *               dilated by the elliptical element of radius min_spacing_pixels
 *               (0: none); sraf_final = sraf_initial and not keep_out;
 *   final     = (target > 0.5) or sraf_final, without components of area
 *               < area_threshold.
 * With no peak every mask is empty (final included), as the reference.
 * height x width device float32 arrays; outputs 0/1; all outputs except
 * final_mask nullable.  info (host, nullable).  Synchronises the context. */
typedef struct {
  double threshold;        /* the adaptive threshold (0 when no peak) */
  long long num_peaks;
} litho_binarize_info;

int litho_binarize_advanced(litho_ctx* ctx, int height, int width, const float* gray_mask, const float* target_pattern,
                            double peak_min_intensity, double quantile_val, int area_threshold,
                            int min_spacing_pixels, float* final_mask, float* initial_binary_mask,
                            float* filtered_binary_mask, float* sraf_mask_initial, float* sraf_mask_final,
                            float* keep_out_zone, litho_binarize_info* info);

/* ---------------------------------------------------------------- on-disk formats
 * Host-only text output of the reference's callers.  Matrices are host
 * row-major double arrays (row = y); paths are used as given.
 *
 * SaveTxt::save_mat (save_txt.cpp:23-37): "# rows=R cols=C", then one line per
 * row of fixed-point values at `precision` decimals separated by one space. */
int litho_save_txt(const char* path, const double* matrix, int rows, int cols, int precision);
/* SaveTxt::save_mat_int (save_txt.cpp:39-51): the same with values rounded to integers. */
int litho_save_txt_int(const char* path, const double* matrix, int rows, int cols);
/* SaveTxt::load_txt (save_txt.cpp:58-105): skips blank and '#' lines, reads
 * whitespace-separated values, requires a constant column count (else
 * LITHO_ERUNTIME, as the reference's runtime_error).  matrix null: only
 * *rows / *cols; else capacity >= rows * cols values are written. */
int litho_load_txt(const char* path, double* matrix, size_t capacity, int* rows, int* cols);
/* The matrix TXT of the full-chip demos and save_full_chip_imaging_comparison
 * (full_chip_imaging_compare.cpp:76-80): "# rows=R cols=C", then the Eigen
 * stream of the matrix at precision 17 (every value right-aligned to the
 * widest value's length, one space between values). */
int litho_save_matrix_txt(const char* path, const double* matrix, int rows, int cols);
/* history.csv of the full-chip CTM (stage 0; demo_full_chip_ctm.cpp:112-122)
 * or LSM demo (stage 1; demo_full_chip_lsm.cpp:170-184), precision 17. */
int litho_write_history_csv(const char* path, int stage, const litho_fullchip_row* rows, int count);
/* summary.csv (demo_full_chip_ctm.cpp:124-128; LSM adds `seconds`, demo_full_chip_lsm.cpp:196-199). */
int litho_write_summary_csv(const char* path, int stage, int best_iteration, double best_loss, int final_iteration,
                            double final_loss, double seconds);
/* metrics.csv and sampling.csv of save_full_chip_imaging_comparison
 * (full_chip_imaging_compare.cpp:94-101); micro_core != 0 writes "micro_core". */
int litho_write_imaging_metrics_csv(const char* path, int stage, double target_pe, double optimized_pe);
int litho_write_sampling_csv(const char* path, int grid_size, int source_grid_size, int source_points, int kernels,
                             double pixel_size_nm, int micro_core, int image_rows, int image_cols, int micro_count);

#ifdef __cplusplus
}
#endif

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#endif /* LITHO_AMD_H */
