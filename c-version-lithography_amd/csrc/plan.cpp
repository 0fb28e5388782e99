// Plan construction and the batched evaluation sequence.
#include "plan.hpp"

#include <atomic>
#include <cstdint>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <string>

namespace litho {

Plan* make_plan(Context* ctx, int n, int k, int wr, int wc, int y0, int x0, const double* kernels,
                const double* weights, double weight_sum, int capacity) {
  // validation mirrors Imaging::compute (imaging.cpp:68-107)
  if (n <= 0) throw Error(1, "Imaging: cache.N must be positive");
  if (k <= 0) throw Error(1, "Imaging: no optical kernels available");
  if (kernels == nullptr || weights == nullptr) throw Error(1, "plan: kernels and weights are required");
  if (wr <= 0 || wc <= 0 || wr > n || wc > n || y0 < 0 || x0 < 0 || y0 + wr > n || x0 + wc > n)
    throw Error(1, "plan: kernel window does not fit the N x N grid");
  if (!std::isfinite(weight_sum) || weight_sum <= 0.0)
    throw Error(1, "Imaging: source weights must be finite, nonnegative, and have a positive sum");
  for (int i = 0; i < k; ++i)
    if (!std::isfinite(weights[i]) || weights[i] < 0.0)
      throw Error(1, "Imaging: SOCS values must be finite and nonnegative");
  if (capacity <= 0) throw Error(1, "plan: batch capacity must be positive");
  // sizes without compiled transforms (or LITHO_SOCS_PATH = generic) run the
  // direct-DFT pipeline, which needs the band-limited column path
  const char* gpath = std::getenv("LITHO_SOCS_PATH");
  const SocsOps* ops = gpath != nullptr && std::string(gpath) == "generic" ? nullptr : socs_ops_for(n);
  const bool generic = ops == nullptr;
  if (generic) {
    int l = 64;
    while (l < 2 * wr - 1) l *= 2;
    if (n > kGenericMaxN || !corr_len_supported(l) || 2 * wr - 1 > n)
      throw Error(4, "plan: tile size " + std::to_string(n) +
                         " has no compiled FFT, and the direct-DFT path needs N <= " + std::to_string(kGenericMaxN) +
                         " and 2 W_r - 1 <= min(N, 1023)");
    ops = generic_socs_ops();
  }

  Plan* p = new Plan();
  try {
    p->ctx = ctx;
    p->n = n; p->k = k; p->wr = wr; p->wc = wc; p->win_y0 = y0; p->win_x0 = x0;
    p->capacity = capacity; p->weight_sum = weight_sum; p->ops = ops;
    p->col_blocks = ops->col_blocks();
    LITHO_HIP(hipSetDevice(ctx->device));

    // scaled kernels H~_k = H_k sqrt(lambda_k / sum w) / N^2
    std::vector<float2> hk((size_t)k * wr * wc);
    const double inv_n2 = 1.0 / ((double)n * (double)n);
    for (int kk = 0; kk < k; ++kk) {
      const double c = std::sqrt(weights[kk] / weight_sum) * inv_n2;
      for (size_t e = 0; e < (size_t)wr * wc; ++e) {
        const size_t idx = (size_t)kk * wr * wc + e;
        const double re = kernels[2 * idx], im = kernels[2 * idx + 1];
        if (!std::isfinite(re) || !std::isfinite(im)) throw Error(1, "plan: kernels must be finite");
        hk[idx] = make_float2((float)(re * c), (float)(im * c));
      }
    }
    std::vector<float2> tw(n);
    for (int m = 0; m < n; ++m) {
      const double ang = -2.0 * 3.14159265358979323846 * (double)m / (double)n;
      tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    // DC-corner frequency index of centred index c: (c - N/2) mod N
    std::vector<int> rf(wr), cf(wc);
    for (int i = 0; i < wr; ++i) rf[i] = (((y0 + i) - n / 2) % n + n) % n;
    for (int j = 0; j < wc; ++j) cf[j] = (((x0 + j) - n / 2) % n + n) % n;

    // frequency -> window row table and the pruning precondition: every window
    // frequency lies in the first or the last 1/R_1 of the axis
    std::vector<int> rou(n, -1);
    for (int i = 0; i < wr; ++i) rou[rf[i]] = i;
    const int r1 = first_radix(n);
    bool prunable = r1 > 0;
    if (prunable) {
      const int t1 = n / r1;
      for (int v : rf) prunable = prunable && (v < t1 || v >= n - t1);
    }
    const char* path = std::getenv("LITHO_COLUMN_PATH");
    p->pruned = prunable && !(path != nullptr && std::string(path) == "generic") ? 1 : 0;
    p->row_of_u.alloc(rou.size() * sizeof(int));
    LITHO_HIP(hipMemcpy(p->row_of_u.ptr, rou.data(), p->row_of_u.bytes, hipMemcpyHostToDevice));

    // the same for the window columns and the row transforms; the real-row
    // forward (two rows per complex line) also reads frequency N - v, so the
    // upper bound is strict
    std::vector<int> cou(n, -1);
    for (int j = 0; j < wc; ++j) cou[cf[j]] = j;
    bool row_prunable = r1 > 0;
    if (row_prunable) {
      const int t1 = n / r1;
      for (int v : cf) row_prunable = row_prunable && (v < t1 || v > n - t1);
    }
    const char* rpath = std::getenv("LITHO_ROW_PATH");
    p->row_pruned = row_prunable && !(rpath != nullptr && std::string(rpath) == "generic") ? 1 : 0;
    p->col_of_u.alloc(cou.size() * sizeof(int));
    LITHO_HIP(hipMemcpy(p->col_of_u.ptr, cou.data(), p->col_of_u.bytes, hipMemcpyHostToDevice));
    const std::vector<float2> rtw = row_stage_twiddles(n);
    p->rtw.alloc((rtw.empty() ? 1 : rtw.size()) * sizeof(float2));
    if (!rtw.empty())
      LITHO_HIP(hipMemcpy(p->rtw.ptr, rtw.data(), rtw.size() * sizeof(float2), hipMemcpyHostToDevice));

    p->h.alloc(hk.size() * sizeof(float2));
    p->tw.alloc(tw.size() * sizeof(float2));
    p->row_freq.alloc(rf.size() * sizeof(int));
    p->col_freq.alloc(cf.size() * sizeof(int));
    LITHO_HIP(hipMemcpy(p->h.ptr, hk.data(), p->h.bytes, hipMemcpyHostToDevice));
    LITHO_HIP(hipMemcpy(p->tw.ptr, tw.data(), p->tw.bytes, hipMemcpyHostToDevice));
    LITHO_HIP(hipMemcpy(p->row_freq.ptr, rf.data(), p->row_freq.bytes, hipMemcpyHostToDevice));
    LITHO_HIP(hipMemcpy(p->col_freq.ptr, cf.data(), p->col_freq.bytes, hipMemcpyHostToDevice));

    // band-limited column path: L = the smallest supported power of two >= 2 W_r - 1
    // (LITHO_COLUMN_PATH = fused / generic selects the transforms of length N instead)
    int L = 64;
    while (L < 2 * wr - 1) L *= 2;
    const bool corr_ok = corr_len_supported(L) && 2 * wr - 1 <= n;
    const bool corr_off = !generic && path != nullptr && (std::string(path) == "fused" || std::string(path) == "generic");
    p->corr_len = corr_ok && !corr_off ? L : 0;
    p->generic = generic ? 1 : 0;
    // the band-limited adjoint: the length-L transforms, or with
    // LITHO_COLUMN_ADJOINT = mfma a Toeplitz GEMM on the matrix cores.  The
    // GEMM (bf16 pairs, three products per term) is faster only up to
    // W_r = 80 (640^2 tiles, W_r = 67: 216 against 238 us per batch of 32;
    // 896^2: 396 against 332 us; 2048^2 alone: 82 against 62 us), and its
    // 16-bit operands move the fixed-step GD trajectory of the 512^2 CPU
    // configuration by 2e-4 at iteration 19, against 1e-7 for the FP32
    // transforms, so it is not the default.
    const char* adj = std::getenv("LITHO_COLUMN_ADJOINT");
    const bool want_mfma = adj != nullptr && std::string(adj) == "mfma";
    p->adj_mfma = p->corr_len > 0 && corr_adjoint_mfma_supported(wr, k) && want_mfma ? 1 : 0;
    if (p->corr_len > 0) {
      // pass twiddle tables of the register-resident length-L transforms (the inverse applies conjugates)
      const std::vector<float2> twl = corr_twiddles(L);
      p->twl.alloc(twl.size() * sizeof(float2));
      LITHO_HIP(hipMemcpy(p->twl.ptr, twl.data(), p->twl.bytes, hipMemcpyHostToDevice));
      // the same tables at length n for the register-resident image kernel (powers of two)
      const std::vector<float2> twi = p->generic ? std::vector<float2>{} : corr_twiddles(n);
      if (!twi.empty()) {
        p->twi.alloc(twi.size() * sizeof(float2));
        LITHO_HIP(hipMemcpy(p->twi.ptr, twi.data(), p->twi.bytes, hipMemcpyHostToDevice));
      }
    }

    const size_t c = (size_t)capacity, N = (size_t)n;
    p->rt.alloc(c * wc * N * sizeof(float2));
    p->mhat.alloc(c * wr * wc * sizeof(float2));
    p->t.alloc(c * k * wr * N * sizeof(float2));
    p->v.alloc(c * k * wr * N * sizeof(float2));
    p->y.alloc(c * k * wr * wc * sizeof(float2));
    p->p.alloc(c * wr * N * sizeof(float2));
    if (p->corr_len > 0) {
      p->acorr.alloc(c * wr * N * sizeof(float2));
      p->ghat.alloc(c * wr * N * sizeof(float2));
      p->spart.alloc(corr_power_bytes(n, p->corr_len, k, capacity));
    }
    p->img_blocks = p->corr_len > 0 ? ops->img_blocks() : p->col_blocks;
    p->img_partials.alloc(c * p->img_blocks * 2 * sizeof(double));
    p->grad_partials.alloc(c * p->col_blocks * 2 * sizeof(double));
    p->red.alloc(c * 4 * sizeof(double));
    p->coef.alloc(c * sizeof(float));
    p->host_red = new HostPinned(c * 4 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent);
    p->host_coef = new HostPinned(c * sizeof(float));
    p->device_bytes = p->h.bytes + p->tw.bytes + p->rtw.bytes + p->row_of_u.bytes + p->col_of_u.bytes +
                      p->row_freq.bytes + p->col_freq.bytes + p->rt.bytes +
                      p->mhat.bytes + p->t.bytes + p->v.bytes + p->y.bytes + p->p.bytes + p->twl.bytes + p->twi.bytes +
                      p->acorr.bytes + p->ghat.bytes + p->spart.bytes + p->img_partials.bytes +
                      p->grad_partials.bytes + p->red.bytes + p->coef.bytes;
  } catch (...) {
    delete p;
    throw;
  }
  return p;
}

void run_eval(Plan& plan, int batch, const EvalRequest& rq) {
  if (batch <= 0 || batch > plan.capacity) throw Error(1, "batch exceeds the plan capacity");
  hipStream_t s = plan.ctx->stream;
  SocsArgs a = plan.args();
  a.active = rq.active;
  a.ring = rq.ring;
  a.ring_stride = rq.ring_stride;
  a.ring_points = rq.ring_points;
  plan.ops->forward_half(a, batch, rq.mode, rq.base, rq.dir, rq.coef, rq.theta_out, rq.mask_out, s);
  ImageArgs im = rq.im;
  im.partials = plan.img_partials.as<double>();
  SocsArgs a2 = a;
  if (!rq.adjoint) a2.v = nullptr;
  plan.ops->image_adjoint(a2, im, batch, s, kColumnAll);
  double* red = rq.red_out != nullptr ? rq.red_out : plan.red.as<double>();
  // only the evaluated tiles' reductions are written: the host may wait on
  // them by value (wait_reductions) while other tiles' entries are rearmed.
  // With the adjoint, the image partials are reduced together with the
  // gradient's at the end (one launch, off the adjoint's path).
  if (rq.adjoint && rq.im.target != nullptr) {
    GradArgs ga = rq.ga;
    ga.partials = plan.grad_partials.as<double>();
    ga.nt_grad_theta = batch == 1 ? 1 : 0;
    plan.ops->adjoint_tail(a, ga, batch, s);
    launch_reduce_eval(plan.img_partials.as<double>(), plan.img_blocks, plan.grad_partials.as<double>(),
                       plan.col_blocks, red, batch, s, rq.active, rq.extra_partials, rq.extra_blocks, rq.extra_out);
  } else {
    reduce_pairs(plan.img_partials.as<double>(), plan.img_blocks, red, 4, batch, s, rq.active);
  }
  LITHO_HIP(hipGetLastError());
}

EvalGraph::EvalGraph(Plan& plan, int batch, const EvalRequest& rq) : plan_(plan), batch_(batch), rq_(rq) {
  if (batch <= 0 || batch > plan.capacity) throw Error(1, "batch exceeds the plan capacity");
  if (!plan.ctx->graphs) {
    plan.ctx->graphs_direct.fetch_add(1);
    return;
  }
  capture();
  if (exec_ != nullptr)
    plan.ctx->graphs_captured.fetch_add(1);
  else
    plan.ctx->graphs_failed.fetch_add(1);
}

void EvalGraph::capture() {
  Plan& plan = plan_;
  hipStream_t s = plan.ctx->stream;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  bool ok = true;
  try {
    run_eval(plan, batch_, rq_);
  } catch (const Error&) {
    ok = false;
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (!ok || e != hipSuccess || g == nullptr) {
    if (g != nullptr) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (!ok) throw Error(3, "kernel launch failed while capturing the evaluation graph");
    return;
  }
  graph_ = g;
  if (hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    exec_ = nullptr;
  }
}

EvalGraph::~EvalGraph() {
  if (exec_ != nullptr) (void)hipGraphExecDestroy(exec_);
  if (graph_ != nullptr) (void)hipGraphDestroy(graph_);
}

void EvalGraph::launch() {
  if (exec_ != nullptr) {
    LITHO_HIP(hipGraphLaunch(exec_, plan_.ctx->stream));
  } else {
    run_eval(plan_, batch_, rq_);
  }
}

void profile_eval(Plan& plan, int batch, const EvalRequest& rq, int repeats, double* stage_ms) {
  if (repeats <= 0) throw Error(1, "profile: repeats must be positive");
  hipStream_t s = plan.ctx->stream;
  hipEvent_t ev[5];
  for (auto& e : ev) LITHO_HIP(hipEventCreate(&e));
  double acc[5] = {0, 0, 0, 0, 0};
  SocsArgs a = plan.args();
  a.active = rq.active;
  a.ring = rq.ring;
  a.ring_stride = rq.ring_stride;
  a.ring_points = rq.ring_points;
  ImageArgs im = rq.im;
  im.partials = plan.img_partials.as<double>();
  GradArgs ga = rq.ga;
  ga.partials = plan.grad_partials.as<double>();
  ga.nt_grad_theta = batch == 1 ? 1 : 0;   // as run_eval sets it
  // untimed warm-up: the first launch of each kernel loads its code object
  plan.ops->forward_half(a, batch, rq.mode, rq.base, rq.dir, rq.coef, rq.theta_out, rq.mask_out, s);
  plan.ops->image_adjoint(a, im, batch, s, kColumnAll);
  plan.ops->adjoint_tail(a, ga, batch, s);
  LITHO_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < repeats; ++r) {
    LITHO_HIP(hipEventRecord(ev[0], s));
    plan.ops->forward_half(a, batch, rq.mode, rq.base, rq.dir, rq.coef, rq.theta_out, rq.mask_out, s);
    LITHO_HIP(hipEventRecord(ev[1], s));
    plan.ops->image_adjoint(a, im, batch, s, kColumnImage);
    LITHO_HIP(hipEventRecord(ev[2], s));
    plan.ops->image_adjoint(a, im, batch, s, kColumnAdjoint);
    LITHO_HIP(hipEventRecord(ev[3], s));
    plan.ops->adjoint_tail(a, ga, batch, s);
    LITHO_HIP(hipEventRecord(ev[4], s));
    LITHO_HIP(hipEventSynchronize(ev[4]));
    float ms = 0.f;
    for (int i = 0; i < 4; ++i) {
      LITHO_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      acc[i] += ms;
    }
    LITHO_HIP(hipEventElapsedTime(&ms, ev[0], ev[4]));
    acc[4] += ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int i = 0; i < 5; ++i) stage_ms[i] = acc[i] / repeats;
}

// Host-side wait for the stream by polling: the optimizer's host turn between
// an evaluation's reductions and the next round's upload is on the critical
// path of every round, and polling answers sooner than a blocking wait.
void spin_wait(hipStream_t s) {
  hipError_t e;
  while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
  }
  LITHO_HIP(e);
}

namespace {
// A signalling-NaN pattern that no reduction produces (FP64 arithmetic yields
// quiet NaNs), marking a reduction as not yet written.
constexpr uint64_t kUnwritten = 0x7FF4C0FFEE0DD00DULL;
}  // namespace

void arm_values(double* host_values, const int* flags, int count) {
  uint64_t* v = reinterpret_cast<uint64_t*>(host_values);
  for (int i = 0; i < count; ++i)
    if (flags == nullptr || flags[i] != 0) __atomic_store_n(&v[i], kUnwritten, __ATOMIC_RELAXED);
  std::atomic_thread_fence(std::memory_order_seq_cst);
}

void wait_values(const double* host_values, const int* flags, int count, hipStream_t s) {
  const uint64_t* v = reinterpret_cast<const uint64_t*>(host_values);
  for (int i = 0; i < count; ++i) {
    if (flags != nullptr && flags[i] == 0) continue;
    for (unsigned spins = 1; __atomic_load_n(&v[i], __ATOMIC_ACQUIRE) == kUnwritten; ++spins) {
      if ((spins & 4095u) != 0) continue;
      const hipError_t e = hipStreamQuery(s);
      if (e == hipErrorNotReady) continue;
      LITHO_HIP(e);
      if (__atomic_load_n(&v[i], __ATOMIC_ACQUIRE) == kUnwritten)
        throw Error(3, "evaluation finished without writing its reductions");
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
}

void arm_reductions(Plan& plan, const int* flags, int batch) {
  uint64_t* red = static_cast<uint64_t*>(plan.host_red->ptr);
  for (int b = 0; b < batch; ++b)
    if (flags[b] != 0)
      for (int k = 0; k < 4; ++k) __atomic_store_n(&red[4 * b + k], kUnwritten, __ATOMIC_RELAXED);
  std::atomic_thread_fence(std::memory_order_seq_cst);
}

const double* wait_reductions(Plan& plan, const int* flags, int batch) {
  const uint64_t* red = static_cast<const uint64_t*>(plan.host_red->ptr);
  hipStream_t s = plan.ctx->stream;
  for (int b = 0; b < batch; ++b) {
    if (flags[b] == 0) continue;
    for (int k = 0; k < 4; ++k) {
      for (unsigned spins = 1; __atomic_load_n(&red[4 * b + k], __ATOMIC_ACQUIRE) == kUnwritten; ++spins) {
        if ((spins & 4095u) != 0) continue;
        // a failed or finished stream with the value still unwritten ends the wait
        const hipError_t e = hipStreamQuery(s);
        if (e == hipErrorNotReady) continue;
        LITHO_HIP(e);
        if (__atomic_load_n(&red[4 * b + k], __ATOMIC_ACQUIRE) == kUnwritten)
          throw Error(3, "evaluation finished without writing its reductions");
      }
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return plan.host_red->as<double>();
}

const double* fetch_reductions(Plan& plan, int batch) {
  hipStream_t s = plan.ctx->stream;
  LITHO_HIP(hipMemcpyAsync(plan.host_red->ptr, plan.red.ptr, (size_t)batch * 4 * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  spin_wait(s);
  return plan.host_red->as<double>();
}

}  // namespace litho
