// C ABI entry points (include/litho_amd.h).  Every call converts exceptions
// into status codes and a thread-local message.
#include "../../include/litho_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batch.hpp"
#include "binarize.hpp"
#include "epe.hpp"
#include "ctm_driver.hpp"
#include "fullchip.hpp"
#include "io.hpp"
#include "lsm.hpp"
#include "noalloc.hpp"
#include "optics.hpp"
#include "opt_kernels.hpp"
#include "plan.hpp"
#include "server.hpp"
#include "tiling.hpp"

struct litho_ctx {
  litho::Context c;
};

struct litho_ep_set {
  litho::EpSet* s;
};

struct litho_plan {
  litho::Plan* p;
};

struct litho_server {
  litho::Server* s;
};

namespace {

thread_local std::string g_last_error;
thread_local double g_last_eigen_gap = 0.0;

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return LITHO_OK;
  } catch (const litho::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "out of host memory";
    return LITHO_ENOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return LITHO_ERUNTIME;
  }
}

void require(bool c, const char* msg) {
  if (!c) throw litho::Error(LITHO_EINVAL, msg);
}

litho::Plan& plan_of(litho_plan* p) {
  require(p != nullptr && p->p != nullptr, "null plan");
  return *p->p;
}

void set_device(const litho::Plan& p) { LITHO_HIP(hipSetDevice(p.ctx->device)); }

litho::Context& ctx_of(litho_ctx* c) {
  require(c != nullptr, "null context");
  LITHO_HIP(hipSetDevice(c->c.device));
  return c->c;
}

}  // namespace

namespace {

litho::CtmOptions ctm_options(const litho_ctm_options& o) {
  litho::CtmOptions c;
  c.method = o.method;
  c.iterations = o.iterations;
  c.learning_rate = o.learning_rate;
  c.history = o.lbfgs_history;
  c.tolerance = o.gradient_tolerance;
  c.threshold = o.threshold;
  c.alpha = o.alpha;
  return c;
}
litho::LsmOptions lsm_options(const litho_lsm_options& o) {
  litho::LsmOptions l;
  l.dx = o.dx;
  l.dy = o.dy;
  l.cfl = o.cfl;
  l.curvature_weight = o.curvature_weight;
  l.mask_type = o.mask_type;
  l.render_scale = o.render_scale;
  return l;
}
void fill_info(const litho::CtmTileResult& r, litho_ctm_info* info) {
  info->evaluations = r.evaluations;
  info->iterations = r.iterations;
  info->final_pe = r.final_pe;
  info->final_grad_norm = r.final_grad_norm;
  info->status = r.status;
}
}  // namespace

extern "C" {

const char* litho_last_error(void) { return g_last_error.c_str(); }

int litho_supported_size(int n) { return litho::socs_ops_for(n) != nullptr ? 1 : 0; }

int litho_ctx_create(int device, litho_ctx** out) {
  return guarded([&] {
    require(out != nullptr, "null output handle");
    int count = 0;
    LITHO_HIP(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, "device index out of range");
    LITHO_HIP(hipSetDevice(device));
    litho::noalloc_init(device);   // this device's address of the allocation-free entry points' scratch
    auto* c = new litho_ctx();
    c->c.device = device;
    if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      throw litho::Error(LITHO_EDEVICE, "hipStreamCreate failed");
    }
    *out = c;
  });
}

int litho_ctx_destroy(litho_ctx* ctx) {
  return guarded([&] {
    if (ctx == nullptr) return;
    const int live = ctx->c.live_plans.load();
    if (live != 0)
      throw litho::Error(LITHO_EINVAL, "context still has " + std::to_string(live) +
                                           " plan(s): destroy them before their context");
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
  });
}

int litho_ctx_synchronize(litho_ctx* ctx) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    LITHO_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

void* litho_ctx_stream(litho_ctx* ctx) { return ctx != nullptr ? (void*)ctx->c.stream : nullptr; }

int litho_ctx_set_graphs(litho_ctx* ctx, int enable) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    ctx->c.graphs = enable != 0;
  });
}

int litho_ctx_graph_counts(litho_ctx* ctx, long long* captured, long long* failed, long long* direct) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    if (captured != nullptr) *captured = c.graphs_captured.load();
    if (failed != nullptr) *failed = c.graphs_failed.load();
    if (direct != nullptr) *direct = c.graphs_direct.load();
  });
}

int litho_ctx_time_pair_kernel(litho_ctx* ctx, int enable) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    c.time_pair = enable != 0;
    c.pair_launches.store(0);
    c.pair_ms.store(0.0);
  });
}

int litho_ctx_pair_kernel_time(litho_ctx* ctx, long long* launches, double* total_ms) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    if (launches != nullptr) *launches = c.pair_launches.load();
    if (total_ms != nullptr) *total_ms = c.pair_ms.load();
  });
}

int litho_socs_window(int n, double pixel_nm, double wavelength_nm, double na, double sigma_in, double sigma_out,
                      int* win_rows, int* win_cols, int* win_y0, int* win_x0, int* num_source_points,
                      double* weight_sum) {
  return guarded([&] {
    litho::OpticsParams op{n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out, {}};
    const litho::SourceSet src = litho::make_source_set(op);
    const litho::Window w = litho::socs_window(op, src);
    if (win_rows) *win_rows = w.rows;
    if (win_cols) *win_cols = w.cols;
    if (win_y0) *win_y0 = w.y0;
    if (win_x0) *win_x0 = w.x0;
    if (num_source_points) *num_source_points = (int)src.fs.size();
    if (weight_sum) *weight_sum = src.weight_sum;
  });
}

double litho_socs_last_eigen_gap(void) { return g_last_eigen_gap; }

int litho_socs_precompute(litho_ctx* ctx, int n, double pixel_nm, double wavelength_nm, double na, double sigma_in,
                          double sigma_out, int k, double* kernels_out, double* weights_out,
                          double* singular_values_out) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    require(kernels_out != nullptr && weights_out != nullptr, "null output arrays");
    LITHO_HIP(hipSetDevice(ctx->c.device));
    litho::OpticsParams op{n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out, {}};
    g_last_eigen_gap = 0.0;
    litho::socs_precompute_device(op, k, ctx->c.stream, kernels_out, weights_out, singular_values_out,
                                  &g_last_eigen_gap);
  });
}

int litho_socs_precompute_zernike(litho_ctx* ctx, int n, double pixel_nm, double wavelength_nm, double na,
                                  double sigma_in, double sigma_out, const int* zernike_index,
                                  const double* zernike_coeff, int num_terms, int k, double* kernels_out,
                                  double* weights_out, double* singular_values_out) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    require(kernels_out != nullptr && weights_out != nullptr, "null output arrays");
    require(num_terms >= 0 && (num_terms == 0 || (zernike_index != nullptr && zernike_coeff != nullptr)),
            "Zernike terms: null arrays");
    LITHO_HIP(hipSetDevice(ctx->c.device));
    litho::OpticsParams op{n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out, {}};
    for (int i = 0; i < num_terms; ++i) op.zernike.emplace_back(zernike_index[i], zernike_coeff[i]);
    g_last_eigen_gap = 0.0;
    litho::socs_precompute_device(op, k, ctx->c.stream, kernels_out, weights_out, singular_values_out,
                                  &g_last_eigen_gap);
  });
}

int litho_plan_create(litho_ctx* ctx, int n, int k, int win_rows, int win_cols, int win_y0, int win_x0,
                      const double* kernels, const double* weights, double weight_sum, int batch_capacity,
                      litho_plan** out) {
  return guarded([&] {
    require(ctx != nullptr && out != nullptr, "null context or output handle");
    litho::Plan* p = litho::make_plan(&ctx->c, n, k, win_rows, win_cols, win_y0, win_x0, kernels, weights,
                                      weight_sum, batch_capacity);
    ctx->c.live_plans.fetch_add(1);
    *out = new litho_plan{p};
  });
}

int litho_plan_destroy(litho_plan* plan) {
  return guarded([&] {
    if (plan == nullptr) return;
    if (plan->p) {
      litho::Context* c = plan->p->ctx;
      (void)hipSetDevice(c->device);
      (void)hipStreamSynchronize(c->stream);
      delete plan->p;
      c->live_plans.fetch_sub(1);
    }
    delete plan;
  });
}

size_t litho_plan_device_bytes(const litho_plan* plan) {
  return plan != nullptr && plan->p != nullptr ? plan->p->device_bytes : 0;
}

int litho_plan_column_path(const litho_plan* plan, int* direct_dft, int* corr_len, int* adjoint_mfma) {
  return guarded([&] {
    require(plan != nullptr && plan->p != nullptr, "plan is null");
    if (direct_dft != nullptr) *direct_dft = plan->p->generic;
    if (corr_len != nullptr) *corr_len = plan->p->corr_len;
    if (adjoint_mfma != nullptr) *adjoint_mfma = plan->p->adj_mfma;
  });
}

int litho_socs_forward(litho_plan* plan, int batch, const float* mask, float* aerial, float* wafer, double threshold,
                       double alpha) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(mask != nullptr, "Imaging: mask is required");
    require(std::isfinite(threshold) && std::isfinite(alpha), "Imaging: mask, threshold, and alpha must be finite");
    set_device(p);
    litho::EvalRequest rq;
    rq.mode = 0;
    rq.base = mask;
    rq.adjoint = false;
    rq.im.thr = (float)threshold;
    rq.im.alpha = (float)alpha;
    rq.im.aerial = aerial;
    rq.im.wafer = wafer;
    litho::run_eval(p, batch, rq);
    LITHO_HIP(hipStreamSynchronize(p.ctx->stream));
  });
}

int litho_pe_gradient(litho_plan* plan, int batch, const float* mask, const float* target, const float* pe_window,
                      const float* penalty_window, double pe_weight, double penalty, double penalty_threshold,
                      double threshold, double alpha, float* aerial, float* wafer, float* grad_mask, double* pe_loss,
                      double* bg_loss) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(mask != nullptr && target != nullptr, "Gradient: mask and target are required");
    require(std::isfinite(threshold) && std::isfinite(alpha), "Imaging: mask, threshold, and alpha must be finite");
    set_device(p);
    litho::EvalRequest rq;
    rq.mode = 0;
    rq.base = mask;
    rq.im.thr = (float)threshold;
    rq.im.alpha = (float)alpha;
    rq.im.aerial = aerial;
    rq.im.wafer = wafer;
    rq.im.target = target;
    rq.im.pe_win = pe_window;
    rq.im.pen_win = penalty_window;
    rq.im.pe_weight = (float)pe_weight;
    rq.im.penalty = (float)penalty;
    rq.im.pen_thr = (float)penalty_threshold;
    rq.ga.scale = (float)(4.0 * alpha);
    rq.ga.grad_mask = grad_mask;
    litho::run_eval(p, batch, rq);
    const double* red = litho::fetch_reductions(p, batch);
    for (int b = 0; b < batch; ++b) {
      if (pe_loss) pe_loss[b] = red[4 * b + 0];
      if (bg_loss) bg_loss[b] = red[4 * b + 1];
    }
  });
}

int litho_ctm_eval(litho_plan* plan, int batch, const float* theta, const float* target, double threshold,
                   double alpha, float* mask, float* grad_theta, double* pe, double* grad_norm2) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(theta != nullptr && target != nullptr, "CTM: theta and target are required");
    set_device(p);
    litho::EvalRequest rq;
    rq.mode = 1;
    rq.base = theta;
    rq.mask_out = mask;
    rq.im.thr = (float)threshold;
    rq.im.alpha = (float)alpha;
    rq.im.target = target;
    rq.im.pe_weight = 1.f;
    rq.ga.scale = (float)(4.0 * alpha);
    rq.ga.theta = theta;
    rq.ga.grad_theta = grad_theta;
    litho::run_eval(p, batch, rq);
    const double* red = litho::fetch_reductions(p, batch);
    for (int b = 0; b < batch; ++b) {
      if (pe) pe[b] = red[4 * b + 0];
      if (grad_norm2) grad_norm2[b] = red[4 * b + 2];
    }
  });
}

int litho_batch_ctm(litho_plan* const* plans, int num_plans, const litho_batch_tile* tiles, int num_tiles,
                    const litho_ctm_options* options, int chunk, litho_ctm_info* info, int* chunk_used) {
  return guarded([&] {
    require(plans != nullptr && num_plans > 0 && options != nullptr, "null argument");
    std::vector<litho::Plan*> ps;
    for (int i = 0; i < num_plans; ++i) ps.push_back(&plan_of(plans[i]));
    set_device(*ps[0]);
    require(tiles != nullptr && num_tiles > 0, "no tiles");
    std::vector<litho::BatchTile> bt(num_tiles);
    for (int i = 0; i < num_tiles; ++i) {
      bt[i].plan = tiles[i].plan_index;
      bt[i].target = tiles[i].target;
      bt[i].theta0 = tiles[i].theta0;
      bt[i].out_mask = tiles[i].out_mask;
      bt[i].out_theta = tiles[i].out_theta;
    }
    litho::CtmOptions o;
    o.method = options->method;
    o.iterations = options->iterations;
    o.learning_rate = options->learning_rate;
    o.history = options->lbfgs_history;
    o.tolerance = options->gradient_tolerance;
    o.threshold = options->threshold;
    o.alpha = options->alpha;
    require(o.method == 0 || o.method == 1, "CTM optimizer must be gradient_descent or lbfgs");
    std::vector<int> used;
    auto res = litho::run_batch(ps, bt.data(), num_tiles, o, chunk, &used);
    for (int i = 0; i < num_tiles; ++i)
      if (info != nullptr) {
        info[i].evaluations = res[i].evaluations;
        info[i].iterations = res[i].iterations;
        info[i].final_pe = res[i].final_pe;
        info[i].final_grad_norm = res[i].final_grad_norm;
        info[i].status = res[i].status;
      }
    if (chunk_used != nullptr)
      for (int i = 0; i < num_plans; ++i) chunk_used[i] = used[i];
  });
}

int litho_server_create(litho_plan* const* plans, int num_plans, const litho_server_options* options,
                        litho_server** out) {
  return guarded([&] {
    require(plans != nullptr && num_plans > 0 && options != nullptr && out != nullptr, "null argument");
    std::vector<litho::Plan*> ps;
    for (int i = 0; i < num_plans; ++i) ps.push_back(&plan_of(plans[i]));
    set_device(*ps[0]);
    litho::ServerOptions so;
    so.ctm = ctm_options(options->ctm);
    so.memory_budget = options->memory_budget_bytes;
    auto* h = new litho_server();
    try {
      h->s = new litho::Server(ps, so);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int litho_server_destroy(litho_server* server) {
  return guarded([&] {
    if (server == nullptr) return;
    delete server->s;
    delete server;
  });
}

int litho_server_submit(litho_server* server, int plan_index, const float* target, const float* theta0,
                        float* out_mask, float* out_theta, long long* ticket) {
  return guarded([&] {
    require(server != nullptr && server->s != nullptr && ticket != nullptr, "null argument");
    *ticket = server->s->submit(plan_index, target, theta0, out_mask, out_theta);
  });
}

int litho_server_wait(litho_server* server, long long ticket, double timeout_s, litho_server_result* result) {
  int timed_out = 0;
  const int rc = guarded([&] {
    require(server != nullptr && server->s != nullptr, "null server");
    litho::ServedResult r;
    if (!server->s->wait(ticket, timeout_s, &r)) {
      timed_out = 1;
      return;
    }
    if (r.error != 0) throw litho::Error(r.error, "server: request failed: " + server->s->error_message());
    if (result != nullptr) {
      fill_info(r.result, &result->info);
      result->queued_ms = r.queued_ms;
      result->service_ms = r.service_ms;
    }
  });
  if (rc == LITHO_OK && timed_out) {
    g_last_error = "request not finished";
    return LITHO_ETIMEDOUT;
  }
  return rc;
}

size_t litho_server_request_bytes(const litho_server* server, int plan_index) {
  if (server == nullptr || server->s == nullptr || plan_index < 0) return 0;
  try {
    return server->s->request_bytes(plan_index);
  } catch (...) {
    return 0;
  }
}

int litho_ctm_optimize(litho_plan* plan, const float* target, const float* theta0, const litho_ctm_options* options,
                       float* out_mask, float* out_theta, double* history, int history_capacity,
                       litho_ctm_info* info) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(options != nullptr, "null options");
    set_device(p);
    litho::CtmOptions o;
    o.method = options->method;
    o.iterations = options->iterations;
    o.learning_rate = options->learning_rate;
    o.history = options->lbfgs_history;
    o.tolerance = options->gradient_tolerance;
    o.threshold = options->threshold;
    o.alpha = options->alpha;
    auto res = litho::run_ctm(p, 1, target, theta0, o, out_mask, out_theta);
    const auto& r = res[0];
    if (history != nullptr)
      for (int i = 0; i < (int)r.history.size() && i < history_capacity; ++i) history[i] = r.history[i];
    if (info != nullptr) {
      info->evaluations = r.evaluations;
      info->iterations = r.iterations;
      info->final_pe = r.final_pe;
      info->final_grad_norm = r.final_grad_norm;
      info->status = r.status;
    }
  });
}

int litho_profile_ctm_eval(litho_plan* plan, int batch, const float* theta, const float* target, double threshold,
                           double alpha, int repeats, double* stage_ms) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(theta != nullptr && target != nullptr && stage_ms != nullptr, "null argument");
    require(batch >= 1 && batch <= p.capacity, "batch exceeds the plan capacity");
    set_device(p);
    hipStream_t s = p.ctx->stream;
    // the evaluation of an L-BFGS trial (LbfgsPool in ctm_driver.cpp):
    // theta = base + step * d (step 0, d = 0), theta written out, the theta
    // gradient stored and g . d formed
    const size_t n2 = p.tile_elems() * (size_t)batch;
    litho::DevMem x(n2 * sizeof(float)), g(n2 * sizeof(float)), d(n2 * sizeof(float)), step(batch * sizeof(float));
    LITHO_HIP(hipMemsetAsync(d.ptr, 0, d.bytes, s));
    LITHO_HIP(hipMemsetAsync(step.ptr, 0, step.bytes, s));
    litho::EvalRequest rq;
    rq.mode = 1;
    rq.base = theta;
    rq.dir = d.as<float>();
    rq.coef = step.as<float>();
    rq.theta_out = x.as<float>();
    rq.im.thr = (float)threshold;
    rq.im.alpha = (float)alpha;
    rq.im.target = target;
    rq.im.pe_weight = 1.f;
    rq.ga.scale = (float)(4.0 * alpha);
    rq.ga.theta = x.as<float>();
    rq.ga.grad_theta = g.as<float>();
    rq.ga.dir = d.as<float>();
    litho::profile_eval(p, batch, rq, repeats, stage_ms);
  });
}

int litho_profile_lbfgs_pair(litho_ctx* ctx, size_t n, int m, int repeats, double* ms) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    require(ms != nullptr && repeats > 0, "null argument or non-positive repeats");
    require(n > 0 && n % 4 == 0, "n must be a positive multiple of 4");
    require(m >= 1 && m <= litho::kMaxHist, "history m out of range");
    hipStream_t s = c.stream;
    const int P = litho::lbfgs_points(m);
    litho::DevMem X, G, need, meta, part;
    for (litho::DevMem* v : {&X, &G}) {
      v->alloc((size_t)P * n * sizeof(float));
      LITHO_HIP(hipMemsetAsync(v->ptr, 0, v->bytes, s));
    }
    // one tile with a full history of m chained pairs (points 0 .. m, pair j
    // from point j to j + 1), base point m, trial point m + 1
    std::vector<int> hm(litho::kMetaStride, -1);
    hm[0] = m;
    hm[litho::kMetaBase] = m;
    hm[litho::kMetaTrial] = m + 1;
    for (int j = 0; j < m; ++j) {
      hm[litho::kMetaStart + j] = j;
      hm[litho::kMetaEnd + j] = j + 1;
    }
    const int one = 1;
    need.alloc(sizeof(int));
    meta.alloc(hm.size() * sizeof(int));
    part.alloc((size_t)litho::kOptBlocks * litho::kPairDots * sizeof(double));
    LITHO_HIP(hipMemcpyAsync(need.ptr, &one, sizeof(int), hipMemcpyHostToDevice, s));
    LITHO_HIP(hipMemcpyAsync(meta.ptr, hm.data(), meta.bytes, hipMemcpyHostToDevice, s));
    litho::DevMem sink;
    sink.alloc(sizeof(float));
    auto launch = [&] {
      litho::launch_lbfgs_pair(X.as<float>(), G.as<float>(), P, need.as<int>(), meta.as<int>(), n, 1, m,
                               part.as<double>(), s);
    };
    // each timed launch follows a 1 GiB read (four times the 256 MB
    // last-level cache), so it streams its vectors cold, as the sweeps of a
    // solve do after an evaluation; back to back, part of them would still be
    // cached from the launch before
    constexpr size_t kEvictBytes = size_t(1) << 30;
    litho::DevMem evict;
    evict.alloc(kEvictBytes);
    LITHO_HIP(hipMemsetAsync(evict.ptr, 0, evict.bytes, s));
    launch();   // untimed: loads the code object
    hipEvent_t e0, e1;
    LITHO_HIP(hipEventCreate(&e0));
    LITHO_HIP(hipEventCreate(&e1));
    double total = 0.0;
    for (int r = 0; r < repeats; ++r) {
      litho::launch_cache_evict(evict.as<float>(), kEvictBytes / 16, sink.as<float>(), s);
      LITHO_HIP(hipEventRecord(e0, s));
      launch();
      LITHO_HIP(hipEventRecord(e1, s));
      LITHO_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      LITHO_HIP(hipEventElapsedTime(&t, e0, e1));
      total += t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = total / repeats;
  });
}

int litho_profile_stream(litho_ctx* ctx, size_t n, int reads, int writes, int repeats, double* ms) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    require(ms != nullptr && repeats > 0, "null argument or non-positive repeats");
    require(n > 0 && n % 4 == 0, "n must be a positive multiple of 4");
    require((reads == 24 && writes == 0) || (reads == 22 && writes == 1) || (reads == 24 && writes == 4) ||
                (reads == 21 && writes == 1),
            "stream pattern: (reads, writes) must be (24, 0), (22, 1), (24, 4) or (21, 1)");
    hipStream_t s = c.stream;
    std::vector<litho::DevMem> bufs((size_t)(reads + writes));
    std::vector<const float*> in;
    std::vector<float*> out;
    for (size_t j = 0; j < bufs.size(); ++j) {
      bufs[j].alloc(n * sizeof(float));
      LITHO_HIP(hipMemsetAsync(bufs[j].ptr, 0, bufs[j].bytes, s));
      if ((int)j < reads)
        in.push_back(bufs[j].as<float>());
      else
        out.push_back(bufs[j].as<float>());
    }
    litho::DevMem sink;
    sink.alloc(sizeof(float));
    auto launch = [&] {
      if (!litho::launch_stream_pattern(reads, writes, in.data(), out.data(), n / 4, sink.as<float>(), s))
        throw litho::Error(LITHO_EINVAL, "stream pattern not supported");
    };
    // each timed launch follows a 1 GiB read, as the pair kernel's profile
    constexpr size_t kEvictBytes = size_t(1) << 30;
    litho::DevMem evict;
    evict.alloc(kEvictBytes);
    LITHO_HIP(hipMemsetAsync(evict.ptr, 0, evict.bytes, s));
    launch();   // untimed: loads the code object
    hipEvent_t e0, e1;
    LITHO_HIP(hipEventCreate(&e0));
    LITHO_HIP(hipEventCreate(&e1));
    double total = 0.0;
    for (int r = 0; r < repeats; ++r) {
      litho::launch_cache_evict(evict.as<float>(), kEvictBytes / 16, sink.as<float>(), s);
      LITHO_HIP(hipEventRecord(e0, s));
      launch();
      LITHO_HIP(hipEventRecord(e1, s));
      LITHO_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      LITHO_HIP(hipEventElapsedTime(&t, e0, e1));
      total += t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = total / repeats;
  });
}

int litho_lbfgs_direction(litho_ctx* ctx, size_t n, int m, int num_points, const float* x_points,
                          const float* g_points, float* d_out, double* gd_out, int* kept_out) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    require(x_points != nullptr && g_points != nullptr && d_out != nullptr, "null argument");
    require(n > 0 && n % 4 == 0, "n must be a positive multiple of 4");
    require(m >= 1 && m <= litho::kMaxHist, "history m out of range");
    require(num_points >= 1, "at least one point is required");
    hipStream_t s = c.stream;
    LITHO_HIP(hipSetDevice(c.device));
    // the buffers of one L-BFGS slot (LbfgsPool in ctm_driver.cpp)
    const int P = litho::lbfgs_points(m);
    litho::DevMem X, G, need, meta, sy, yy, coef, pair_part, dir_part, vec_red, gd;
    for (litho::DevMem* v : {&X, &G}) v->alloc((size_t)P * n * sizeof(float));
    need.alloc(sizeof(int));
    meta.alloc(litho::kMetaStride * sizeof(int));
    sy.alloc(litho::kMaxHist * litho::kMaxHist * sizeof(double));
    yy.alloc(litho::kMaxHist * litho::kMaxHist * sizeof(double));
    coef.alloc(litho::kCoefStride * sizeof(double));
    pair_part.alloc((size_t)litho::kOptBlocks * litho::kPairDots * sizeof(double));
    dir_part.alloc((size_t)litho::kOptBlocks * sizeof(double));
    vec_red.alloc(litho::kPairDots * sizeof(double));
    gd.alloc(sizeof(double));
    for (litho::DevMem* v : {&meta, &sy, &yy, &coef}) LITHO_HIP(hipMemsetAsync(v->ptr, 0, v->bytes, s));
    litho::launch_lbfgs_reset(meta.as<int>(), 0, s);
    const int one = 1;
    LITHO_HIP(hipMemcpyAsync(need.ptr, &one, sizeof(int), hipMemcpyHostToDevice, s));
    // point p of the slot <- iterate j
    auto put = [&](int p, int j) {
      LITHO_HIP(hipMemcpyAsync(X.as<float>() + (size_t)p * n, x_points + (size_t)j * n, n * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
      LITHO_HIP(hipMemcpyAsync(G.as<float>() + (size_t)p * n, g_points + (size_t)j * n, n * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
    };
    put(0, 0);   // the base
    int mb = 0;
    std::vector<int> hm(litho::kMetaStride);
    for (int j = 1; j < num_points; ++j) {
      // iterate j as the accepted trial: pair (x_j - x_{j-1}, g_j - g_{j-1});
      // the dots of the history with g_j give the coefficients of its direction
      LITHO_HIP(hipMemcpyAsync(hm.data(), meta.ptr, meta.bytes, hipMemcpyDeviceToHost, s));
      LITHO_HIP(hipStreamSynchronize(s));
      put(hm[litho::kMetaTrial], j);
      mb = std::min(j - 1, m);
      litho::launch_lbfgs_pair(X.as<float>(), G.as<float>(), P, need.as<int>(), meta.as<int>(), n, 1, mb,
                               pair_part.as<double>(), s);
      litho::launch_reduce_fixed(pair_part.as<double>(), litho::kOptBlocks, 4 * (mb + 1), vec_red.as<double>(),
                                 litho::kPairDots, 1, s);
      litho::launch_lbfgs_update(1, need.as<int>(), vec_red.as<double>(), mb, m, meta.as<int>(), sy.as<double>(),
                                 yy.as<double>(), coef.as<double>(), s);
    }
    if (num_points == 1) {
      // no pair: d = -g (cg = -1)
      const double cg = -1.0;
      LITHO_HIP(hipMemcpyAsync(coef.ptr, &cg, sizeof(double), hipMemcpyHostToDevice, s));
      mb = -1;
    }
    const int mb_dir = std::min(mb + 1, m);
    litho::launch_lbfgs_direction(X.as<float>(), G.as<float>(), P, need.as<int>(), meta.as<int>(),
                                  coef.as<double>(), n, d_out, 1, mb_dir, dir_part.as<double>(), s);
    litho::launch_reduce_fixed(dir_part.as<double>(), litho::kOptBlocks, 1, gd.as<double>(), 1, 1, s);
    double gdv = 0.0;
    int kept = 0;
    LITHO_HIP(hipMemcpyAsync(&gdv, gd.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    LITHO_HIP(hipMemcpyAsync(&kept, meta.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    LITHO_HIP(hipStreamSynchronize(s));
    if (gd_out != nullptr) *gd_out = gdv;
    if (kept_out != nullptr) *kept_out = kept;
  });
}

int litho_initialize_theta(litho_ctx* ctx, const float* pattern, float* theta, size_t count) {
  return guarded([&] {
    require(ctx != nullptr && pattern != nullptr && theta != nullptr, "null argument");
    LITHO_HIP(hipSetDevice(ctx->c.device));
    litho::launch_init_theta(pattern, theta, count, ctx->c.stream);
    LITHO_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int litho_ls_terms(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, const float* vn, double dx,
                   double dy, double curvature_weight, float* delta, float* kappa) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    litho::ls_terms(c, batch, rows, cols, phi, vn, dx, dy, curvature_weight, delta, kappa);
    LITHO_HIP(hipStreamSynchronize(c.stream));
  });
}

double litho_ls_dt(double cfl, double dx, double dy, double max_h1, double max_h2, double curvature_weight) {
  return litho::ls_dt(cfl, dx, dy, max_h1, max_h2, curvature_weight);
}

int litho_ls_reinit_sdf(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, double dx, double dy,
                        float* sdf) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    litho::ls_reinit_sdf(c, batch, rows, cols, phi, dx, dy, sdf);
    LITHO_HIP(hipStreamSynchronize(c.stream));
  });
}

int litho_ls_reinit_sdf_noalloc(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, double dx,
                                double dy, float* sdf) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    LITHO_HIP(hipSetDevice(c.device));
    litho::ls_reinit_sdf_noalloc(c, batch, rows, cols, phi, dx, dy, sdf);
  });
}

int litho_render_levelset_mask(litho_ctx* ctx, int batch, int rows, int cols, const float* phi, int type, int scale,
                               float* mask_out) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    litho::render_levelset_mask(c, batch, rows, cols, phi, type, scale, mask_out);
    LITHO_HIP(hipStreamSynchronize(c.stream));
  });
}

int litho_lsm_step(litho_ctx* ctx, int batch, int n, const float* phi, const float* mask_gradient,
                   const int* update_rects, const litho_lsm_options* options, int reinit, float* phi_out,
                   float* mask_out, double* dt_out) {
  return guarded([&] {
    require(options != nullptr, "null options");
    litho::Context& c = ctx_of(ctx);
    const litho::LsmOptions o = lsm_options(*options);
    litho::lsm_step(c, batch, n, phi, mask_gradient, update_rects, o, reinit != 0, phi_out, mask_out, dt_out);
    LITHO_HIP(hipStreamSynchronize(c.stream));
  });
}

int litho_partition_build(int height, int width, int grid_rows, int grid_cols, int halo, int guard,
                          int* tile_size, litho_tile* tiles) {
  return guarded([&] {
    require(tile_size != nullptr && tiles != nullptr, "null output");
    std::vector<litho_tile> t;
    litho::partition_build(height, width, grid_rows, grid_cols, halo, guard, tile_size, t);
    std::memcpy(tiles, t.data(), t.size() * sizeof(litho_tile));
  });
}

int litho_partition_extract(litho_ctx* ctx, const float* global, int height, int width, const litho_tile* tiles,
                            int num_tiles, int tile_size, float* local_out) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    require(global != nullptr && local_out != nullptr && height > 0 && width > 0 && tile_size > 0, "invalid argument");
    litho::validate_tiles(tiles, num_tiles, tile_size);
    const std::vector<int> tab = litho::tile_table(tiles, num_tiles);
    litho::DevMem d(tab.size() * sizeof(int));
    LITHO_HIP(hipMemcpyAsync(d.ptr, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
    litho::launch_extract(global, height, width, d.as<int>(), num_tiles, tile_size, local_out, c.stream);
    LITHO_HIP(hipGetLastError());
    LITHO_HIP(hipStreamSynchronize(c.stream));
  });
}

int litho_partition_assemble(litho_ctx* ctx, const float* local, const litho_tile* tiles, int num_tiles,
                             int tile_size, int height, int width, float* global_out) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    require(local != nullptr && global_out != nullptr && height > 0 && width > 0, "invalid argument");
    litho::validate_tiles(tiles, num_tiles, tile_size);
    for (int i = 0; i < num_tiles; ++i) {
      const litho_rect& r = tiles[i].core_global;
      require(r.y0 >= 0 && r.x0 >= 0 && r.y1 <= height && r.x1 <= width, "Micro Core outside the chip");
    }
    const std::vector<int> tab = litho::tile_table(tiles, num_tiles);
    litho::DevMem d(tab.size() * sizeof(int));
    LITHO_HIP(hipMemcpyAsync(d.ptr, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
    litho::launch_assemble(local, d.as<int>(), num_tiles, tile_size, width, global_out, tile_size * tile_size,
                           c.stream);
    LITHO_HIP(hipGetLastError());
    LITHO_HIP(hipStreamSynchronize(c.stream));
  });
}

int litho_tile_exchange(litho_plan* plan, const litho_tile* tiles, int num_tiles, const float* local_masks,
                        const float* local_targets, const litho_exchange_options* options, float* mask_gradient,
                        float* raw_gradient, float* wafer, litho_tile_loss* per_tile, double* totals) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(options != nullptr, "null options");
    set_device(p);
    litho::Exchanger ex(p, tiles, num_tiles);
    ex.prepare(local_targets, *options);
    ex.run(local_masks, *options, mask_gradient, raw_gradient, wafer, per_tile, totals);
  });
}

int litho_fullchip_ctm(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                       const float* initial_pattern, const float* global_target, const litho_exchange_options* options,
                       int iterations, double learning_rate, float* global_mask_out, float* local_masks_out,
                       litho_fullchip_row* history, int* best_iteration, double* best_loss) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(options != nullptr, "null options");
    set_device(p);
    const litho::FullChipResult r =
        litho::fullchip_ctm(p, tiles, num_tiles, height, width, initial_pattern, global_target, *options, iterations,
                            learning_rate, global_mask_out, local_masks_out, history);
    if (best_iteration != nullptr) *best_iteration = r.best_iteration;
    if (best_loss != nullptr) *best_loss = r.best_loss;
  });
}

int litho_micro_imaging(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                        const float* global_mask, double threshold, double alpha, float* global_aerial,
                        float* global_wafer, float* local_aerial, float* local_wafer) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    set_device(p);
    litho::micro_imaging(p, tiles, num_tiles, height, width, global_mask, threshold, alpha, global_aerial,
                         global_wafer, local_aerial, local_wafer);
  });
}

int litho_fullchip_lsm(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                       const float* initial_binary, const float* global_target, const litho_exchange_options* options,
                       const litho_lsm_options* lsm, int iterations, int reinit_interval, float* global_mask_out,
                       float* global_phi_out, litho_fullchip_row* history, int* best_iteration, double* best_loss) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(options != nullptr && lsm != nullptr, "null options");
    set_device(p);
    const litho::LsmOptions lo = lsm_options(*lsm);
    const litho::FullChipResult r =
        litho::fullchip_lsm(p, tiles, num_tiles, height, width, initial_binary, global_target, *options, lo,
                            iterations, reinit_interval, global_mask_out, global_phi_out, history);
    if (best_iteration != nullptr) *best_iteration = r.best_iteration;
    if (best_loss != nullptr) *best_loss = r.best_loss;
  });
}


int litho_binarize_advanced(litho_ctx* ctx, int height, int width, const float* gray_mask, const float* target_pattern,
                            double peak_min_intensity, double quantile_val, int area_threshold,
                            int min_spacing_pixels, float* final_mask, float* initial_binary_mask,
                            float* filtered_binary_mask, float* sraf_mask_initial, float* sraf_mask_final,
                            float* keep_out_zone, litho_binarize_info* info) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    LITHO_HIP(hipSetDevice(c.device));
    litho::BinarizeOutputs o;
    o.final_mask = final_mask;
    o.initial_mask = initial_binary_mask;
    o.filtered_mask = filtered_binary_mask;
    o.sraf_initial = sraf_mask_initial;
    o.sraf_final = sraf_mask_final;
    o.keep_out = keep_out_zone;
    const litho::BinarizeInfo r = litho::binarize_advanced(c, height, width, gray_mask, target_pattern,
                                                           peak_min_intensity, quantile_val, area_threshold,
                                                           min_spacing_pixels, o);
    if (info != nullptr) {
      info->threshold = r.threshold;
      info->num_peaks = r.num_peaks;
    }
  });
}

int litho_compare_micro_imaging(litho_plan* plan, const litho_tile* tiles, int num_tiles, int height, int width,
                                const float* global_target, const float* optimized_local_masks, double threshold,
                                double alpha, float* optimized_mask, float* target_aerial, float* target_wafer,
                                float* optimized_aerial, float* optimized_wafer, float* optimized_local_aerial,
                                double* target_pe, double* optimized_pe) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    set_device(p);
    require(tiles != nullptr && num_tiles > 0, "no Micro tiles");
    litho::MicroComparisonOutputs o;
    o.optimized_mask = optimized_mask;
    o.target_aerial = target_aerial;
    o.target_wafer = target_wafer;
    o.optimized_aerial = optimized_aerial;
    o.optimized_wafer = optimized_wafer;
    o.optimized_local_aerial = optimized_local_aerial;
    const litho::MicroComparison r = litho::compare_micro_imaging(p, tiles, num_tiles, height, width, global_target,
                                                                  optimized_local_masks, threshold, alpha, o);
    if (target_pe != nullptr) *target_pe = r.target_pe;
    if (optimized_pe != nullptr) *optimized_pe = r.optimized_pe;
  });
}

int litho_evaluate_mask_continuity(litho_ctx* ctx, const litho_tile* tiles, int num_tiles, int tile_size, int height,
                                   int width, const float* local_masks, const float* target,
                                   litho_mask_continuity* out) {
  return guarded([&] {
    litho::Context& c = ctx_of(ctx);
    require(tiles != nullptr && num_tiles > 0 && out != nullptr, "null tiles or output");
    LITHO_HIP(hipSetDevice(c.device));
    *out = litho::mask_continuity(c, tiles, num_tiles, tile_size, height, width, local_masks, target);
  });
}

int litho_ep_set_create(litho_plan* plan, int count, const double* points_yx, const double* weights,
                        litho_ep_set** out) {
  return guarded([&] {
    litho::Plan& p = plan_of(plan);
    require(out != nullptr, "null output handle");
    set_device(p);
    litho::EpSet* e = litho::make_ep_set(p, count, points_yx, weights);
    *out = new litho_ep_set{e};
  });
}

int litho_ep_set_destroy(litho_ep_set* set) {
  return guarded([&] {
    if (set == nullptr) return;
    if (set->s != nullptr) {
      (void)hipSetDevice(set->s->plan->ctx->device);
      (void)hipStreamSynchronize(set->s->plan->ctx->stream);
      delete set->s;
    }
    delete set;
  });
}

int litho_epe_gradient(litho_ep_set* set, const float* mask, const float* target, const float* pe_window,
                       double pe_weight, double epe_weight, double threshold, double alpha, double pixel_size_nm,
                       float* grad_mask, float* aerial, float* wafer, litho_epe_result* result) {
  return guarded([&] {
    require(set != nullptr && set->s != nullptr, "null edge-point set");
    set_device(*set->s->plan);
    const litho::EpeResult r = litho::epe_gradient(*set->s, mask, target, pe_window, pe_weight, epe_weight, threshold,
                                                   alpha, pixel_size_nm, grad_mask, aerial, wafer);
    if (result != nullptr) {
      result->pe_loss = r.pe_loss;
      result->epe_loss = r.epe_loss;
      result->epe_abs_sum_nm = r.epe_abs_sum_nm;
      result->active_count = r.active;
    }
  });
}

int litho_save_txt(const char* path, const double* matrix, int rows, int cols, int precision) {
  return guarded([&] { litho::save_txt(path, matrix, rows, cols, precision); });
}

int litho_save_txt_int(const char* path, const double* matrix, int rows, int cols) {
  return guarded([&] { litho::save_txt_int(path, matrix, rows, cols); });
}

int litho_load_txt(const char* path, double* matrix, size_t capacity, int* rows, int* cols) {
  return guarded([&] {
    require(rows != nullptr && cols != nullptr, "null row / column outputs");
    std::vector<double> v;
    int r = 0, c = 0;
    litho::load_txt(path, v, &r, &c);
    *rows = r;
    *cols = c;
    if (matrix == nullptr) return;
    require(capacity >= v.size(), "load_txt: the output array is smaller than the matrix");
    std::copy(v.begin(), v.end(), matrix);
  });
}

int litho_save_matrix_txt(const char* path, const double* matrix, int rows, int cols) {
  return guarded([&] { litho::save_matrix_txt(path, matrix, rows, cols); });
}

int litho_write_history_csv(const char* path, int stage, const litho_fullchip_row* rows, int count) {
  return guarded([&] { litho::write_history_csv(path, stage, rows, count); });
}

int litho_write_summary_csv(const char* path, int stage, int best_iteration, double best_loss, int final_iteration,
                            double final_loss, double seconds) {
  return guarded(
      [&] { litho::write_summary_csv(path, stage, best_iteration, best_loss, final_iteration, final_loss, seconds); });
}

int litho_write_imaging_metrics_csv(const char* path, int stage, double target_pe, double optimized_pe) {
  return guarded([&] { litho::write_imaging_metrics_csv(path, stage, target_pe, optimized_pe); });
}

int litho_write_sampling_csv(const char* path, int grid_size, int source_grid_size, int source_points, int kernels,
                             double pixel_size_nm, int micro_core, int image_rows, int image_cols, int micro_count) {
  return guarded([&] {
    litho::write_sampling_csv(path, grid_size, source_grid_size, source_points, kernels, pixel_size_nm, micro_core,
                              image_rows, image_cols, micro_count);
  });
}

}  // extern "C"
