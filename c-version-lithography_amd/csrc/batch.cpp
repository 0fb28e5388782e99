// Batched CTM over tiles of several sizes (batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <thread>

#include "opt_kernels.hpp"

namespace litho {

namespace {

void copy_tile(float* dst, const float* src, size_t n, hipStream_t s) {
  LITHO_HIP(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s));
}

// One evaluation of `batch` tiles (theta = initialize_theta(target)) timed with
// HIP events on the plan stream after one untimed warm launch.
double time_eval(Plan& plan, int batch, const float* targets, float* theta, float* grad, const CtmOptions& opt) {
  hipStream_t s = plan.ctx->stream;
  launch_init_theta(targets, theta, (size_t)batch * plan.tile_elems(), s);
  EvalRequest rq;
  rq.mode = 1;
  rq.base = theta;
  rq.im.thr = (float)opt.threshold;
  rq.im.alpha = (float)opt.alpha;
  rq.im.target = targets;
  rq.im.pe_weight = 1.f;
  rq.ga.scale = (float)(4.0 * opt.alpha);
  rq.ga.theta = theta;
  rq.ga.grad_theta = grad;
  EvalGraph g(plan, batch, rq);
  g.launch();
  hipEvent_t e0, e1;
  LITHO_HIP(hipEventCreate(&e0));
  LITHO_HIP(hipEventCreate(&e1));
  constexpr int reps = 3;
  LITHO_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) g.launch();
  LITHO_HIP(hipEventRecord(e1, s));
  LITHO_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LITHO_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

}  // namespace

int autotune_chunk(Plan& plan, const std::vector<const float*>& targets, int limit, const CtmOptions& opt) {
  limit = std::max(1, std::min({limit, plan.capacity, (int)targets.size()}));
  if (plan.tuned_chunk > 0 && plan.tuned_limit == limit) return plan.tuned_chunk;
  hipStream_t s = plan.ctx->stream;
  const size_t n2 = plan.tile_elems();
  DevMem tg((size_t)limit * n2 * sizeof(float)), th((size_t)limit * n2 * sizeof(float)),
      gr((size_t)limit * n2 * sizeof(float));
  for (int i = 0; i < limit; ++i) copy_tile(tg.as<float>() + i * n2, targets[i], n2, s);
  std::vector<std::pair<int, double>> cand;
  for (int c = 1;; c *= 2) {
    const int cc = std::min(c, limit);
    const double ms = time_eval(plan, cc, tg.as<float>(), th.as<float>(), gr.as<float>(), opt);
    cand.emplace_back(cc, ms / cc);
    if (cc == limit) break;
  }
  // the largest chunk within 3 % of the best time per tile: a batch run
  // takes fewer rounds (and launches) with larger chunks
  double best = cand[0].second;
  for (const auto& c : cand) best = std::min(best, c.second);
  int pick = cand.front().first;
  for (const auto& c : cand)
    if (c.second <= 1.03 * best) pick = c.first;
  plan.tuned_chunk = pick;
  plan.tuned_limit = limit;
  return pick;
}

namespace {

// The tiles of one plan through a refilling pool of c slots.
void run_plan_group(Plan& plan, const std::vector<int>& ids, const BatchTile* tiles, const CtmOptions& opt,
                    int chunk, std::vector<CtmTileResult>& res, int* chunk_used) {
  int c = chunk;
  if (c == 0) {
    std::vector<const float*> tg;
    for (int i : ids) tg.push_back(tiles[i].target);
    c = autotune_chunk(plan, tg, (int)ids.size(), opt);
  }
  c = std::max(1, std::min({c, plan.capacity, (int)ids.size()}));
  *chunk_used = c;
  // c slots; a slot whose tile finished takes the next queued tile at once
  CtmPool pool(plan, c, opt);
  size_t next = 0;
  while (next < ids.size() || pool.busy() > 0) {
    while (next < ids.size()) {
      const BatchTile& t = tiles[ids[next]];
      CtmJob j;
      j.id = ids[next];
      j.target = t.target;
      j.theta0 = t.theta0;
      j.out_mask = t.out_mask;
      j.out_theta = t.out_theta;
      if (!pool.admit(j)) break;
      ++next;
    }
    pool.step();
    for (auto& f : pool.take_finished()) res[(size_t)f.first] = std::move(f.second);
  }
  LITHO_HIP(hipStreamSynchronize(plan.ctx->stream));
}

}  // namespace

std::vector<CtmTileResult> run_batch(const std::vector<Plan*>& plans, const BatchTile* tiles, int num_tiles,
                                     const CtmOptions& opt, int chunk, std::vector<int>* chunk_used) {
  validate_ctm_options(opt);
  if (plans.empty()) throw Error(1, "batch: no plans");
  if (num_tiles <= 0 || tiles == nullptr) throw Error(1, "batch: no tiles");
  if (chunk < 0) throw Error(1, "batch: chunk must be >= 0 (0 = autotune)");
  for (const Plan* p : plans)
    if (p == nullptr) throw Error(1, "batch: null plan");
  for (int i = 0; i < num_tiles; ++i) {
    const BatchTile& t = tiles[i];
    if (t.plan < 0 || t.plan >= (int)plans.size()) throw Error(1, "batch: tile plan index out of range");
    if (t.target == nullptr || t.out_mask == nullptr) throw Error(1, "batch: tile target and output are required");
  }
  std::vector<CtmTileResult> res(num_tiles);
  std::vector<int> used(plans.size(), 0);
  // chunk autotune first, one plan at a time: timed side by side, the plans'
  // streams would perturb each other's measurements
  std::vector<int> chunk_of(plans.size(), chunk);
  if (chunk == 0) {
    for (int pi = 0; pi < (int)plans.size(); ++pi) {
      std::vector<const float*> tg;
      for (int i = 0; i < num_tiles; ++i)
        if (tiles[i].plan == pi) tg.push_back(tiles[i].target);
      if (tg.empty()) continue;
      LITHO_HIP(hipSetDevice(plans[pi]->ctx->device));
      chunk_of[pi] = autotune_chunk(*plans[pi], tg, (int)tg.size(), opt);
    }
  }
  // plans on one context run one after another on its stream; distinct
  // contexts run side by side, one host thread each
  std::vector<Context*> ctxs;
  for (Plan* p : plans)
    if (std::find(ctxs.begin(), ctxs.end(), p->ctx) == ctxs.end()) ctxs.push_back(p->ctx);
  auto run_ctx = [&](Context* c) {
    LITHO_HIP(hipSetDevice(c->device));
    for (int pi = 0; pi < (int)plans.size(); ++pi) {
      if (plans[pi]->ctx != c) continue;
      std::vector<int> ids;
      for (int i = 0; i < num_tiles; ++i)
        if (tiles[i].plan == pi) ids.push_back(i);
      if (!ids.empty()) run_plan_group(*plans[pi], ids, tiles, opt, chunk_of[pi], res, &used[pi]);
    }
  };
  if (ctxs.size() == 1) {
    run_ctx(ctxs[0]);
  } else {
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(ctxs.size());
    for (size_t k = 0; k < ctxs.size(); ++k)
      th.emplace_back([&, k] {
        try {
          run_ctx(ctxs[k]);
        } catch (...) {
          err[k] = std::current_exception();
        }
      });
    for (auto& t : th) t.join();
    for (auto& e : err)
      if (e) std::rethrow_exception(e);
  }
  if (chunk_used != nullptr) *chunk_used = used;
  return res;
}

}  // namespace litho
