// CTM gradient descent and L-BFGS with device-resident state, as a pool of
// optimizer slots (ctm_driver.hpp).
//
// Gradient descent restates CTM_Optimizer::_optimize_gradient_descent
// (CTM_Optimizer.cpp:109-173).  L-BFGS restates the published algorithm of
// LBFGS++ v0.4.0 (LBFGSSolver<double> with its default line search,
// LineSearchNocedalWright, as constructed at CTM_Optimizer.cpp:194-200): the
// inverse-Hessian product uses the compact representation of Byrd, Nocedal &
// Schnabel (1994), which equals LBFGS++'s two-loop recursion with
// H0 = (s'y / y'y) I; pairs are kept only when s'y > eps * y'y; the first step
// is 1/||d||, later steps 1; the strong-Wolfe line search brackets with
// expansion 2 (Nocedal & Wright Alg. 3.5) and zooms with the
// quadratic-interpolation trial of ls_quad_interp (Alg. 3.6); when
// max_linesearch trials run out it takes the best point of sufficient decrease
// and the solve goes on.  Every objective evaluation (line-search trials
// included) is one grad-eval and appends one PE to the history, like the
// reference's _error_history (CTM_Optimizer.cpp:204).
#include "ctm_driver.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <string>

namespace litho {

double ls_quad_interp(double step_lo, double step_hi, double fx_lo, double fx_hi, double dg_lo) {
  const double fdiff = fx_hi - fx_lo;
  const double sdiff = step_hi - step_lo;
  const double smid = (step_hi + step_lo) / 2.0;
  double cand = fdiff * step_lo - smid * sdiff * dg_lo;
  cand = cand / (fdiff - sdiff * dg_lo);
  if (!std::isfinite(cand)) return smid;
  const double end_dist = std::min(std::abs(cand - step_lo), std::abs(cand - step_hi));
  const bool near_end = end_dist < 0.01 * std::abs(sdiff);
  if (cand <= std::min(step_lo, step_hi) || cand >= std::max(step_lo, step_hi) || near_end) return smid;
  return cand;
}

void validate_ctm_options(const CtmOptions& opt) {
  if (opt.method != 0 && opt.method != 1) throw Error(1, "CTM optimizer must be gradient_descent or lbfgs");
  if (opt.iterations <= 0) throw Error(1, "CTM max_iteration must be positive");
  if (!std::isfinite(opt.learning_rate) || opt.learning_rate <= 0.0)
    throw Error(1, "CTM learning_rate must be finite and positive");
  if (opt.history <= 0) throw Error(1, "CTM lbfgs_history_size must be positive");
  if (opt.method == 1 && opt.history > kMaxHist)
    throw Error(1, "CTM lbfgs_history_size must be in [1, " + std::to_string(kMaxHist) + "]");
  if (!std::isfinite(opt.tolerance) || opt.tolerance < 0.0)
    throw Error(1, "CTM gradient_tolerance must be finite and non-negative");
}

namespace {

// Host-written parameters of the slots (active flags, step coefficients and
// the L-BFGS need-a-new-direction flags), staged in one pinned block and
// uploaded with one transfer per round.  The host rewrites the block only
// after a stream synchronisation that follows its last upload.
struct Params {
  size_t off_step, off_need, bytes;
  HostPinned host;
  DevMem dev;
  explicit Params(int cap)
      : off_step((size_t)cap * sizeof(int)), off_need(off_step + (size_t)cap * sizeof(float)),
        bytes(off_need + (size_t)cap * sizeof(int)), host(bytes), dev(bytes) {}
  int* active() const { return host.as<int>(); }
  float* step() const { return reinterpret_cast<float*>(host.as<char>() + off_step); }
  int* need() const { return reinterpret_cast<int*>(host.as<char>() + off_need); }
  const int* d_active() const { return dev.as<int>(); }
  const float* d_step() const { return reinterpret_cast<const float*>(dev.as<char>() + off_step); }
  const int* d_need() const { return reinterpret_cast<const int*>(dev.as<char>() + off_need); }
  void upload_active_step(hipStream_t s) const {
    LITHO_HIP(hipMemcpyAsync(dev.ptr, host.ptr, bytes, hipMemcpyHostToDevice, s));
  }
};

void copy_tile(float* dst, const float* src, size_t n, hipStream_t s) {
  LITHO_HIP(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s));
}

void alloc_slots(DevMem& m, int slots, size_t n2, size_t per = 1) { m.alloc((size_t)slots * per * n2 * sizeof(float)); }

struct LbfgsTile {
  // INIT: first evaluation; BRACKET / ZOOM: line-search trials; REEVAL: the
  // point of the best sufficient-decrease trial, evaluated again (uncounted)
  // after the trials ran out, since only the last trial's gradient is kept
  enum Phase { INIT, BRACKET, ZOOM, REEVAL, NEED_PAIR, DONE };
  Phase phase = INIT;
  int k = 0;         // LBFGS++ iteration counter
  int ls_iter = 0;   // LineSearchNocedalWright's trial counter
  int pairs = 0;     // pair updates formed so far: the history holds at most min(pairs, m) pairs
  double f = 0, gnorm = 0, f0 = 0, dg0 = 0, test_decr = 0, test_curv = 0, step = 0, a = 0;
  double lo_a = 0, lo_f = 0, lo_g = 0, hi_a = 0, hi_f = 0;
  bool final_is_xp = false;
};

}  // namespace

// ------------------------------------------------------------------ pool base
struct CtmPool::Impl {
  Plan& plan;
  const int S;
  const CtmOptions opt;
  hipStream_t s;
  const size_t n2;
  std::vector<CtmJob> job;
  std::vector<char> running;
  std::vector<CtmTileResult> res;
  std::vector<std::pair<long long, CtmTileResult>> finished;
  DevMem tgt;        // [S][N^2] targets of the slots
  Params pm;

  Impl(Plan& p, int slots, const CtmOptions& o)
      : plan(p), S(slots), opt(o), s(p.ctx->stream), n2(p.tile_elems()), job(slots), running(slots, 0),
        res(slots), pm(slots) {
    alloc_slots(tgt, S, n2);
    for (int b = 0; b < S; ++b) {
      pm.active()[b] = 0;
      pm.step()[b] = 0.f;
      pm.need()[b] = 0;
    }
  }
  virtual ~Impl() = default;
  int busy() const { return (int)std::count(running.begin(), running.end(), 1); }
  int find_free() const {
    for (int b = 0; b < S; ++b)
      if (!running[b]) return b;
    return -1;
  }
  // Copies the job's target into slot b and its initial theta to theta_dst (n2 floats).
  void load_slot(int b, const CtmJob& j, float* theta_dst) {
    if (j.target == nullptr || j.out_mask == nullptr) throw Error(1, "CTM: target and output mask are required");
    copy_tile(tgt.as<float>() + b * n2, j.target, n2, s);
    if (j.theta0 != nullptr)
      copy_tile(theta_dst, j.theta0, n2, s);
    else
      launch_init_theta(j.target, theta_dst, n2, s);
    job[b] = j;
    res[b] = CtmTileResult{};
    running[b] = 1;
  }
  void finish(int b) {
    finished.emplace_back(job[b].id, std::move(res[b]));
    res[b] = CtmTileResult{};
    running[b] = 0;
  }
  EvalRequest base_request(float* mask_out, float* grad_theta) const {
    EvalRequest rq;
    rq.mode = 1;
    rq.mask_out = mask_out;
    rq.im.thr = (float)opt.threshold;
    rq.im.alpha = (float)opt.alpha;
    rq.im.target = tgt.as<float>();
    rq.im.pe_weight = 1.f;
    rq.ga.scale = (float)(4.0 * opt.alpha);
    rq.ga.grad_theta = grad_theta;
    rq.active = pm.d_active();
    rq.red_out = plan.host_red->dev_as<double>();   // read by wait_reductions, no copy
    return rq;
  }
  virtual bool admit(const CtmJob& j) = 0;
  virtual void step() = 0;
};

// ------------------------------------------------------------ gradient descent
namespace {

struct GdPool final : CtmPool::Impl {
  DevMem x, g, mask, best;
  std::vector<double> min_pe;
  std::vector<char> first;
  std::unique_ptr<EvalGraph> graph;

  GdPool(Plan& p, int slots, const CtmOptions& o) : Impl(p, slots, o), min_pe(slots), first(slots, 1) {
    alloc_slots(x, S, n2);
    alloc_slots(g, S, n2);
    alloc_slots(mask, S, n2);
    alloc_slots(best, S, n2);
    // every evaluation: theta = x + coef * g (coef 0 for a slot's first
    // evaluation, -lr after it), gradient to g
    EvalRequest rq = base_request(mask.as<float>(), g.as<float>());
    rq.base = x.as<float>();
    rq.dir = g.as<float>();
    rq.coef = pm.d_step();
    rq.theta_out = x.as<float>();
    rq.ga.theta = x.as<float>();
    graph = std::make_unique<EvalGraph>(plan, S, rq);
  }

  bool admit(const CtmJob& j) override {
    const int b = find_free();
    if (b < 0) return false;
    load_slot(b, j, x.as<float>() + b * n2);
    LITHO_HIP(hipMemsetAsync(g.as<float>() + b * n2, 0, n2 * sizeof(float), s));   // x + 0 * g must be x
    min_pe[b] = std::numeric_limits<double>::max();
    first[b] = 1;
    return true;
  }

  void step() override {
    if (busy() == 0) return;
    int* flags = pm.active();
    float* coef = pm.step();
    for (int b = 0; b < S; ++b) {
      flags[b] = running[b];
      coef[b] = (running[b] && !first[b]) ? (float)(-opt.learning_rate) : 0.f;
    }
    pm.upload_active_step(s);
    arm_reductions(plan, flags, S);
    graph->launch();
    const double* red = wait_reductions(plan, flags, S);
    std::vector<int> done;
    bool update = false;
    for (int b = 0; b < S; ++b) {
      if (!running[b]) continue;
      first[b] = 0;
      CtmTileResult& r = res[b];
      const double pe = red[4 * b + 0];
      const double gn = std::sqrt(red[4 * b + 2]);
      r.history.push_back(pe);
      r.evaluations++;
      r.iterations = r.evaluations;
      r.final_grad_norm = gn;
      if (pe < min_pe[b]) {
        min_pe[b] = pe;
        copy_tile(best.as<float>() + b * n2, mask.as<float>() + b * n2, n2, s);
      }
      if (gn <= opt.tolerance) {             // CTM_Optimizer.cpp:156-159
        r.status = 0;
        done.push_back(b);
      } else if (r.evaluations == opt.iterations) {
        r.status = 1;
        done.push_back(b);
        update = true;
      }
    }
    if (done.empty()) return;
    // the reference applies the update after the last evaluation too (CTM_Optimizer.cpp:164-165)
    if (update) {
      for (int b = 0; b < S; ++b) {
        flags[b] = 0;
        coef[b] = 0.f;
      }
      for (int b : done)
        if (res[b].status == 1) {
          flags[b] = 1;
          coef[b] = (float)(-opt.learning_rate);
        }
      pm.upload_active_step(s);
      launch_axpy(g.as<float>(), x.as<float>(), pm.d_step(), pm.d_active(), n2, S, s);
    }
    for (int b : done) {
      res[b].final_pe = min_pe[b];
      copy_tile(job[b].out_mask, best.as<float>() + b * n2, n2, s);
      if (job[b].out_theta != nullptr) copy_tile(job[b].out_theta, x.as<float>() + b * n2, n2, s);
      finish(b);
    }
    spin_wait(s);   // the staged flags are rewritten next step
  }
};

// ------------------------------------------------------------------ L-BFGS
struct LbfgsPool final : CtmPool::Impl {
  // X, G: the point slots of every tile (iterates and their gradients,
  // [S][P][N^2], opt_kernels.hpp); d: the direction of every tile
  DevMem X, G, d, pair_part, dir_part, vec_red;
  // device-resident history of every slot: meta ints (pairs, base, trial,
  // pair points), S'Y and Y'Y (kMaxHist^2 doubles each) and the direction
  // coefficients
  DevMem d_meta, d_sy, d_yy, d_coef;
  HostPinned hdg;
  std::vector<LbfgsTile> st;
  // g . d of a new direction arrives with the evaluation of its first trial
  // (step 1): the line search starts once both are on the host
  std::vector<char> pending_dg0;
  int mb_pair = 0;
  std::unique_ptr<EvalGraph> graph;
  // events around the round's pair-kernel launch when the context times it
  hipEvent_t pair_ev[2] = {nullptr, nullptr};
  bool pair_timed = false;
  const LineSearchParams lp;
  const int m_max;
  const int P;   // point slots per tile

  LbfgsPool(Plan& p, int slots, const CtmOptions& o)
      : Impl(p, slots, o), hdg((size_t)slots * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent), st(slots), pending_dg0(slots, 0),
        m_max(o.history), P(lbfgs_points(o.history)) {
    alloc_slots(X, S * P, n2);
    alloc_slots(G, S * P, n2);
    alloc_slots(d, S, n2);
    pair_part.alloc((size_t)S * kOptBlocks * kPairDots * sizeof(double));
    dir_part.alloc((size_t)S * kOptBlocks * sizeof(double));
    LITHO_HIP(hipMemsetAsync(dir_part.ptr, 0, dir_part.bytes, s));   // reduced by every evaluation
    vec_red.alloc((size_t)S * kPairDots * sizeof(double));
    d_meta.alloc((size_t)S * kMetaStride * sizeof(int));
    d_sy.alloc((size_t)S * kMaxHist * kMaxHist * sizeof(double));
    d_yy.alloc((size_t)S * kMaxHist * kMaxHist * sizeof(double));
    d_coef.alloc((size_t)S * kCoefStride * sizeof(double));
    LITHO_HIP(hipMemsetAsync(d_meta.ptr, 0, d_meta.bytes, s));
    LITHO_HIP(hipMemsetAsync(d_sy.ptr, 0, d_sy.bytes, s));
    LITHO_HIP(hipMemsetAsync(d_yy.ptr, 0, d_yy.bytes, s));
    // every evaluation: theta = X[base] + step * d for the active slots into
    // X[trial], its gradient into G[trial], g . d for the line search; base
    // and trial are read per tile from meta when the kernels run (the masks
    // of the trials are not kept: the result is rebuilt from theta)
    EvalRequest rq = base_request(nullptr, G.as<float>());
    rq.base = X.as<float>();
    rq.dir = d.as<float>();
    rq.coef = pm.d_step();
    rq.theta_out = X.as<float>();
    rq.ga.theta = X.as<float>();
    rq.ga.dir = d.as<float>();
    rq.ring = d_meta.as<int>() + kMetaBase;   // [base, trial] at kMetaBase, kMetaBase + 1
    rq.ring_stride = kMetaStride;
    rq.ring_points = P;
    // the direction's g . d partials, reduced with the evaluation's losses
    // into mapped host memory (meaningful for the slots with a new direction)
    rq.extra_partials = dir_part.as<double>();
    rq.extra_blocks = kOptBlocks;
    rq.extra_out = hdg.dev_as<double>();
    graph = std::make_unique<EvalGraph>(plan, S, rq);
    LITHO_HIP(hipEventCreate(&pair_ev[0]));
    LITHO_HIP(hipEventCreate(&pair_ev[1]));
  }
  ~LbfgsPool() override {
    (void)hipEventDestroy(pair_ev[0]);
    (void)hipEventDestroy(pair_ev[1]);
  }

  bool admit(const CtmJob& j) override {
    const int b = find_free();
    if (b < 0) return false;
    load_slot(b, j, X.as<float>() + (size_t)b * P * n2);   // point 0, the base
    LITHO_HIP(hipMemsetAsync(d.as<float>() + b * n2, 0, n2 * sizeof(float), s));   // first trial: theta0 + 0 * d
    launch_lbfgs_reset(d_meta.as<int>(), b, s);   // no pairs, base point 0, trial point 1
    st[b] = LbfgsTile{};
    pending_dg0[b] = 0;
    return true;
  }

  // the base point must already hold the accepted iterate and its gradient
  // (the history update moves the trial there; the first line search swaps)
  void begin_line_search(int b, double dg0) {
    LbfgsTile& t = st[b];
    t.f0 = t.f;
    t.dg0 = dg0;
    t.test_decr = lp.ftol * dg0;
    t.test_curv = -lp.wolfe * dg0;
    t.ls_iter = 0;
    t.lo_a = 0.0;
    t.lo_f = t.f0;
    t.lo_g = dg0;
    t.hi_a = 0.0;
    t.hi_f = t.f0;
    t.a = t.step;
    if (dg0 > 0.0) {  // LBFGS++ throws: "the moving direction increases the objective function value"
      t.phase = LbfgsTile::DONE;
      t.final_is_xp = true;
      res[b].status = 2;
      return;
    }
    t.phase = LbfgsTile::BRACKET;
  }
  void fail(int b) {
    st[b].phase = LbfgsTile::DONE;
    st[b].final_is_xp = true;
    st[b].f = st[b].f0;
    res[b].status = 2;
  }
  void accept(int b) {
    LbfgsTile& t = st[b];
    if (t.gnorm <= opt.tolerance) {
      t.phase = LbfgsTile::DONE;
      res[b].status = 0;
    } else if (opt.iterations != 0 && t.k >= opt.iterations) {
      t.phase = LbfgsTile::DONE;
      res[b].status = 1;
    } else {
      t.phase = LbfgsTile::NEED_PAIR;
    }
  }

  // (A) new directions for the slots that accepted a step, formed on the
  // device after the round's upload: pair dots, history update and
  // compact-form coefficients, the direction and g . d, without a host round
  // trip.  g . d is copied to the host ahead of the evaluation of the first
  // trial (step 1), and the line search starts when both have arrived.
  bool prepare_directions() {
    int* need = pm.need();
    bool any = false;
    mb_pair = 0;
    for (int b = 0; b < S; ++b) {
      need[b] = running[b] && st[b].phase == LbfgsTile::NEED_PAIR ? 1 : 0;
      if (!need[b]) continue;
      any = true;
      LbfgsTile& t = st[b];
      mb_pair = std::max(mb_pair, std::min(t.pairs, m_max));
      t.pairs += 1;
      t.step = 1.0;
      t.k += 1;
      t.a = t.step;
      t.phase = LbfgsTile::BRACKET;
      pending_dg0[b] = 1;
    }
    return any;
  }
  // mb_pair bounds the history length before the update over the slots of
  // the round (it selects the instantiation); one more pair after it
  void launch_directions() {
    const int mb_dir = std::min(mb_pair + 1, m_max);
    // timed: launches over a full history (the k_lbfgs_pair<m> instantiation)
    pair_timed = plan.ctx->time_pair && mb_pair == m_max;
    if (pair_timed) LITHO_HIP(hipEventRecord(pair_ev[0], s));
    launch_lbfgs_pair(X.as<float>(), G.as<float>(), P, pm.d_need(), d_meta.as<int>(), n2, S, mb_pair,
                      pair_part.as<double>(), s);
    if (pair_timed) LITHO_HIP(hipEventRecord(pair_ev[1], s));
    launch_reduce_fixed(pair_part.as<double>(), kOptBlocks, 4 * (mb_pair + 1), vec_red.as<double>(), kPairDots, S,
                        s);
    launch_lbfgs_update(S, pm.d_need(), vec_red.as<double>(), mb_pair, m_max, d_meta.as<int>(), d_sy.as<double>(),
                        d_yy.as<double>(), d_coef.as<double>(), s);
    launch_lbfgs_direction(X.as<float>(), G.as<float>(), P, pm.d_need(), d_meta.as<int>(), d_coef.as<double>(), n2,
                           d.as<float>(), S, mb_dir, dir_part.as<double>(), s);
    // g . d is reduced into mapped host memory by the evaluation's final
    // reduction (off the path to the evaluation)
  }

  // (B) one batched evaluation of every slot with a pending trial
  void evaluate(bool directions) {
    int* flags = pm.active();
    float* stepv = pm.step();
    bool any = false;
    for (int b = 0; b < S; ++b) {
      const LbfgsTile& t = st[b];
      const bool ev = running[b] && (t.phase == LbfgsTile::INIT || t.phase == LbfgsTile::BRACKET ||
                                     t.phase == LbfgsTile::ZOOM || t.phase == LbfgsTile::REEVAL);
      flags[b] = ev ? 1 : 0;
      stepv[b] = (ev && t.phase != LbfgsTile::INIT) ? (float)t.a : 0.f;
      any = any || ev;
    }
    if (!any) return;   // a new direction always comes with its first trial
    pm.upload_active_step(s);
    if (directions) launch_directions();
    // the evaluation writes the losses and g . d of every evaluated slot
    // (g . d is used where a new direction is pending); waiting for all of
    // them keeps a late write out of the next round's armed entries
    arm_reductions(plan, flags, S);
    arm_values(hdg.as<double>(), flags, S);
    graph->launch();
    const double* red = wait_reductions(plan, flags, S);
    wait_values(hdg.as<double>(), flags, S, s);
    if (directions && pair_timed) {   // both events completed before the evaluation did
      float ms = 0.f;
      LITHO_HIP(hipEventElapsedTime(&ms, pair_ev[0], pair_ev[1]));
      plan.ctx->pair_launches.fetch_add(1);
      plan.ctx->add_pair_ms(ms);
    }
    std::vector<int> started;
    for (int b = 0; b < S; ++b) {
      if (!flags[b]) continue;
      LbfgsTile& t = st[b];
      if (pending_dg0[b]) {
        // the direction's g . d (copied before this evaluation on the same
        // stream); a non-descent direction ends the tile without counting
        // the trial, as if it had not been evaluated
        pending_dg0[b] = 0;
        begin_line_search(b, hdg.as<double>()[b]);
        if (t.phase == LbfgsTile::DONE) continue;
      }
      const double f = red[4 * b + 0];
      const double gn2 = red[4 * b + 2];
      const double dg = red[4 * b + 3];
      if (t.phase == LbfgsTile::REEVAL) {
        // the same theta as the earlier trial at lo_a, so the same f; the
        // reference keeps that point's gradient instead of evaluating again,
        // so this evaluation is neither counted nor recorded
        t.f = f;
        t.gnorm = std::sqrt(gn2);
        accept(b);
        continue;
      }
      res[b].history.push_back(f);
      res[b].evaluations++;
      if (t.phase == LbfgsTile::INIT) {
        t.f = f;
        t.gnorm = std::sqrt(gn2);
        t.k = 1;
        if (t.gnorm <= opt.tolerance) {   // LBFGS++ early exit, 1 evaluation
          t.phase = LbfgsTile::DONE;
          res[b].status = 0;
          continue;
        }
        t.step = 1.0 / t.gnorm;
        started.push_back(b);             // d = -g, then the first line search
        continue;
      }
      const double a = t.a;
      if (t.phase == LbfgsTile::BRACKET) {
        // Alg. 3.5: leave for the zoom on a failed decrease test (or no
        // decrease against the last trial), stop at strong Wolfe, else expand
        if (f - t.f0 > a * t.test_decr || (0.0 < t.lo_a && f >= t.lo_f)) {
          t.hi_a = a; t.hi_f = f;
          t.phase = LbfgsTile::ZOOM;
        } else if (std::abs(dg) <= t.test_curv) {
          t.f = f; t.gnorm = std::sqrt(gn2);
          accept(b);
          continue;
        } else {
          t.hi_a = t.lo_a; t.hi_f = t.lo_f;
          t.lo_a = a; t.lo_f = f; t.lo_g = dg;
          if (dg >= 0.0) {
            t.phase = LbfgsTile::ZOOM;
          } else {
            t.ls_iter += 1;
            if (t.ls_iter >= lp.max_linesearch) {   // trials used up: the best point is this one
              t.f = f; t.gnorm = std::sqrt(gn2);
              accept(b);
            } else {
              t.a = 2.0 * a;
            }
            continue;
          }
        }
      } else {  // ZOOM, Alg. 3.6
        if (f - t.f0 > a * t.test_decr || f >= t.lo_f) {
          if (a == t.hi_a) { fail(b); continue; }   // LBFGS++ throws: insufficient numeric precision
          t.hi_a = a; t.hi_f = f;
        } else {
          if (std::abs(dg) <= t.test_curv) {
            t.f = f; t.gnorm = std::sqrt(gn2);
            accept(b);
            continue;
          }
          if (dg * (t.hi_a - t.lo_a) >= 0.0) {
            t.hi_a = t.lo_a; t.hi_f = t.lo_f;
          }
          if (a == t.lo_a) { fail(b); continue; }   // LBFGS++ throws: insufficient numeric precision
          t.lo_a = a; t.lo_f = f; t.lo_g = dg;
        }
        t.ls_iter += 1;
        if (t.ls_iter >= lp.max_linesearch) {
          // trials used up: the best point of sufficient decrease (lo), or
          // LBFGS++ throws when there is none
          if (t.lo_a <= 0.0) { fail(b); continue; }
          if (t.lo_a == a) {
            t.f = f; t.gnorm = std::sqrt(gn2);
            accept(b);
          } else {
            t.phase = LbfgsTile::REEVAL;
            t.a = t.lo_a;
          }
          continue;
        }
      }
      t.a = ls_quad_interp(t.lo_a, t.hi_a, t.lo_f, t.hi_f, t.lo_g);
    }
    if (started.empty()) return;
    // slots whose first evaluation just ended: its point becomes the base,
    // d = -g and the first line search
    for (int b = 0; b < S; ++b) {
      flags[b] = 0;
      stepv[b] = 0.f;
    }
    for (int b : started) flags[b] = 1;
    pm.upload_active_step(s);
    launch_lbfgs_start(G.as<float>(), P, pm.d_active(), d_meta.as<int>(), n2, S, d.as<float>(), s);
    for (int b : started) begin_line_search(b, -st[b].gnorm * st[b].gnorm);
    spin_wait(s);   // the staged flags / steps are rewritten next round
  }

  void step() override {
    if (busy() == 0) return;
    evaluate(prepare_directions());
    // write out the slots that finished: the last trial, or the base after a
    // failed line search
    for (int b = 0; b < S; ++b) {
      if (!running[b] || st[b].phase != LbfgsTile::DONE) continue;
      launch_lbfgs_point_out(X.as<float>(), P, d_meta.as<int>(), b, st[b].final_is_xp ? kMetaBase : kMetaTrial, n2,
                             job[b].out_theta, job[b].out_mask, s);
      res[b].iterations = std::max(st[b].k, 1);
      res[b].final_pe = st[b].f;
      res[b].final_grad_norm = st[b].gnorm;
      finish(b);
    }
  }
};

}  // namespace

// ------------------------------------------------------------------ CtmPool
CtmPool::CtmPool(Plan& plan, int slots, const CtmOptions& opt) {
  validate_ctm_options(opt);
  if (slots <= 0 || slots > plan.capacity) throw Error(1, "CTM pool: slots must be in [1, plan capacity]");
  if (opt.method == 0)
    impl_ = std::make_unique<GdPool>(plan, slots, opt);
  else
    impl_ = std::make_unique<LbfgsPool>(plan, slots, opt);
}
CtmPool::~CtmPool() = default;
int CtmPool::slots() const { return impl_->S; }
int CtmPool::busy() const { return impl_->busy(); }
bool CtmPool::admit(const CtmJob& job) { return impl_->admit(job); }
void CtmPool::step() { impl_->step(); }
std::vector<std::pair<long long, CtmTileResult>> CtmPool::take_finished() {
  std::vector<std::pair<long long, CtmTileResult>> out;
  out.swap(impl_->finished);
  return out;
}
size_t CtmPool::slot_bytes(const Plan& plan, const CtmOptions& opt) {
  const size_t n2 = plan.tile_elems();
  return (opt.method == 0 ? 5 * n2 : (2 * (size_t)lbfgs_points(opt.history) + 2) * n2) * sizeof(float);
}

std::vector<CtmTileResult> run_ctm(Plan& plan, int batch, const float* targets, const float* theta0,
                                   const CtmOptions& opt, float* out_masks, float* out_theta) {
  validate_ctm_options(opt);
  if (batch <= 0 || batch > plan.capacity) throw Error(1, "batch exceeds the plan capacity");
  if (targets == nullptr || out_masks == nullptr) throw Error(1, "CTM: target and output mask are required");
  const size_t n2 = plan.tile_elems();
  CtmPool pool(plan, batch, opt);
  for (int b = 0; b < batch; ++b) {
    CtmJob j;
    j.id = b;
    j.target = targets + b * n2;
    j.theta0 = theta0 != nullptr ? theta0 + b * n2 : nullptr;
    j.out_mask = out_masks + b * n2;
    j.out_theta = out_theta != nullptr ? out_theta + b * n2 : nullptr;
    pool.admit(j);
  }
  std::vector<CtmTileResult> res(batch);
  while (pool.busy() > 0) {
    pool.step();
    for (auto& f : pool.take_finished()) res[(size_t)f.first] = std::move(f.second);
  }
  LITHO_HIP(hipStreamSynchronize(plan.ctx->stream));
  return res;
}

}  // namespace litho
