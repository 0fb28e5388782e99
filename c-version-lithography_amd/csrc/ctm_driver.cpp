// CTM gradient descent and L-BFGS with device-resident state.
//
// Gradient descent restates CTM_Optimizer::_optimize_gradient_descent
// (CTM_Optimizer.cpp:109-173).  L-BFGS restates the published algorithm of
// LBFGS++ (LBFGSSolver::minimize as called at CTM_Optimizer.cpp:194-212): the
// inverse-Hessian product uses the compact representation of Byrd, Nocedal &
// Schnabel (1994), which equals LBFGS++'s two-loop recursion with
// H0 = (s'y / y'y) I; pairs are kept only when s'y > eps * y'y; the first step
// is 1/||d||, later steps 1; the strong-Wolfe line search follows Nocedal &
// Wright Alg. 3.5 / 3.6 with a safeguarded cubic zoom.  Every objective
// evaluation (line-search trials included) is one grad-eval and appends one PE
// to the history, like the reference's _error_history (CTM_Optimizer.cpp:204).
#include "ctm_driver.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace litho {

double zoom_trial(double a_lo, double f_lo, double g_lo, double a_hi, double f_hi, double g_hi) {
  const double lo = std::min(a_lo, a_hi), hi = std::max(a_lo, a_hi);
  const double width = hi - lo;
  double a = std::numeric_limits<double>::quiet_NaN();
  const double d1 = g_lo + g_hi - 3.0 * (f_lo - f_hi) / (a_lo - a_hi);
  const double rad = d1 * d1 - g_lo * g_hi;
  if (rad >= 0.0) {
    const double d2 = std::copysign(std::sqrt(rad), a_hi - a_lo);
    const double den = g_hi - g_lo + 2.0 * d2;
    if (den != 0.0) a = a_hi - (a_hi - a_lo) * (g_hi + d2 - d1) / den;
  }
  if (!std::isfinite(a) || a < lo + 0.1 * width || a > hi - 0.1 * width) a = 0.5 * (lo + hi);
  return a;
}

namespace {

struct Buffers {
  DevMem x, xp, g, gp, d, mask, hs, hy, pair_part, dir_part, vec_red;
};

// Host-written parameters of a batch (active flags, step coefficients, L-BFGS
// slot lists and direction coefficients), staged in one pinned block and
// uploaded region by region with one transfer each.  The host rewrites a
// region only after a stream synchronisation that follows its last upload.
struct Params {
  size_t off_step, off_meta, off_coef, bytes;
  HostPinned host;
  DevMem dev;
  static size_t layout(int cap, size_t* step, size_t* meta, size_t* coef) {
    *step = (size_t)cap * sizeof(int);
    *meta = *step + (size_t)cap * sizeof(float);
    *coef = (*meta + (size_t)cap * kMetaStride * sizeof(int) + 7) / 8 * 8;
    return *coef + (size_t)cap * kCoefStride * sizeof(double);
  }
  explicit Params(int cap)
      : bytes(layout(cap, &off_step, &off_meta, &off_coef)), host(bytes), dev(bytes) {}
  int* active() const { return host.as<int>(); }
  float* step() const { return reinterpret_cast<float*>(host.as<char>() + off_step); }
  int* meta() const { return reinterpret_cast<int*>(host.as<char>() + off_meta); }
  double* coef() const { return reinterpret_cast<double*>(host.as<char>() + off_coef); }
  const int* d_active() const { return dev.as<int>(); }
  const float* d_step() const { return reinterpret_cast<const float*>(dev.as<char>() + off_step); }
  const int* d_meta() const { return reinterpret_cast<const int*>(dev.as<char>() + off_meta); }
  const double* d_coef() const { return reinterpret_cast<const double*>(dev.as<char>() + off_coef); }
  void upload(size_t off, size_t end, hipStream_t s) const {
    LITHO_HIP(hipMemcpyAsync(dev.as<char>() + off, host.as<char>() + off, end - off, hipMemcpyHostToDevice, s));
  }
  void upload_active_step(hipStream_t s) const { upload(0, off_meta, s); }
  void upload_meta(hipStream_t s) const { upload(off_meta, off_coef, s); }
  void upload_meta_coef(hipStream_t s) const { upload(off_meta, bytes, s); }
};

void copy_tile(float* dst, const float* src, size_t n, hipStream_t s) {
  LITHO_HIP(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s));
}

// ------------------------------------------------------------ gradient descent
std::vector<CtmTileResult> run_gd(Plan& plan, int batch, const float* targets, const float* theta0,
                                  const CtmOptions& opt, float* out_masks, float* out_theta) {
  hipStream_t s = plan.ctx->stream;
  const size_t n2 = plan.tile_elems();
  Buffers bf;
  bf.x.alloc(batch * n2 * sizeof(float));
  bf.g.alloc(batch * n2 * sizeof(float));
  bf.mask.alloc(batch * n2 * sizeof(float));
  Params pm(batch);
  float* x = bf.x.as<float>();
  if (theta0 != nullptr)
    LITHO_HIP(hipMemcpyAsync(x, theta0, batch * n2 * sizeof(float), hipMemcpyDeviceToDevice, s));
  else
    launch_init_theta(targets, x, batch * n2, s);
  float* best = out_masks;

  std::vector<CtmTileResult> res(batch);
  std::vector<int> running(batch, 1);
  std::vector<double> min_pe(batch, std::numeric_limits<double>::max());
  int* flags = pm.active();
  float* coef = pm.step();
  for (int b = 0; b < batch; ++b) {
    flags[b] = 1;
    coef[b] = 0.f;
  }

  for (int it = 0; it < opt.iterations; ++it) {
    EvalRequest rq;
    rq.mode = 1;
    rq.base = x;
    if (it > 0) {
      for (int b = 0; b < batch; ++b) {
        coef[b] = running[b] ? (float)(-opt.learning_rate) : 0.f;
        flags[b] = running[b];
      }
      rq.dir = bf.g.as<float>();
      rq.coef = pm.d_step();
      rq.theta_out = x;
    }
    pm.upload_active_step(s);
    rq.mask_out = bf.mask.as<float>();
    rq.im.thr = (float)opt.threshold;
    rq.im.alpha = (float)opt.alpha;
    rq.im.target = targets;
    rq.im.pe_weight = 1.f;
    rq.ga.scale = (float)(4.0 * opt.alpha);
    rq.ga.theta = x;
    rq.ga.grad_theta = bf.g.as<float>();
    rq.active = pm.d_active();
    run_eval(plan, batch, rq);
    const double* red = fetch_reductions(plan, batch);
    bool any = false;
    for (int b = 0; b < batch; ++b) {
      if (!running[b]) continue;
      const double pe = red[4 * b + 0];
      const double gn = std::sqrt(red[4 * b + 2]);
      res[b].history.push_back(pe);
      res[b].evaluations++;
      res[b].iterations = it + 1;
      res[b].final_grad_norm = gn;
      if (pe < min_pe[b]) {
        min_pe[b] = pe;
        copy_tile(best + b * n2, bf.mask.as<float>() + b * n2, n2, s);
      }
      if (gn <= opt.tolerance) {       // CTM_Optimizer.cpp:156-159
        running[b] = 0;
        res[b].status = 0;
      } else {
        any = true;
      }
    }
    if (!any) break;
  }
  // the reference applies the update after the last evaluation too (CTM_Optimizer.cpp:164-165)
  for (int b = 0; b < batch; ++b) coef[b] = running[b] ? (float)(-opt.learning_rate) : 0.f;
  pm.upload_active_step(s);
  launch_axpy(bf.g.as<float>(), x, pm.d_step(), nullptr, n2, batch, s);
  for (int b = 0; b < batch; ++b) res[b].final_pe = min_pe[b];
  if (out_theta != nullptr)
    LITHO_HIP(hipMemcpyAsync(out_theta, x, batch * n2 * sizeof(float), hipMemcpyDeviceToDevice, s));
  LITHO_HIP(hipStreamSynchronize(s));
  return res;
}

// ------------------------------------------------------------------ L-BFGS
struct LbfgsTile {
  enum Phase { INIT, BRACKET, ZOOM, NEED_PAIR, DONE };
  Phase phase = INIT;
  int k = 0, trials = 0;
  double f = 0, gnorm = 0, f0 = 0, dg0 = 0, step = 0, a = 0;
  double a_prev = 0, f_prev = 0, dg_prev = 0;
  double lo_a = 0, lo_f = 0, lo_g = 0, hi_a = 0, hi_f = 0, hi_g = 0;
  bool final_is_xp = false;
  // history: chronological slot list, dot matrices (row/col = chronological)
  std::vector<int> slots;
  std::vector<std::vector<double>> sy, yy;   // sy[i][j] = s_i . y_j, yy symmetric
  int spare = 0;
};

void solve_compact(const LbfgsTile& t, const std::vector<double>& p, const std::vector<double>& q, double* coef) {
  // d = -gamma g - S u + gamma Y w,  w = R^{-1} p,  u = R^{-T}((D + gamma Y'Y) w - gamma q)
  const int m = (int)t.slots.size();
  for (int i = 0; i < kCoefStride; ++i) coef[i] = 0.0;
  if (m == 0) {
    coef[0] = -1.0;
    return;
  }
  const double gamma = t.sy[m - 1][m - 1] / t.yy[m - 1][m - 1];
  std::vector<double> w(m), r(m), u(m);
  for (int i = m - 1; i >= 0; --i) {
    double acc = p[i];
    for (int j = i + 1; j < m; ++j) acc -= t.sy[i][j] * w[j];
    w[i] = acc / t.sy[i][i];
  }
  for (int i = 0; i < m; ++i) {
    double acc = t.sy[i][i] * w[i] - gamma * q[i];
    for (int j = 0; j < m; ++j) acc += gamma * t.yy[i][j] * w[j];
    r[i] = acc;
  }
  for (int i = 0; i < m; ++i) {
    double acc = r[i];
    for (int j = 0; j < i; ++j) acc -= t.sy[j][i] * u[j];
    u[i] = acc / t.sy[i][i];
  }
  coef[0] = -gamma;
  for (int j = 0; j < m; ++j) {
    coef[1 + j] = -u[j];
    coef[1 + kMaxHist + j] = gamma * w[j];
  }
}

std::vector<CtmTileResult> run_lbfgs(Plan& plan, int batch, const float* targets, const float* theta0,
                                     const CtmOptions& opt, float* out_masks, float* out_theta) {
  if (opt.history <= 0 || opt.history > kMaxHist)
    throw Error(1, "CTM lbfgs_history_size must be in [1, " + std::to_string(kMaxHist) + "]");
  hipStream_t s = plan.ctx->stream;
  const size_t n2 = plan.tile_elems();
  const LineSearchParams lp;
  const int m_max = opt.history;
  Buffers bf;
  bf.x.alloc(batch * n2 * sizeof(float));
  bf.xp.alloc(batch * n2 * sizeof(float));
  bf.g.alloc(batch * n2 * sizeof(float));
  bf.gp.alloc(batch * n2 * sizeof(float));
  bf.d.alloc(batch * n2 * sizeof(float));
  bf.mask.alloc(batch * n2 * sizeof(float));
  bf.hs.alloc(batch * (size_t)(kMaxHist + 1) * n2 * sizeof(float));
  bf.hy.alloc(batch * (size_t)(kMaxHist + 1) * n2 * sizeof(float));
  bf.pair_part.alloc(batch * (size_t)kOptBlocks * kPairDots * sizeof(double));
  bf.dir_part.alloc(batch * (size_t)kOptBlocks * sizeof(double));
  bf.vec_red.alloc(batch * kPairDots * sizeof(double));
  Params pm(batch);
  HostPinned hvec(batch * kPairDots * sizeof(double));
  float* x = bf.x.as<float>();
  float* xp = bf.xp.as<float>();
  float* g = bf.g.as<float>();
  float* gp = bf.gp.as<float>();
  float* d = bf.d.as<float>();
  LITHO_HIP(hipMemsetAsync(d, 0, batch * n2 * sizeof(float), s));
  if (theta0 != nullptr)
    LITHO_HIP(hipMemcpyAsync(xp, theta0, batch * n2 * sizeof(float), hipMemcpyDeviceToDevice, s));
  else
    launch_init_theta(targets, xp, batch * n2, s);

  std::vector<LbfgsTile> st(batch);
  std::vector<CtmTileResult> res(batch);
  for (int b = 0; b < batch; ++b) st[b].spare = 0;
  float* stepv = pm.step();
  int* flags = pm.active();
  int* meta = pm.meta();
  double* coefs = pm.coef();
  for (int b = 0; b < batch; ++b) {
    stepv[b] = 0.f;
    flags[b] = 1;
  }
  for (int i = 0; i < batch * kMetaStride; ++i) meta[i] = -1;
  for (int i = 0; i < batch * kCoefStride; ++i) coefs[i] = 0.0;
  const double eps = std::numeric_limits<double>::epsilon();

  // xp / gp must already hold the accepted point and its gradient (the pair
  // kernel shifts them; the first line search copies them)
  auto begin_line_search = [&](int b, double dg0) {
    LbfgsTile& t = st[b];
    t.f0 = t.f;
    t.dg0 = dg0;
    t.trials = 0;
    t.a_prev = 0.0;
    t.f_prev = t.f0;
    t.dg_prev = dg0;
    t.a = t.step;
    if (dg0 >= 0.0) {  // LBFGS++: "the moving direction does not decrease the objective"
      t.phase = LbfgsTile::DONE;
      t.final_is_xp = true;
      res[b].status = 2;
      return;
    }
    if (t.a < lp.min_step) {
      t.phase = LbfgsTile::DONE;
      t.final_is_xp = true;
      res[b].status = 2;
      return;
    }
    t.phase = LbfgsTile::BRACKET;
  };
  auto fail = [&](int b) {
    st[b].phase = LbfgsTile::DONE;
    st[b].final_is_xp = true;
    st[b].f = st[b].f0;
    res[b].status = 2;
  };
  auto accept = [&](int b) {
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
  };

  for (;;) {
    // ---- (A) new directions for tiles that accepted a step
    bool need_pair = false;
    int max_m = 0;
    for (int b = 0; b < batch; ++b) {
      int* mb = meta + b * kMetaStride;
      for (int i = 0; i < kMetaStride; ++i) mb[i] = 0;
      mb[0] = -1;
      if (st[b].phase == LbfgsTile::NEED_PAIR) {
        need_pair = true;
        const LbfgsTile& t = st[b];
        mb[0] = (int)t.slots.size();
        mb[1] = t.spare;
        for (size_t j = 0; j < t.slots.size(); ++j) mb[2 + j] = t.slots[j];
        max_m = std::max(max_m, mb[0]);
      }
    }
    if (need_pair) {
      const int pair_nv = 4 * (max_m + 1);   // dots per tile: 4 per old slot, then the new pair
      pm.upload_meta(s);
      launch_lbfgs_pair(x, xp, g, gp, bf.hs.as<float>(), bf.hy.as<float>(), pm.d_meta(), n2, batch, max_m, true,
                        bf.pair_part.as<double>(), s);
      launch_reduce_fixed(bf.pair_part.as<double>(), kOptBlocks, pair_nv, bf.vec_red.as<double>(), kPairDots, batch,
                          s);
      LITHO_HIP(hipMemcpyAsync(hvec.ptr, bf.vec_red.ptr, batch * kPairDots * sizeof(double), hipMemcpyDeviceToHost, s));
      LITHO_HIP(hipStreamSynchronize(s));
      const double* dots = hvec.as<double>();
      for (int b = 0; b < batch; ++b) {
        if (st[b].phase != LbfgsTile::NEED_PAIR) continue;
        LbfgsTile& t = st[b];
        const double* v = dots + (size_t)b * kPairDots;
        const int m = (int)t.slots.size();
        const double ys = v[4 * max_m + 0], yy = v[4 * max_m + 1];
        std::vector<double> p(m), q(m);
        for (int j = 0; j < m; ++j) {
          p[j] = v[4 * j + 2];
          q[j] = v[4 * j + 3];
        }
        if (ys > eps * yy) {          // BFGSMat::add_correction condition
          std::vector<double> col_sy(m), col_yy(m);
          for (int j = 0; j < m; ++j) {
            col_sy[j] = v[4 * j + 0];
            col_yy[j] = v[4 * j + 1];
          }
          int drop = m == m_max ? 1 : 0;
          int freed = drop ? t.slots[0] : -1;
          std::vector<int> ns(t.slots.begin() + drop, t.slots.end());
          ns.push_back(t.spare);
          const int m2 = (int)ns.size();
          std::vector<std::vector<double>> sy2(m2, std::vector<double>(m2, 0.0)), yy2 = sy2;
          for (int i = 0; i < m2 - 1; ++i)
            for (int j = 0; j < m2 - 1; ++j) {
              sy2[i][j] = t.sy[i + drop][j + drop];
              yy2[i][j] = t.yy[i + drop][j + drop];
            }
          for (int i = 0; i < m2 - 1; ++i) {
            sy2[i][m2 - 1] = col_sy[i + drop];
            yy2[i][m2 - 1] = col_yy[i + drop];
            yy2[m2 - 1][i] = col_yy[i + drop];
          }
          sy2[m2 - 1][m2 - 1] = ys;
          yy2[m2 - 1][m2 - 1] = yy;
          std::vector<double> p2(m2), q2(m2);
          for (int i = 0; i < m2 - 1; ++i) {
            p2[i] = p[i + drop];
            q2[i] = q[i + drop];
          }
          p2[m2 - 1] = v[4 * max_m + 2];
          q2[m2 - 1] = v[4 * max_m + 3];
          // spare slot for the next pair: the freed slot or an unused index
          if (freed >= 0) {
            t.spare = freed;
          } else {
            int cand = 0;
            while (std::find(ns.begin(), ns.end(), cand) != ns.end()) ++cand;
            t.spare = cand;
          }
          t.slots = ns;
          t.sy = sy2;
          t.yy = yy2;
          p = p2;
          q = q2;
        }
        solve_compact(t, p, q, coefs + (size_t)b * kCoefStride);
      }
      // rebuild meta with the committed slot lists for the direction kernel
      int dir_m = 0;
      for (int b = 0; b < batch; ++b) {
        int* mb = meta + b * kMetaStride;
        if (st[b].phase != LbfgsTile::NEED_PAIR) continue;
        mb[0] = (int)st[b].slots.size();
        mb[1] = st[b].spare;
        for (size_t j = 0; j < st[b].slots.size(); ++j) mb[2 + j] = st[b].slots[j];
        dir_m = std::max(dir_m, mb[0]);
      }
      pm.upload_meta_coef(s);
      launch_lbfgs_direction(g, bf.hs.as<float>(), bf.hy.as<float>(), pm.d_meta(), pm.d_coef(), n2, d, batch, dir_m,
                             bf.dir_part.as<double>(), s);
      launch_reduce_fixed(bf.dir_part.as<double>(), kOptBlocks, 1, bf.vec_red.as<double>(), 1, batch, s);
      LITHO_HIP(hipMemcpyAsync(hvec.ptr, bf.vec_red.ptr, batch * sizeof(double), hipMemcpyDeviceToHost, s));
      LITHO_HIP(hipStreamSynchronize(s));
      for (int b = 0; b < batch; ++b) {
        if (st[b].phase != LbfgsTile::NEED_PAIR) continue;
        st[b].step = 1.0;
        st[b].k += 1;
        begin_line_search(b, hvec.as<double>()[b]);
      }
    }
    // ---- (B) one batched evaluation for every tile with a pending trial
    bool any = false;
    for (int b = 0; b < batch; ++b) {
      const LbfgsTile& t = st[b];
      const bool ev = t.phase == LbfgsTile::INIT || t.phase == LbfgsTile::BRACKET || t.phase == LbfgsTile::ZOOM;
      flags[b] = ev ? 1 : 0;
      stepv[b] = t.phase == LbfgsTile::INIT ? 0.f : (float)t.a;
      any = any || ev;
    }
    if (!any) break;
    pm.upload_active_step(s);
    EvalRequest rq;
    rq.mode = 1;
    rq.base = xp;
    rq.dir = d;
    rq.coef = pm.d_step();
    rq.theta_out = x;
    rq.mask_out = bf.mask.as<float>();
    rq.im.thr = (float)opt.threshold;
    rq.im.alpha = (float)opt.alpha;
    rq.im.target = targets;
    rq.im.pe_weight = 1.f;
    rq.ga.scale = (float)(4.0 * opt.alpha);
    rq.ga.theta = x;
    rq.ga.grad_theta = g;
    rq.ga.dir = d;
    rq.active = pm.d_active();
    run_eval(plan, batch, rq);
    const double* red = fetch_reductions(plan, batch);
    for (int b = 0; b < batch; ++b) {
      if (!flags[b]) continue;
      LbfgsTile& t = st[b];
      const double f = red[4 * b + 0];
      const double gn2 = red[4 * b + 2];
      const double dg = red[4 * b + 3];
      res[b].history.push_back(f);
      res[b].evaluations++;
      if (t.phase == LbfgsTile::INIT) {
        t.f = f;
        t.gnorm = std::sqrt(gn2);
        if (t.gnorm <= opt.tolerance) {   // LBFGS++ early exit, 1 evaluation
          t.phase = LbfgsTile::DONE;
          t.k = 1;
          res[b].status = 0;
          continue;
        }
        stepv[b] = -1.f;                  // d = -g
        t.step = 1.0 / t.gnorm;
        t.k = 1;
        t.phase = LbfgsTile::NEED_PAIR;   // marker: handled right below
        continue;
      }
      t.trials += 1;
      const double a = t.a;
      if (t.phase == LbfgsTile::BRACKET) {
        if (f > t.f0 + lp.ftol * a * t.dg0 || (t.trials > 1 && f >= t.f_prev)) {
          t.lo_a = t.a_prev; t.lo_f = t.f_prev; t.lo_g = t.dg_prev;
          t.hi_a = a; t.hi_f = f; t.hi_g = dg;
          t.phase = LbfgsTile::ZOOM;
        } else if (std::abs(dg) <= -lp.wolfe * t.dg0) {
          t.f = f; t.gnorm = std::sqrt(gn2);
          accept(b);
          continue;
        } else if (dg >= 0.0) {
          t.lo_a = a; t.lo_f = f; t.lo_g = dg;
          t.hi_a = t.a_prev; t.hi_f = t.f_prev; t.hi_g = t.dg_prev;
          t.phase = LbfgsTile::ZOOM;
        } else {
          t.a_prev = a; t.f_prev = f; t.dg_prev = dg;
          t.a = std::min(2.0 * a, lp.max_step);
          if (t.trials >= lp.max_linesearch || t.a < lp.min_step) fail(b);
          continue;
        }
      } else {  // ZOOM
        if (f > t.f0 + lp.ftol * a * t.dg0 || f >= t.lo_f) {
          t.hi_a = a; t.hi_f = f; t.hi_g = dg;
        } else {
          if (std::abs(dg) <= -lp.wolfe * t.dg0) {
            t.f = f; t.gnorm = std::sqrt(gn2);
            accept(b);
            continue;
          }
          if (dg * (t.hi_a - t.lo_a) >= 0.0) {
            t.hi_a = t.lo_a; t.hi_f = t.lo_f; t.hi_g = t.lo_g;
          }
          t.lo_a = a; t.lo_f = f; t.lo_g = dg;
        }
      }
      // next zoom trial
      if (t.trials >= lp.max_linesearch) {
        fail(b);
        continue;
      }
      t.a = zoom_trial(t.lo_a, t.lo_f, t.lo_g, t.hi_a, t.hi_f, t.hi_g);
    }
    // tiles that finished INIT: d = -g and start the first line search
    bool init_dir = false;
    for (int b = 0; b < batch; ++b) {
      flags[b] = (st[b].phase == LbfgsTile::NEED_PAIR && st[b].k == 1 && st[b].slots.empty() &&
                  res[b].evaluations == 1) ? 1 : 0;
      init_dir = init_dir || flags[b];
    }
    if (init_dir) {
      pm.upload_active_step(s);
      launch_scale(g, d, pm.d_step(), pm.d_active(), n2, batch, s);
      for (int b = 0; b < batch; ++b) {
        if (!flags[b]) continue;
        copy_tile(xp + b * n2, x + b * n2, n2, s);
        copy_tile(gp + b * n2, g + b * n2, n2, s);
        begin_line_search(b, -st[b].gnorm * st[b].gnorm);
      }
      // the staged flags / steps are rewritten next iteration: wait for the upload
      LITHO_HIP(hipStreamSynchronize(s));
    }
  }
  // final theta: x, or xp after a failed line search
  for (int b = 0; b < batch; ++b) {
    const float* src = st[b].final_is_xp ? xp + b * n2 : x + b * n2;
    if (out_theta != nullptr) copy_tile(out_theta + b * n2, src, n2, s);
    launch_mask_from_theta(src, out_masks + b * n2, n2, s);
    res[b].iterations = std::max(st[b].k, 1);
    res[b].final_pe = st[b].f;
    res[b].final_grad_norm = st[b].gnorm;
  }
  LITHO_HIP(hipStreamSynchronize(s));
  return res;
}

}  // namespace

std::vector<CtmTileResult> run_ctm(Plan& plan, int batch, const float* targets, const float* theta0,
                                   const CtmOptions& opt, float* out_masks, float* out_theta) {
  if (opt.method != 0 && opt.method != 1) throw Error(1, "CTM optimizer must be gradient_descent or lbfgs");
  if (opt.iterations <= 0) throw Error(1, "CTM max_iteration must be positive");
  if (!std::isfinite(opt.learning_rate) || opt.learning_rate <= 0.0)
    throw Error(1, "CTM learning_rate must be finite and positive");
  if (opt.history <= 0) throw Error(1, "CTM lbfgs_history_size must be positive");
  if (!std::isfinite(opt.tolerance) || opt.tolerance < 0.0)
    throw Error(1, "CTM gradient_tolerance must be finite and non-negative");
  if (batch <= 0 || batch > plan.capacity) throw Error(1, "batch exceeds the plan capacity");
  if (targets == nullptr || out_masks == nullptr) throw Error(1, "CTM: target and output mask are required");
  return opt.method == 0 ? run_gd(plan, batch, targets, theta0, opt, out_masks, out_theta)
                         : run_lbfgs(plan, batch, targets, theta0, opt, out_masks, out_theta);
}

}  // namespace litho
