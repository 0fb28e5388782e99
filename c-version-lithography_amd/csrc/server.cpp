// Served CTM with continuous batching (server.hpp).
#include "server.hpp"

#include <algorithm>

namespace litho {

Server::Server(const std::vector<Plan*>& plans, const ServerOptions& opt) : plans_(plans), opt_(opt) {
  validate_ctm_options(opt.ctm);
  if (plans.empty()) throw Error(1, "server: no plans");
  for (const Plan* p : plans)
    if (p == nullptr) throw Error(1, "server: null plan");
  if (!(opt.memory_budget > 0.0)) throw Error(1, "server: memory budget must be positive");
  for (Plan* p : plans) {
    pools_.push_back(std::make_unique<CtmPool>(*p, p->capacity, opt.ctm));
    // optimizer state of one slot plus the plan's per-tile share of its workspace
    bytes_.push_back(CtmPool::slot_bytes(*p, opt.ctm) + p->device_bytes / (size_t)p->capacity);
  }
  for (size_t b : bytes_)
    if ((double)b > opt.memory_budget) throw Error(2, "server: one request exceeds the memory budget");
  std::vector<Context*> ctxs;
  for (Plan* p : plans)
    if (std::find(ctxs.begin(), ctxs.end(), p->ctx) == ctxs.end()) ctxs.push_back(p->ctx);
  for (Context* c : ctxs) workers_.emplace_back([this, c] { work(c); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto& w : workers_)
    if (w.joinable()) w.join();
}

size_t Server::request_bytes(int plan_index) const { return bytes_.at((size_t)plan_index); }

long long Server::submit(int plan_index, const float* target, const float* theta0, float* out_mask,
                         float* out_theta) {
  if (plan_index < 0 || plan_index >= (int)plans_.size()) throw Error(1, "server: plan index out of range");
  if (target == nullptr || out_mask == nullptr) throw Error(1, "server: target and output mask are required");
  long long t;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw Error(5, "server: stopped");
    if (error_ != 0) throw Error(error_, "server: worker failed: " + error_msg_);
    t = next_ticket_++;
    Pending p;
    p.ticket = t;
    p.plan = plan_index;
    p.job.id = t;
    p.job.target = target;
    p.job.theta0 = theta0;
    p.job.out_mask = out_mask;
    p.job.out_theta = out_theta;
    p.submitted = Clock::now();
    queue_.push_back(p);
  }
  cv_work_.notify_all();
  return t;
}

bool Server::wait(long long ticket, double timeout_s, ServedResult* out) {
  std::unique_lock<std::mutex> lk(mu_);
  auto ready = [&] { return done_.count(ticket) != 0; };
  if (timeout_s < 0.0) {
    cv_done_.wait(lk, ready);
  } else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    return false;
  }
  if (out != nullptr) *out = std::move(done_[ticket]);
  done_.erase(ticket);
  return true;
}

void Server::work(Context* ctx) {
  std::vector<int> mine;
  for (int i = 0; i < (int)plans_.size(); ++i)
    if (plans_[i]->ctx == ctx) mine.push_back(i);
  auto is_mine = [&](int plan) { return std::find(mine.begin(), mine.end(), plan) != mine.end(); };
  auto busy = [&] {
    for (int i : mine)
      if (pools_[(size_t)i]->busy() > 0) return true;
    return false;
  };
  auto queued = [&] {
    for (const auto& q : queue_)
      if (is_mine(q.plan)) return true;
    return false;
  };
  try {
    LITHO_HIP(hipSetDevice(ctx->device));
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] { return queued() || busy() || stop_ || error_ != 0; });
        if (error_ != 0) break;
        if (stop_ && !queued() && !busy()) break;
        // admission in arrival order: a request waits only for its own size's slots or the budget
        const auto now = Clock::now();
        for (auto it = queue_.begin(); it != queue_.end();) {
          const double b = (double)bytes_[(size_t)it->plan];
          if (is_mine(it->plan) && in_flight_ + b <= opt_.memory_budget && pools_[(size_t)it->plan]->admit(it->job)) {
            in_flight_ += b;
            running_[it->ticket] = Running{it->plan, it->submitted, now};
            it = queue_.erase(it);
          } else {
            ++it;
          }
        }
        // nothing admitted and nothing running: the budget is held by other sizes
        if (!busy()) {
          cv_work_.wait_for(lk, std::chrono::milliseconds(1));
          continue;
        }
      }
      // one round of every busy pool (outside the lock: submissions continue meanwhile)
      std::vector<std::pair<long long, CtmTileResult>> fin;
      for (int i : mine) {
        CtmPool& p = *pools_[(size_t)i];
        if (p.busy() == 0) continue;
        p.step();
        for (auto& f : p.take_finished()) fin.push_back(std::move(f));
      }
      if (fin.empty()) continue;
      LITHO_HIP(hipStreamSynchronize(ctx->stream));    // outputs of the finished requests are written
      const auto now = Clock::now();
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& f : fin) {
          const Running r = running_[f.first];
          running_.erase(f.first);
          in_flight_ -= (double)bytes_[(size_t)r.plan];
          ServedResult sr;
          sr.result = std::move(f.second);
          sr.queued_ms = std::chrono::duration<double, std::milli>(r.admitted - r.submitted).count();
          sr.service_ms = std::chrono::duration<double, std::milli>(now - r.admitted).count();
          done_[f.first] = std::move(sr);
        }
      }
      cv_done_.notify_all();
      cv_work_.notify_all();     // freed budget may let another size admit
    }
  } catch (const Error& e) {
    std::lock_guard<std::mutex> lk(mu_);
    error_ = e.code;
    error_msg_ = e.what();
  } catch (const std::exception& e) {
    std::lock_guard<std::mutex> lk(mu_);
    if (error_ == 0) {
      error_ = 5;
      error_msg_ = e.what();
    }
  }
  // after a failure, fail whatever is still pending so that no waiter blocks forever
  std::lock_guard<std::mutex> lk(mu_);
  if (error_ != 0) {
    for (auto& q : queue_) done_[q.ticket].error = error_;
    for (auto& r : running_) done_[r.first].error = error_;
    queue_.clear();
    running_.clear();
  }
  cv_done_.notify_all();
  cv_work_.notify_all();
}

}  // namespace litho
