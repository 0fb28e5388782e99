// Served CTM with continuous batching: requests (one tile each, any of the
// plans' sizes) arrive at any time; worker threads keep one optimizer pool per
// tile size, admit queued requests into free slots at every round (in arrival
// order, within a device-memory budget shared by all sizes) and release each
// request as soon as its own optimizer finishes.  Plans that share a context
// are served by one worker on its stream; each further context gets its own
// worker, so the rounds of different sizes overlap on the GPU.  A request's
// result does not depend on what it shares its batches with.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ctm_driver.hpp"

namespace litho {

struct ServerOptions {
  CtmOptions ctm;
  double memory_budget = 64.0 * (1ull << 30);   // bytes of request state in flight
};

struct ServedResult {
  CtmTileResult result;
  double queued_ms = 0.0;    // submit -> admission
  double service_ms = 0.0;   // admission -> completion
  int error = 0;             // status code when the worker failed
};

class Server {
 public:
  Server(const std::vector<Plan*>& plans, const ServerOptions& opt);
  ~Server();                 // finishes the admitted and queued requests, then stops
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;
  long long submit(int plan_index, const float* target, const float* theta0, float* out_mask, float* out_theta);
  // Waits up to timeout_s (negative: forever) for the ticket; true when done
  // (the result is removed from the server and moved to *out).
  bool wait(long long ticket, double timeout_s, ServedResult* out);
  size_t request_bytes(int plan_index) const;

 private:
  using Clock = std::chrono::steady_clock;
  struct Pending {
    long long ticket;
    int plan;
    CtmJob job;
    Clock::time_point submitted;
  };
  struct Running {
    int plan;
    Clock::time_point submitted, admitted;
  };
  void work(Context* ctx);

  std::vector<Plan*> plans_;
  ServerOptions opt_;
  std::vector<std::unique_ptr<CtmPool>> pools_;
  std::vector<size_t> bytes_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Pending> queue_;
  std::map<long long, Running> running_;
  std::map<long long, ServedResult> done_;
  double in_flight_ = 0.0;
  long long next_ticket_ = 1;
  bool stop_ = false;
  int error_ = 0;
  std::string error_msg_;
  std::vector<std::thread> workers_;

 public:
  const std::string& error_message() const { return error_msg_; }
};

}  // namespace litho
