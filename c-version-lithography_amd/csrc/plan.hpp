// Host-side context / plan objects behind the C ABI and the device evaluation
// helper shared by the single-tile optimizers, the batch driver and the server.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "opt_kernels.hpp"
#include "socs_kernels.hpp"

namespace litho {

// Exceptions carry the C status code the ABI returns.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LITHO_HIP(call)                                                                         \
  do {                                                                                          \
    hipError_t _e = (call);                                                                     \
    if (_e != hipSuccess)                                                                       \
      throw ::litho::Error(3, std::string(#call) + ": " + hipGetErrorString(_e));               \
  } while (0)

// Owning device allocation.
struct DevMem {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  explicit DevMem(size_t b) { alloc(b); }
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  DevMem(DevMem&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DevMem& operator=(DevMem&& o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevMem() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    if (hipMalloc(&ptr, b) != hipSuccess) {
      ptr = nullptr;
      throw Error(2, "hipMalloc of " + std::to_string(b) + " bytes failed");
    }
    bytes = b;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Pinned host staging.  With mapped | coherent flags the kernels can write it
// directly (dev is the device-side address): the writes are visible to the
// host once the writing kernel has completed.
struct HostPinned {
  void* ptr = nullptr;
  void* dev = nullptr;
  explicit HostPinned(size_t b, unsigned flags = hipHostMallocDefault) {
    if (hipHostMalloc(&ptr, b, flags) != hipSuccess) throw Error(2, "hipHostMalloc failed");
    if ((flags & hipHostMallocMapped) && hipHostGetDevicePointer(&dev, ptr, 0) != hipSuccess) {
      (void)hipHostFree(ptr);
      ptr = nullptr;
      throw Error(2, "hipHostGetDevicePointer failed");
    }
  }
  HostPinned(const HostPinned&) = delete;
  HostPinned& operator=(const HostPinned&) = delete;
  ~HostPinned() { if (ptr) (void)hipHostFree(ptr); }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
  template <class T>
  T* dev_as() const { return static_cast<T*>(dev); }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool graphs = true;    // replay repeated evaluations from a captured hipGraph
  // evaluation graphs built on this context: captured, failed to capture
  // (launched directly instead) and direct by choice (graphs off)
  std::atomic<long long> graphs_captured{0}, graphs_failed{0}, graphs_direct{0};
  std::atomic<int> live_plans{0};   // plans made on this context and not yet destroyed
  // HIP-event timing of the L-BFGS pair-kernel launches of this context's
  // optimizer rounds (off by default)
  bool time_pair = false;
  std::atomic<long long> pair_launches{0};
  std::atomic<double> pair_ms{0.0};
  void add_pair_ms(double v) {
    double cur = pair_ms.load();
    while (!pair_ms.compare_exchange_weak(cur, cur + v)) {
    }
  }
  DevMem scratch;        // grow-only workspace of the level-set and tiling calls
  // At least `bytes` of workspace.  Growing it waits for the stream first, so
  // work still queued on the old buffer finishes before it is freed.
  void* workspace(size_t bytes) {
    if (scratch.bytes < bytes) {
      if (stream != nullptr) LITHO_HIP(hipStreamSynchronize(stream));
      scratch.alloc(bytes);
    }
    return scratch.ptr;
  }
};

// Device state of one tile size: scaled kernels, tables and the workspace of
// up to `capacity` tiles.
struct Plan {
  Context* ctx = nullptr;
  int n = 0, k = 0, wr = 0, wc = 0, win_y0 = 0, win_x0 = 0, capacity = 0;
  double weight_sum = 0.0;
  const SocsOps* ops = nullptr;
  int col_blocks = 0;
  int img_blocks = 0;   // image-kernel blocks (FP64 partials of PE and penalty)
  DevMem h, tw, rtw, row_freq, col_freq, row_of_u, col_of_u;
  int pruned = 0;                            // window-pruned column kernels in use
  int row_pruned = 0;                        // window-pruned row kernels in use
  int corr_len = 0;                          // L of the band-limited column path, 0 = fused column kernel
  DevMem twl, acorr, ghat, spart;
  DevMem twi;                                // pass twiddles of the length-n register transforms (powers of two)
  int adj_mfma = 0;                          // band-limited adjoint on the matrix cores (Toeplitz GEMM)
  int generic = 0;                           // direct-DFT pipeline (no compiled transforms for n)
  int tuned_chunk = 0;                       // batch size chosen by autotune_chunk (0 = not tuned)
  int tuned_limit = 0;                       // the limit it was tuned for
  DevMem rt, mhat, t, v, y, p;
  DevMem img_partials, grad_partials, red;   // red: [capacity][4] = pe, bg, |g|^2, g.d
  DevMem coef;                               // [capacity] float, K1 step coefficients
  HostPinned* host_red = nullptr;            // [capacity][4]
  HostPinned* host_coef = nullptr;           // [capacity] float
  size_t device_bytes = 0;

  ~Plan() {
    delete host_red;
    delete host_coef;
  }
  size_t tile_elems() const { return (size_t)n * n; }
  SocsArgs args() const {
    SocsArgs a;
    a.k = k; a.wr = wr; a.wc = wc;
    a.row_freq = row_freq.as<int>(); a.col_freq = col_freq.as<int>();
    a.h = h.as<float2>(); a.tw = tw.as<float2>();
    a.rt = rt.as<float2>(); a.mhat = mhat.as<float2>(); a.t = t.as<float2>();
    a.v = v.as<float2>(); a.y = y.as<float2>(); a.p = p.as<float2>();
    a.active = nullptr;
    a.row_of_u = row_of_u.as<int>();
    a.pruned = pruned;
    a.col_of_u = col_of_u.as<int>();
    a.rtw = rtw.as<float2>();
    a.row_pruned = row_pruned;
    a.n = n;
    a.corr_len = corr_len;
    a.twl = twl.as<float2>();
    a.acorr = acorr.as<float2>();
    a.ghat = ghat.as<float2>();
    a.spart = spart.as<float>();
    a.twi = twi.bytes > 0 ? twi.as<float2>() : nullptr;
    a.adj_mfma = adj_mfma;
    return a;
  }
};

// Plan construction (validates like imaging.cpp:68-107).
Plan* make_plan(Context* ctx, int n, int k, int wr, int wc, int y0, int x0, const double* kernels,
                const double* weights, double weight_sum, int capacity);

// One batched evaluation.  K1 input: mode 0 = mask, 1 = theta (+ coef*dir);
// need_adjoint adds the adjoint tail.  Reductions land in plan.red; the call
// does not synchronise.
struct EvalRequest {
  int mode = 1;
  const float* base = nullptr;
  const float* dir = nullptr;
  const float* coef = nullptr;       // device [batch]
  float* theta_out = nullptr;
  float* mask_out = nullptr;
  ImageArgs im{};
  GradArgs ga{};
  bool adjoint = true;
  const int* active = nullptr;       // device [batch] flags, null = all tiles
  double* red_out = nullptr;         // device-visible [batch][4] reductions (PE, penalty, |g|^2, g.d); null = plan.red
  // L-BFGS point slots (SocsArgs::ring): base / theta_out / ga.theta /
  // ga.grad_theta are then the point arrays, indexed per tile by ring
  const int* ring = nullptr;
  int ring_stride = 0, ring_points = 0;
  // optional: partials of one more value per tile ([batch][extra_blocks]),
  // reduced with the evaluation's own into extra_out[b] (the L-BFGS
  // direction's g . d, which the host reads with the evaluation's results)
  const double* extra_partials = nullptr;
  int extra_blocks = 0;
  double* extra_out = nullptr;
};

void run_eval(Plan& plan, int batch, const EvalRequest& rq);

// The launch sequence of run_eval(plan, batch, rq) captured once and replayed:
// repeated evaluations on fixed buffers (the optimizer loops) pay one graph
// launch instead of ~10 kernel launches.  The captured sequence has kernel
// nodes only (no memset or copy nodes).  When the context disables graphs, or
// capture fails, launch() issues the same kernels directly.
class EvalGraph {
 public:
  EvalGraph(Plan& plan, int batch, const EvalRequest& rq);
  ~EvalGraph();
  EvalGraph(const EvalGraph&) = delete;
  EvalGraph& operator=(const EvalGraph&) = delete;
  void launch();
  bool captured() const { return exec_ != nullptr; }

 private:
  void capture();
  Plan& plan_;
  int batch_;
  EvalRequest rq_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
};

// Average HIP-event time (ms) of the three launch groups of one evaluation on
// the plan's stream: [0] forward half (3 kernels), [1] fused column
// image/adjoint kernel, [2] adjoint tail (3 kernels), [3] whole evaluation.
void profile_eval(Plan& plan, int batch, const EvalRequest& rq, int repeats, double* stage_ms);

// Copy plan.red of `batch` tiles to the pinned host mirror and wait.
const double* fetch_reductions(Plan& plan, int batch);
// Waits for every command on the stream by polling it.
void spin_wait(hipStream_t s);
// Evaluations whose request writes its reductions to plan.host_red (red_out =
// its device address): arm_reductions marks the entries of the tiles with
// flags[b] != 0 unwritten before the evaluation is launched, and
// wait_reductions spins on those host-memory values until the device has
// written them all (no stream call on the way: a stream poll answers
// ~20 us later than the values arrive).  The evaluation writes only its
// active tiles' entries.
void arm_reductions(Plan& plan, const int* flags, int batch);
// The same for other device-written values in mapped host memory: count
// doubles, those with flags[i] != 0 (all with flags null).
void arm_values(double* host_values, const int* flags, int count);
void wait_values(const double* host_values, const int* flags, int count, hipStream_t s);
const double* wait_reductions(Plan& plan, const int* flags, int batch);

}  // namespace litho
