// litho_bench: a C++ caller of the C ABI (include/litho_amd.h), used for
// profiling runs without Python.  It mirrors the reference's timing harness
// demo/demo_CTM_parallel_benchmark.cpp: set-up (optics, SOCS precompute, plan)
// outside the timer, then timed work on one tile.
//
//   litho_bench eval  N K REPEATS [KFILE]   one CTM grad-eval per repeat (stage timings)
//   litho_bench ctm   N K ITERS   KFILE [S] S L-BFGS CTM solves (default 3), each timed
//   litho_bench evalb N K REPEATS BATCH [KFILE]  the same for BATCH equal tiles in one launch
//   litho_bench save  N K KFILE             precompute the SOCS kernels and save them
//   litho_bench copy  MIB                   device-to-device copy of MIB MiB (counter calibration)
//   litho_bench pair  N M REPEATS           the L-BFGS pair kernel alone (litho_profile_lbfgs_pair)
//   litho_bench stream N R W                litho_profile_stream: a plain kernel that reads R and writes
//                                           W vectors of N^2 floats with the L-BFGS kernels' grid, the
//                                           rate that access pattern reaches (R, W: 24 0, 22 1, 24 4, 21 1)
// With KFILE the kernels are read from a file written by `save` (no precompute).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/litho_amd.h"

namespace {

void check(int rc, const char* what) {
  if (rc != LITHO_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, litho_last_error());
    std::exit(1);
  }
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

// deterministic Manhattan-like target: 20 x 20 px squares on a 30 px lattice
std::vector<float> lattice_target(int n) {
  std::vector<float> t((size_t)n * n, 0.f);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) t[(size_t)y * n + x] = ((y % 30) < 20 && (x % 30) < 20) ? 1.f : 0.f;
  return t;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s eval N K REPEATS | ctm N K ITERS | copy MIB | pair N M REPEATS | stream N R W\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  if (mode == "copy") {
    const size_t bytes = (size_t)std::atol(argv[2]) << 20;
    void *a = nullptr, *b = nullptr;
    hip_check(hipMalloc(&a, bytes), "hipMalloc");
    hip_check(hipMalloc(&b, bytes), "hipMalloc");
    hip_check(hipMemset(a, 1, bytes), "hipMemset");
    hip_check(hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice), "hipMemcpy");   // untimed
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
    for (int i = 0; i < 3; ++i) hip_check(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, nullptr), "hipMemcpyAsync");
    hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    // read + write bytes per copy
    std::printf("copied %zu bytes x3, %.4f ms per copy, %.1f GB/s (read + write)\n", bytes, ms / 3,
                2.0 * (double)bytes / (ms / 3 * 1e-3) / 1e9);
    return 0;
  }
  if (mode == "stream") {
    // litho_bench stream N R W: the plain R-read, W-write pattern over N x N floats
    if (argc < 5) return 2;
    const int sn = std::atoi(argv[2]), r = std::atoi(argv[3]), w = std::atoi(argv[4]);
    const size_t nn = (size_t)sn * (size_t)sn;
    litho_ctx* sctx = nullptr;
    check(litho_ctx_create(0, &sctx), "litho_ctx_create");
    double ms = 0.0;
    check(litho_profile_stream(sctx, nn, r, w, 20, &ms), "litho_profile_stream");
    const double bytes = (double)(r + w) * nn * sizeof(float);
    std::printf("stream n=%d reads=%d writes=%d ms=%.4f bytes=%.0f rate=%.1f GB/s\n", sn, r, w, ms, bytes,
                bytes / (ms * 1e-3) / 1e9);
    check(litho_ctx_destroy(sctx), "litho_ctx_destroy");
    return 0;
  }
  if (mode == "pair") {
    // litho_bench pair N M REPEATS: the L-BFGS pair kernel over one N x N tile
    if (argc < 5) return 2;
    const size_t nn = (size_t)std::atol(argv[2]) * (size_t)std::atol(argv[2]);
    litho_ctx* pctx = nullptr;
    check(litho_ctx_create(0, &pctx), "litho_ctx_create");
    double ms = 0.0;
    check(litho_profile_lbfgs_pair(pctx, nn, std::atoi(argv[3]), std::atoi(argv[4]), &ms), "litho_profile_lbfgs_pair");
    std::printf("pair n=%s m=%s ms=%.4f\n", argv[2], argv[3], ms);
    check(litho_ctx_destroy(pctx), "litho_ctx_destroy");
    return 0;
  }
  if (argc < 5) return 2;
  const int n = std::atoi(argv[2]), k = std::atoi(argv[3]);
  const int reps = mode == "save" ? 0 : std::atoi(argv[4]);
  const int batch = mode == "evalb" && argc > 5 ? std::atoi(argv[5]) : 1;
  if (batch < 1) return 2;
  litho_ctx* ctx = nullptr;
  check(litho_ctx_create(0, &ctx), "litho_ctx_create");
  int wr = 0, wc = 0, y0 = 0, x0 = 0, ns = 0;
  double wsum = 0.0;
  check(litho_socs_window(n, 4.0, 193.0, 1.35, 0.6, 0.9, &wr, &wc, &y0, &x0, &ns, &wsum), "litho_socs_window");
  std::vector<double> kern((size_t)k * wr * wc * 2), lam(k);
  const char* kfile = mode == "evalb" ? (argc > 6 ? argv[6] : nullptr)
                                      : (argc > 5 ? argv[5] : (mode == "save" ? argv[4] : nullptr));
  if (kfile != nullptr && mode != "save") {
    FILE* f = std::fopen(kfile, "rb");
    if (f == nullptr || std::fread(kern.data(), sizeof(double), kern.size(), f) != kern.size() ||
        std::fread(lam.data(), sizeof(double), lam.size(), f) != lam.size()) {
      std::fprintf(stderr, "cannot read kernels from %s\n", kfile);
      return 1;
    }
    std::fclose(f);
  } else {
    check(litho_socs_precompute(ctx, n, 4.0, 193.0, 1.35, 0.6, 0.9, k, kern.data(), lam.data(), nullptr),
          "litho_socs_precompute");
  }
  if (mode == "save") {
    FILE* f = std::fopen(kfile, "wb");
    if (f == nullptr) return 1;
    std::fwrite(kern.data(), sizeof(double), kern.size(), f);
    std::fwrite(lam.data(), sizeof(double), lam.size(), f);
    std::fclose(f);
    std::printf("saved %d kernels (%dx%d window) to %s\n", k, wr, wc, kfile);
    check(litho_ctx_destroy(ctx), "litho_ctx_destroy");
    return 0;
  }
  litho_plan* plan = nullptr;
  check(litho_plan_create(ctx, n, k, wr, wc, y0, x0, kern.data(), lam.data(), wsum, batch, &plan),
        "litho_plan_create");
  const size_t bytes = (size_t)batch * n * n * sizeof(float);
  float *target = nullptr, *theta = nullptr, *mask = nullptr;
  hip_check(hipMalloc(&target, bytes), "hipMalloc");
  hip_check(hipMalloc(&theta, bytes), "hipMalloc");
  hip_check(hipMalloc(&mask, bytes), "hipMalloc");
  const std::vector<float> t = lattice_target(n);
  for (int b = 0; b < batch; ++b)
    hip_check(hipMemcpy(target + (size_t)b * n * n, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy");
  check(litho_initialize_theta(ctx, target, theta, (size_t)batch * n * n), "litho_initialize_theta");
  if (mode == "eval" || mode == "evalb") {
    double st[5];
    check(litho_profile_ctm_eval(plan, batch, theta, target, 0.225, 50.0, reps, st), "litho_profile_ctm_eval");
    std::printf("n=%d k=%d batch=%d window=%dx%d stages_ms forward=%.4f column=%.4f (image %.4f adjoint %.4f) "
                "adjoint_tail=%.4f eval=%.4f\n",
                n, k, batch, wr, wc, st[0], st[1] + st[2], st[1], st[2], st[3], st[4]);
  } else if (mode == "ctm") {
    const int solves = argc > 6 ? std::atoi(argv[6]) : 3;
    litho_ctm_options o{1, reps, 0.9, 10, 1e-2, 0.225, 50.0};
    for (int i = 0; i < solves; ++i) {
      litho_ctm_info info{};
      const auto t0 = std::chrono::steady_clock::now();
      check(litho_ctm_optimize(plan, target, nullptr, &o, mask, nullptr, nullptr, 0, &info), "litho_ctm_optimize");
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("solve %d: n=%d k=%d lbfgs iterations=%d evaluations=%d final_pe=%.6f time=%.4f s (%.1f grad-evals/s)\n",
                  i, n, k, info.iterations, info.evaluations, info.final_pe, s, info.evaluations / s);
    }
  } else {
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  }
  hip_check(hipFree(target), "hipFree");
  hip_check(hipFree(theta), "hipFree");
  hip_check(hipFree(mask), "hipFree");
  check(litho_plan_destroy(plan), "litho_plan_destroy");
  check(litho_ctx_destroy(ctx), "litho_ctx_destroy");
  return 0;
}
