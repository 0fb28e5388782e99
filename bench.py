"""Headline benchmark: CTM grad-evals/s (SOCS forward + adjoint) at 2048^2, K = 24.

Workload (the single-GPU configuration the metric is quoted on): one
2048 x 2048 tile at 4 nm/px, lambda 193 nm, NA 1.35, annular source
sigma 0.6-0.9, K = 24 SOCS kernels from the device precompute (219 x 219
window), seeded synthetic Manhattan target, sigmoid resist alpha 50 /
threshold 0.225, L-BFGS history 10, 50 iterations, gradient tolerance 1e-2.
One step = one full CTM L-BFGS solve; a grad-eval is one objective
evaluation (line-search trials included).  Inputs are resident in HBM before
the timed region.  FP32 fields, FP64 loss / norm / L-BFGS scalars.

Last, with the headline context closed, a short served-tile run (Poisson
requests of 384 / 640 / 1152 / 2176 px tiles at 70 % of capacity) reports
the tile p50 / p99 latency of continuous batching against the plain
one-request-at-a-time path (`served_tiles`; `--no-served` skips it).

The CPU baseline is the FP64 numpy/scipy oracle (the reference's full-FFT
dataflow, 3K+1 N x N FFTs per grad-eval) timed on bounded samples on this
host: the 2048^2 workload on all cores (its rate is `value`), and the
512^2 CPU configuration on 1 thread and on all cores; each figure is the
median of 3 repeats.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N, K, W = 2048, 24, 219
OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)
THR, ALPHA = 0.225, 50.0
HBM_PEAK = 8.0e12


def _load(name, path):
    import importlib.util
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def algorithmic_bytes_per_eval(n, k, w):
    """16 ((K+1) N^2 + K W^2): theta/target/gradient traffic, K complex64 fields
    written and read once, kernel windows read twice."""
    return 16 * ((k + 1) * n * n + k * w * w)


def corr_len(w):
    """Length of the band-limited column correlations: the smallest power of
    two >= max(64, 2 W - 1)."""
    l = 64
    while l < 2 * w - 1:
        l *= 2
    return l


def corr_adjoint_flops(n, k, w):
    """Algorithmic FLOPs of one k_corr_adjoint launch: 5 L log2 L per length-L
    complex transform, 2K + 1 transforms per column (FFT_L(t_k) and the
    inverse after the product for every k, FFT_L(h) once), N columns."""
    l = corr_len(w)
    return (2 * k + 1) * n * 5 * l * (l.bit_length() - 1)


def corr_adjoint_bytes(n, k, w):
    """Compulsory HBM bytes of the same launch: the K window-row spectra T_k
    read and V_k written (complex64, W x N each), ghat read (W x N)."""
    return 16 * k * w * n + 8 * w * n


def pair_kernel_bytes(n, m):
    """Compulsory HBM bytes of one L-BFGS pair-kernel launch over a chained
    history of m pairs: the m + 2 points (the chain's m + 1 iterates and the
    trial) read from the iterate slots and from the gradient slots, nothing
    written (N^2 floats each)."""
    return (2 * m + 4) * 4 * n * n


HISTORY = 10
PAIR_KERNEL = f"k_lbfgs_pair<{HISTORY}>"
EVAL_KERNEL = f"k_corr_adjoint<{corr_len(W)}>"
PAIR_TRAFFIC_FILE = ROOT / "profiles" / f"k_lbfgs_pair_{HISTORY}_2048_hbm_traffic.md"
EVAL_TRAFFIC_FILE = ROOT / "profiles" / f"k_corr_adjoint_{corr_len(W)}_2048_hbm_traffic.md"
EVAL_TRAFFIC_FILE_ALL = ROOT / "profiles" / "socs_eval_2048_k24_hbm_traffic.md"
FP32_PEAK = 157.3e12


def measured_traffic(path, kernel):
    """HBM bytes per launch of a kernel from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None."""
    key = f"traffic_bytes_per_launch {kernel}:"
    try:
        for line in path.read_text().splitlines():
            if line.startswith(key):
                return int(line[len(key):].strip())
    except OSError:
        pass
    return None


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(cache, target, threads, evals, repeats=3):
    """Median over `repeats` of evals / wall time of the oracle grad-eval, and its PE."""
    from oracle import ctm as oc
    from oracle import imaging as im
    im.set_threads(threads)
    theta = oc.initialize_theta(target.astype(np.float64))
    t64 = target.astype(np.float64)
    pe = im.ctm_evaluate(cache, theta, t64, THR, ALPHA)[0]          # warm-up (plans, page-in)
    rates = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for _ in range(evals):
            im.ctm_evaluate(cache, theta, t64, THR, ALPHA)
        rates.append(evals / (time.perf_counter() - t0))
    return float(np.median(rates)), pe


def cpu_baseline(dk, target, ctx, tg, evals_2048=2):
    """The oracle timed on this host (bounded: about 20-40 s in all)."""
    from oracle import imaging as im
    from oracle import optics as op
    cores = len(os.sched_getaffinity(0))
    sk = op.SocsKernels(n=N, win_y0=dk["y0"], win_x0=dk["x0"], kernels=dk["kernels"], weights=dk["weights"],
                        weight_sum=dk["weight_sum"])
    rate, _ = _oracle_rate(im.Cache.from_socs(sk), target, cores, evals_2048)
    # the CPU configuration (512^2, K = 24): 1 thread and all cores, same result
    d5 = ctx.socs_precompute(512, OPT["pixel_nm"], OPT["wavelength_nm"], OPT["na"], OPT["sigma_in"],
                             OPT["sigma_out"], K)
    c5 = im.Cache.from_socs(op.SocsKernels(n=512, win_y0=d5["y0"], win_x0=d5["x0"], kernels=d5["kernels"],
                                           weights=d5["weights"], weight_sum=d5["weight_sum"]))
    t5 = tg.manhattan_target(512, 512)
    r1, pe1 = _oracle_rate(c5, t5, 1, 3)
    rn, pen = _oracle_rate(c5, t5, cores, 3)
    im.set_threads(1)
    return {"value": round(rate, 4), "unit": "grad-evals/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"FP64 numpy/scipy oracle (full N x N FFTs, 3K+1 per grad-eval): median of 3 repeats of "
                      f"{evals_2048} grad-evals of the 2048^2 K=24 workload on {cores} threads",
            "c1_512": {"threads_1": round(r1, 3), f"threads_{cores}": round(rn, 3), "unit": "grad-evals/s",
                       "sample": "median of 3 repeats of 3 grad-evals, 512^2 K=24",
                       "pe_rel_diff_1_vs_all": abs(pe1 - pen) / abs(pen)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true",
                    help="skip the served-tile latency measurement (tile p99 vs plain)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(tiles would shard across GPUs, one optimizer pool per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch

    la = _load("litho_amd", ROOT / "c-version-lithography_amd" / "litho_amd.py")
    tg = _load("litho_targets", ROOT / "c-version-lithography_amd" / "targets.py")

    ctx = la.Context(0)
    t0 = time.perf_counter()
    dk = ctx.socs_precompute(N, OPT["pixel_nm"], OPT["wavelength_nm"], OPT["na"], OPT["sigma_in"],
                             OPT["sigma_out"], K)
    t_pre = time.perf_counter() - t0
    plan = la.Plan(ctx, N, dk["kernels"], dk["weights"], dk["weight_sum"], dk["y0"], dk["x0"], capacity=1)
    target = tg.manhattan_target(N, 2048)
    d_target = torch.from_numpy(target).cuda()
    out = torch.zeros_like(d_target)
    torch.cuda.synchronize()
    print(f"setup: precompute {t_pre:.2f} s, window {dk['rows']}x{dk['cols']}, eigen gap at K "
          f"{dk['eigen_gap']:.4f}, plan {plan.device_bytes()/2**20:.0f} MiB",
          file=sys.stderr, flush=True)

    def solve():
        return plan.ctm_optimize(d_target, out, method="lbfgs", iterations=args.iterations, history=10, tol=1e-2,
                                 threshold=THR, alpha=ALPHA)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    graphs0 = ctx.graph_counts()
    evals = 0
    infos = []
    # HIP events around every pair-kernel launch of the timed solves, on the
    # library stream the kernel runs on
    ctx.time_pair_kernel(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = solve()
        evals += r["evaluations"]
        infos.append(r)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pair_launches, pair_total_ms = ctx.pair_kernel_time()
    ctx.time_pair_kernel(False)
    graphs1 = ctx.graph_counts()
    value = evals / elapsed
    print(f"timed: {args.steps} solves, {evals} grad-evals in {elapsed:.3f} s; final PE {infos[-1]['final_pe']:.3f} "
          f"status {infos[-1]['status']}", file=sys.stderr, flush=True)

    # per-stage HIP-event timing on the library stream
    theta0 = torch.zeros_like(d_target)
    la.initialize_theta(ctx, d_target, theta0)
    st = plan.profile_eval(theta0, d_target, threshold=THR, alpha=ALPHA, repeats=20)
    if st[2] <= 0.0:
        raise RuntimeError("the plan did not take the band-limited column path (no correlation adjoint timed)")
    adj_s = st[2] * 1e-3
    adj_flops = corr_adjoint_flops(N, K, W)
    adj_bytes = corr_adjoint_bytes(N, K, W)
    eval_bytes = algorithmic_bytes_per_eval(N, K, W)
    # the pair kernel's mean launch time inside the timed solves (one launch per L-BFGS iteration)
    if pair_launches <= 0:
        raise RuntimeError("no pair-kernel launch was timed in the solves")
    pair_ms = pair_total_ms / pair_launches
    pair_bytes = pair_kernel_bytes(N, HISTORY)
    # the yardstick: a plain kernel with the pair kernel's access pattern
    # (2m + 4 = 24 vectors read, none written, the same grid and loads),
    # each launch after a read that evicts the last-level cache, as the
    # sweeps of a solve run after an evaluation; HIP events
    pattern_ms = ctx.profile_stream(N * N, 2 * HISTORY + 4, 0, repeats=20)
    print(f"stages (ms): forward {st[0]:.3f} column image {st[1]:.3f} column adjoint {st[2]:.3f} "
          f"adjoint-tail {st[3]:.3f} eval {st[4]:.3f}; {PAIR_KERNEL} {pair_ms:.4f} (mean of {pair_launches} "
          f"launches in the timed solves)", file=sys.stderr, flush=True)
    line = {
        "metric": "CTM grad-evals/s (SOCS fwd+adjoint) 2048^2, K=24",
        "value": round(value, 3), "unit": "grad-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (fp64 reductions)", "data": "synthetic",
        "config": {"workload": "single-tile CTM L-BFGS, 2048x2048 @4nm, K=24 (219x219 window), m=10, "
                               "50 iterations, tol 1e-2, seeded Manhattan target",
                   "n": N, "k": K, "window": W, "evals_per_step": round(evals / args.steps, 2),
                   "eval_graphs": {k: graphs1[k] - graphs0[k] for k in graphs1}},
        # The kernel with the most time in the timed solves: one pair-kernel
        # launch per L-BFGS iteration, HBM-bound.
        "roofline": {"bound": "hbm", "kernel": PAIR_KERNEL,
                     "achieved": round(pair_bytes / (pair_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(pair_bytes / (pair_ms * 1e-3) / HBM_PEAK, 4),
                     "traffic": measured_traffic(PAIR_TRAFFIC_FILE, PAIR_KERNEL), "algorithmic_bytes": pair_bytes,
                     "launch_ms": round(pair_ms, 4), "launches_timed": pair_launches,
                     # the rate a plain kernel with the same access pattern
                     # reaches on this GPU (litho_profile_stream), and the
                     # in-solve launches' fraction of it
                     "pattern_GBps": round(pair_bytes / (pattern_ms * 1e-3) / 1e9, 1),
                     "frac_of_pattern": round(pattern_ms / pair_ms, 4)},
        # The dominant kernel of a grad-eval, bound by FP32 vector (VALU) FFT
        # arithmetic, priced against MI355X's published FP32 peak (packed FP32
        # is not used, so the issue ceiling is half of it); its compulsory HBM
        # bytes are reported beside.
        "eval_kernel_roofline": {"bound": "fp32 valu", "kernel": EVAL_KERNEL,
                                 "achieved": round(adj_flops / adj_s / 1e12, 2), "peak": FP32_PEAK / 1e12,
                                 "unit": "TFLOP/s", "frac": round(adj_flops / adj_s / FP32_PEAK, 4),
                                 "traffic": measured_traffic(EVAL_TRAFFIC_FILE, EVAL_KERNEL),
                                 "algorithmic_flops": adj_flops, "launch_ms": round(st[2], 4),
                                 "hbm": {"algorithmic_bytes": adj_bytes,
                                         "achieved_GBps": round(adj_bytes / adj_s / 1e9, 1),
                                         "frac": round(adj_bytes / adj_s / HBM_PEAK, 4)}},
        # One whole grad-eval as the solves run it (theta written, gradient
        # stored, g.d formed), priced with the reference's materialised-field
        # bytes, and with the HBM bytes the kernels actually move (counters).
        "eval_roofline": {"algorithmic_bytes": eval_bytes, "eval_ms": round(st[4], 4),
                          "achieved_GBps": round(eval_bytes / (st[4] * 1e-3) / 1e9, 1),
                          "frac": round(eval_bytes / (st[4] * 1e-3) / HBM_PEAK, 4),
                          "traffic": measured_traffic(EVAL_TRAFFIC_FILE_ALL, "grad-eval"),
                          "traffic_frac": (None if measured_traffic(EVAL_TRAFFIC_FILE_ALL, "grad-eval") is None else
                                           round(measured_traffic(EVAL_TRAFFIC_FILE_ALL, "grad-eval") /
                                                 (st[4] * 1e-3) / HBM_PEAK, 4))},
    }
    plan.close()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(dk, target, ctx, tg)
    # The served run's four pools (one stream each) need a hardware queue
    # each, and HIP gives a process four by default: a fifth stream that has
    # run work (this context's) makes two pools share one, which doubled the
    # served p50.  So this context is closed first.
    ctx.close()
    if not args.no_served:
        # the metric's second half, tile p99 latency vs plain: Poisson requests
        # of 384 / 640 / 1152 / 2176 px tiles (20 GD iterations each) at 70 % of
        # the server's measured capacity, continuous batching against one
        # request at a time (bench_tiles.py served, 120 requests)
        bt = _load("litho_bench_tiles", ROOT / "c-version-lithography_amd" / "bench_tiles.py")
        sv = bt.served_measure(requests=120, iterations=20, load=0.7)
        line["served_tiles"] = {k: sv[k] for k in ("capacity_tiles_per_s", "offered_tiles_per_s", "served", "plain",
                                                   "requests", "sizes", "iterations",
                                                   "identical_masks_alone_vs_mixed")}
        print(f"served: p50 {sv['served']['p50_ms']} ms, p99 {sv['served']['p99_ms']} ms vs plain p99 "
              f"{sv['plain']['p99_ms']} ms", file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
