"""Tile-throughput benchmarks of the batched and served paths (one GPU).

  python c-version-lithography_amd/bench_tiles.py batched [--tiles 96] [--iterations 30]
      96 independent tiles, Cores 512 / 768 / 1024 (1:1:1) with a 64-pixel halo
      per side -> N = 640 / 896 / 1152 (SOCS windows 67 / 97 / 123), K = 24,
      30 L-BFGS iterations (m = 10), seeded Manhattan targets.  Plain: every
      tile solved alone (litho_ctm_optimize, one tile per call).  Batched:
      litho_batch_ctm with autotuned chunks, one context (stream) per size, after one untimed
      warm-up run of the whole batch (the chunk autotune is cached in the plans).  Prints one JSON
      line with grad-evals/s of both and their ratio.  --skip-plain times the batched run only.

  python c-version-lithography_amd/bench_tiles.py served [--requests 160] [--iterations 20] [--load 0.7]
      Requests of sides 256 / 512 / 1024 / 2048 plus a 64-pixel halo per side
      (N = 384 / 640 / 1152 / 2176, uniformly drawn, seeded), K = 24, 20 CTM
      gradient-descent iterations each (one evaluation per iteration: a
      deterministic service time).  Capacity = requests/s of the server with
      every request submitted at once; then a seeded Poisson arrival trace at
      `load` x capacity is replayed against (a) the continuous-batching server
      (litho_server_*) and (b) the plain path, one request at a time in arrival
      order (litho_ctm_optimize).  Prints one JSON line with tiles/s, p50 and
      p99 latency (arrival -> result) of both, and whether every request's mask
      is the same in both runs.

  python c-version-lithography_amd/bench_tiles.py fullchip [--chip 8192] [--grid 16] [--updates 20]
      Full-chip Micro optimisation of a seeded Manhattan chip: 16 x 16 Micros
      of Core 512 + halo 64 (tile 640), guard 40, K = 24, PE loss, Average
      Guard fusion 0.5 / 0.5, guard-only CEPT.  Full-chip CTM (5 iterations,
      lr 0.9), the best mask thresholded at 0.5, then full-chip LSM from that
      binary mask (WENO5 / Godunov, CFL 0.5, curvature 0.01, re-init every 20
      updates, Binary render).  Prints one JSON line with LSM updates/s (every
      update = one RK2 step of every Micro + the batched exchange) and CTM
      iterations/s.

Inputs are resident in HBM before the timed regions; the SOCS precompute and
plan creation are set-up, outside the timers.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OPTICS = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)
THR, ALPHA, K = 0.225, 50.0, 24


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


la = _load("litho_amd", HERE / "litho_amd.py")
tg = _load("litho_targets", HERE / "targets.py")


def make_plan(ctx, n, capacity):
    d = ctx.socs_precompute(n, OPTICS["pixel_nm"], OPTICS["wavelength_nm"], OPTICS["na"], OPTICS["sigma_in"],
                            OPTICS["sigma_out"], K)
    return la.Plan(ctx, n, d["kernels"], d["weights"], d["weight_sum"], d["y0"], d["x0"], capacity=capacity), d


def batched(args):
    import torch
    sizes = [512 + 2 * 64, 768 + 2 * 64, 1024 + 2 * 64]
    per = args.tiles // 3
    ctxs = [la.Context(0) for _ in sizes]     # one stream per tile size: the groups run side by side
    plans, wins = [], []
    t0 = time.perf_counter()
    for ctx, n in zip(ctxs, sizes):
        p, d = make_plan(ctx, n, per)
        plans.append(p)
        wins.append(d["rows"])
    print(f"set-up: {time.perf_counter() - t0:.1f} s, windows {wins}", file=sys.stderr, flush=True)
    tiles = []
    for pi, n in enumerate(sizes):
        for j in range(per):
            t = torch.from_numpy(tg.manhattan_target(n, 1000 * pi + j)).cuda()
            tiles.append((pi, t, torch.zeros_like(t)))
    torch.cuda.synchronize()
    kw = dict(method="lbfgs", iterations=args.iterations, history=10, tol=0.0, threshold=THR, alpha=ALPHA)
    # warm-up: one untimed run of the whole batch (first launches, and the
    # per-plan chunk autotune, which is set-up: it is cached in the plans)
    la.batch_ctm(plans, tiles, **kw)
    torch.cuda.synchronize()
    plain_masks = None
    ev_plain, t_plain = 0, float("nan")
    if not args.skip_plain:
        # plain: one tile per call
        t0 = time.perf_counter()
        for pi, t, out in tiles:
            r = plans[pi].ctm_optimize(t, out, **kw)
            ev_plain += r["evaluations"]
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        plain_masks = [out.clone() for _, _, out in tiles]
        print(f"plain: {len(tiles)} tiles, {ev_plain} grad-evals in {t_plain:.3f} s", file=sys.stderr, flush=True)
    # batched
    for _, _, out in tiles:
        out.zero_()
    t0 = time.perf_counter()
    info, used = la.batch_ctm(plans, tiles, **kw)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    ev_batch = sum(i["evaluations"] for i in info)
    same = None if plain_masks is None else all(torch.equal(a, out) for a, (_, _, out) in zip(plain_masks, tiles))
    print(f"batched: chunks {used}, {ev_batch} grad-evals in {t_batch:.3f} s; masks equal to plain: {same}",
          file=sys.stderr, flush=True)
    plain_rate = ev_plain / t_plain if ev_plain else None
    line = {"metric": "batched uneven-tile CTM grad-evals/s vs plain (one tile per call)",
            "value": round(ev_batch / t_batch, 2), "plain": round(plain_rate, 2) if plain_rate else None,
            "speedup": round((ev_batch / t_batch) / plain_rate, 3) if plain_rate else None, "unit": "grad-evals/s",
            "tiles": len(tiles), "sizes": sizes, "windows": wins, "k": K, "iterations": args.iterations,
            "chunks": used, "evaluations": ev_batch, "batched_s": round(t_batch, 4), "plain_s": round(t_plain, 4) if ev_plain else None,
            "identical_masks": same}
    for p in plans:
        p.close()
    for ctx in ctxs:
        ctx.close()
    print(json.dumps(line), flush=True)


def served(args):
    print(json.dumps(served_measure(args.requests, args.iterations, args.load)), flush=True)


def served_measure(requests=160, iterations=20, load=0.7):
    """The served-tile measurement (see the module docstring); returns the JSON dict."""
    import torch
    from types import SimpleNamespace
    args = SimpleNamespace(requests=requests, iterations=iterations, load=load)
    sizes = [256 + 128, 512 + 128, 1024 + 128, 2048 + 128]
    caps = [16, 16, 8, 4]
    ctxs = [la.Context(0) for _ in sizes]     # one worker and stream per tile size
    plans, wins = [], []
    t0 = time.perf_counter()
    for ctx, n, c in zip(ctxs, sizes, caps):
        p, d = make_plan(ctx, n, c)
        plans.append(p)
        wins.append(d["rows"])
    print(f"set-up: {time.perf_counter() - t0:.1f} s, windows {wins}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(2024)
    kinds = rng.integers(0, len(sizes), size=args.requests)
    tgts = [torch.from_numpy(tg.manhattan_target(sizes[k], 5000 + i)).cuda() for i, k in enumerate(kinds)]
    kw = dict(method="gradient_descent", iterations=args.iterations, learning_rate=0.9, threshold=THR, alpha=ALPHA)
    torch.cuda.synchronize()

    def run_server(arrivals):
        outs = [torch.zeros_like(t) for t in tgts]
        srv = la.Server(plans, **kw)
        # warm: first launches and graph capture of every pool
        warm = [srv.submit(k, tgts[i], torch.zeros_like(tgts[i])) for i, k in
                [(int(np.argmax(kinds == j)), j) for j in range(len(sizes)) if (kinds == j).any()]]
        for w in warm:
            srv.wait(w)
        start = time.perf_counter()
        tickets, sub = [], []
        for i, k in enumerate(kinds):
            if arrivals is not None:
                delay = start + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            sub.append(time.perf_counter() - start)
            tickets.append(srv.submit(int(k), tgts[i], outs[i]))
        res = [srv.wait(tk) for tk in tickets]
        end = time.perf_counter() - start
        srv.close()
        arr = np.array(sub) if arrivals is None else np.asarray(arrivals)
        # arrival -> result: submission lag behind the schedule + queue + service
        lat = np.array([(sub[i] - arr[i]) * 1e3 + r["queued_ms"] + r["service_ms"] for i, r in enumerate(res)])
        return outs, lat, end, sum(r["evaluations"] for r in res), arr

    outs_cap, lat_cap, span_cap, ev_cap, _ = run_server(None)
    capacity = args.requests / span_cap
    print(f"server capacity: {args.requests} requests in {span_cap:.3f} s = {capacity:.1f} req/s "
          f"({ev_cap / span_cap:.0f} grad-evals/s)", file=sys.stderr, flush=True)
    rate = args.load * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=args.requests))
    outs_srv, lat_srv, span_srv, ev_srv, _ = run_server(arrivals)
    # plain: one request at a time, in arrival order, same trace
    outs_plain = [torch.zeros_like(t) for t in tgts]
    plans[0].ctm_optimize(tgts[0], torch.zeros_like(tgts[0]), **kw)
    lat_plain = np.zeros(args.requests)
    start = time.perf_counter()
    for i, k in enumerate(kinds):
        delay = start + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        plans[int(k)].ctm_optimize(tgts[i], outs_plain[i], **kw)
        torch.cuda.synchronize()
        lat_plain[i] = (time.perf_counter() - start - arrivals[i]) * 1e3
    span_plain = time.perf_counter() - start
    same = all(torch.equal(a, b) for a, b in zip(outs_srv, outs_plain)) and \
        all(torch.equal(a, b) for a, b in zip(outs_cap, outs_plain))
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)
    line = {"metric": "served CTM tiles/s and latency, continuous batching vs plain",
            "value": round(args.requests / span_srv, 2), "unit": "tiles/s",
            "capacity_tiles_per_s": round(capacity, 2), "offered_load": args.load,
            "offered_tiles_per_s": round(rate, 2),
            "served": {"tiles_per_s": round(args.requests / span_srv, 2), "p50_ms": pct(lat_srv, 50),
                       "p99_ms": pct(lat_srv, 99), "grad_evals_per_s": round(ev_srv / span_srv, 1)},
            "plain": {"tiles_per_s": round(args.requests / span_plain, 2), "p50_ms": pct(lat_plain, 50),
                      "p99_ms": pct(lat_plain, 99)},
            "requests": args.requests, "sizes": sizes, "windows": wins, "slots": caps, "k": K,
            "iterations": args.iterations, "optimizer": "gradient_descent",
            "identical_masks_alone_vs_mixed": same}
    for p in plans:
        p.close()
    for ctx in ctxs:
        ctx.close()
    return line


def fullchip(args):
    import torch
    ctx = la.Context(0)
    core = args.chip // args.grid
    halo, guard = 64, 40
    ts, tiles = la.partition_build(args.chip, args.chip, args.grid, args.grid, halo, guard)
    t0 = time.perf_counter()
    plan, d = make_plan(ctx, ts, len(tiles))
    target = torch.from_numpy(tg.manhattan_target(args.chip, 8192)).cuda()
    torch.cuda.synchronize()
    print(f"set-up: {time.perf_counter() - t0:.1f} s; {len(tiles)} Micros of {ts} (core {core}), window {d['rows']}",
          file=sys.stderr, flush=True)
    opts = la.exchange_options(loss_type="PE", enable_cept=True, cept_guard_only=True, enable_grad_fusion=True,
                               fusion_mode="Average", fusion_weight_owner=0.5, fusion_weight_neighbor=0.5,
                               threshold=THR, alpha=ALPHA)
    gmask = torch.zeros_like(target)
    plan.fullchip_ctm(tiles, target, target, opts, 1, 0.9, gmask)          # warm: first launches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r_ctm = plan.fullchip_ctm(tiles, target, target, opts, args.ctm_iterations, 0.9, gmask)
    torch.cuda.synchronize()
    t_ctm = time.perf_counter() - t0
    # the reference's hand-off: Imbinarize::binarize_advanced(mask, target, 0.125, 0.1, 5, 5)
    # (demo_full_chip_ctm.cpp:109); its final mask seeds the level set
    binary = torch.zeros_like(target)
    t0 = time.perf_counter()
    thr, npk = ctx.binarize_advanced(gmask, target, binary, 0.125, 0.1, 5, 5)
    t_bin = time.perf_counter() - t0
    print(f"CTM: {args.ctm_iterations} iterations in {t_ctm:.3f} s, best loss {r_ctm['best_loss']:.1f} "
          f"(iteration {r_ctm['best_iteration']}); binarize_advanced {1e3 * t_bin:.1f} ms: threshold {thr:.4f} "
          f"from {npk} peaks, {int(binary.sum().item())} mask pixels", file=sys.stderr, flush=True)
    lmask = torch.zeros_like(target)
    lphi = torch.zeros_like(target)
    plan.fullchip_lsm(tiles, binary, target, opts, 1, lmask, lphi, cfl=0.5, curvature_weight=0.01,
                      reinit_interval=20)                                   # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r_lsm = plan.fullchip_lsm(tiles, binary, target, opts, args.updates, lmask, lphi, cfl=0.5, curvature_weight=0.01,
                              reinit_interval=20)
    torch.cuda.synchronize()
    t_lsm = time.perf_counter() - t0
    h = r_lsm["history"]
    print(f"LSM: {args.updates} updates in {t_lsm:.3f} s; loss {h[0][0]:.1f} -> {h[-1][0]:.1f}, "
          f"best {r_lsm['best_loss']:.1f} (update {r_lsm['best_iteration']}), dt {h[1][3]:.3f}..{h[1][4]:.3f}",
          file=sys.stderr, flush=True)
    px = len(tiles) * ts * ts
    line = {"metric": "full-chip LSM updates/s (RK2 WENO5/Godunov of every Micro + batched exchange)",
            "value": round(args.updates / t_lsm, 3), "unit": "updates/s", "ms_per_update": round(1e3 * t_lsm / args.updates, 2),
            "ctm_iterations_per_s": round(args.ctm_iterations / t_ctm, 3),
            "chip": args.chip, "micros": len(tiles), "tile": ts, "core": core, "halo": halo, "guard": guard,
            "k": K, "window": d["rows"], "micro_pixels": px,
            "stencil_GBps_at_40B_per_px": round(40 * px / (t_lsm / args.updates) / 1e9, 1),
            "lsm_loss_first_last": [round(h[0][0], 3), round(h[-1][0], 3)],
            "ctm_best_loss": round(r_ctm["best_loss"], 3), "binarize_ms": round(1e3 * t_bin, 2),
            "binarize_threshold": round(thr, 6)}
    plan.close()
    ctx.close()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("batched")
    b.add_argument("--tiles", type=int, default=96)
    b.add_argument("--iterations", type=int, default=30)
    b.add_argument("--skip-plain", action="store_true", help="batched run only (profiling)")
    v = sub.add_parser("served")
    v.add_argument("--requests", type=int, default=160)
    v.add_argument("--iterations", type=int, default=20)
    v.add_argument("--load", type=float, default=0.7)
    f = sub.add_parser("fullchip")
    f.add_argument("--chip", type=int, default=8192)
    f.add_argument("--grid", type=int, default=16)
    f.add_argument("--updates", type=int, default=20)
    f.add_argument("--ctm-iterations", type=int, default=5)
    args = ap.parse_args()
    if args.cmd == "batched":
        batched(args)
    elif args.cmd == "served":
        served(args)
    else:
        fullchip(args)


if __name__ == "__main__":
    main()
