"""The batched (C3) and served (C4) tile sizes with real K = 24 SOCS kernels
from the device precompute, on the GPU.

* Parity at 384, 640, 896, 1152 and 2176 px against the FP64 oracle on the
  same kernels: aerial relative L-inf <= 1e-5, gradient relative L2 <= 1e-4,
  PE relative <= 1e-5; the eigen-gap at K is printed and must be a real gap.
* litho_batch_ctm over 640 / 896 / 1152 tiles (two each, 3 L-BFGS
  iterations) equals the single-tile solves bit for bit.
* The server over 384 / 640 / 1152 / 2176 (one request each, then a mix)
  equals each request solved alone bit for bit (20 GD iterations, the C4
  optimizer)."""
import os

import numpy as np
import pytest

from conftest import litho_amd, targets
from oracle import imaging as im
from oracle import optics as op
from test_gpu_socs import dev, empty, host, rel_l2, rel_linf

pytestmark = pytest.mark.gpu

_DK = {}


def kernels(ctx, n):
    """Device-precomputed K = 24 kernels at tile size n (cached per session)."""
    if n not in _DK:
        dk = ctx.socs_precompute(n, 4.0, 193.0, 1.35, 0.6, 0.9, 24)
        _DK[n] = (op.SocsKernels(n=n, win_y0=dk["y0"], win_x0=dk["x0"], kernels=dk["kernels"],
                                 weights=dk["weights"], weight_sum=dk["weight_sum"]), dk["eigen_gap"])
    return _DK[n][0]


@pytest.mark.parametrize("n", [384, 640, 896, 1152, 2176])
def test_parity_real_kernels(gpu_ctx, n):
    print(f"case parity real K=24 n={n}", flush=True)
    la = litho_amd()
    sk = kernels(gpu_ctx, n)
    gap = _DK[n][1]
    plan = la.Plan.from_socs(gpu_ctx, sk)
    tgt = targets().manhattan_target(n, 3000 + n)
    rng = np.random.default_rng(n)
    mask = np.clip(tgt + 0.3 * rng.standard_normal(tgt.shape), 0.0, 1.0).astype(np.float32)
    a, w, gm = empty(n), empty(n), empty(n)
    pe, _ = plan.pe_gradient(dev(mask), dev(tgt), grad_mask=gm, aerial=a, wafer=w, threshold=0.225, alpha=50.0)
    im.set_threads(len(os.sched_getaffinity(0)))
    cache = im.Cache.from_socs(sk)
    ra, rw, fields = im.forward(cache, mask.astype(np.float64), 0.225, 50.0)
    rg = im.pe_gradient(cache, rw, tgt.astype(np.float64), fields, 50.0)
    del fields
    im.set_threads(1)
    ea, eg = rel_linf(host(a)[0], ra), rel_l2(host(gm)[0], rg)
    rpe = im.pe_loss(rw, tgt)
    print(f"n={n} window {sk.kernels.shape[1]}x{sk.kernels.shape[2]} eigen gap {gap:.4f}: aerial {ea:.2e}, "
          f"gradient {eg:.2e}, PE {pe[0]:.4f} vs {rpe:.4f}")
    assert gap > 0.05
    assert ea <= 1e-5 and eg <= 1e-4 and abs(pe[0] - rpe) <= 1e-5 * rpe
    plan.close()


def test_batch_c3_sizes_equal_single_tiles(gpu_ctx):
    import torch
    la = litho_amd()
    sizes = [640, 896, 1152]
    plans = [la.Plan.from_socs(gpu_ctx, kernels(gpu_ctx, n), capacity=2) for n in sizes]
    tiles = []
    for pi, n in enumerate(sizes):
        for j in range(2):
            t = dev(targets().manhattan_target(n, 900 + 10 * pi + j))
            tiles.append((pi, t, torch.zeros_like(t)))
    print("case batch C3 sizes 640/896/1152 x2, 3 L-BFGS iterations", flush=True)
    info, used = la.batch_ctm(plans, tiles, method="lbfgs", iterations=3, threshold=0.225, alpha=50.0, chunk=2)
    for (pi, t, out), inf in zip(tiles, info):
        ref = torch.zeros_like(t)
        r = plans[pi].ctm_optimize(t, ref, method="lbfgs", iterations=3, threshold=0.225, alpha=50.0)
        assert r["evaluations"] == inf["evaluations"]
        assert r["final_pe"] == inf["final_pe"]
        assert torch.equal(out, ref)
    print(f"C3 sizes batched = single: evaluations {[i['evaluations'] for i in info]}")
    for p in plans:
        p.close()


def test_server_c4_sizes_equal_alone():
    import torch
    la = litho_amd()
    ctx = la.Context(0)
    sizes = [384, 640, 1152, 2176]
    plans = [la.Plan.from_socs(ctx, kernels(ctx, n), capacity=2) for n in sizes]
    # one request per size, then a mix of sizes
    order = [0, 1, 2, 3, 0, 1, 0, 2, 1, 3]
    reqs = []
    for i, pi in enumerate(order):
        t = dev(targets().manhattan_target(sizes[pi], 4000 + i))
        reqs.append((pi, t, torch.zeros_like(t)))
    refs = []
    for pi, t, _ in reqs:
        out = torch.zeros_like(t)
        r = plans[pi].ctm_optimize(t, out, method="gradient_descent", iterations=20, learning_rate=0.9,
                                   threshold=0.225, alpha=50.0)
        refs.append((r, out))
    torch.cuda.synchronize()
    print("case server C4 sizes 384/640/1152/2176, 20 GD iterations", flush=True)
    srv = la.Server(plans, method="gradient_descent", iterations=20, threshold=0.225, alpha=50.0)
    tickets = [srv.submit(pi, t, out) for pi, t, out in reqs[:4]]
    results = [srv.wait(tk) for tk in tickets]
    tickets = [srv.submit(pi, t, out) for pi, t, out in reqs[4:]]
    results += [srv.wait(tk) for tk in tickets]
    srv.close()
    for (pi, t, out), r, (ref, ref_out) in zip(reqs, results, refs):
        assert r["evaluations"] == ref["evaluations"] == 20
        assert r["final_pe"] == ref["final_pe"]
        assert torch.equal(out, ref_out)
    print(f"C4 sizes served = alone: {len(reqs)} requests")
    for p in plans:
        p.close()
    ctx.close()
