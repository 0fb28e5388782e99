"""Batched CTM over tiles of several sizes (litho_batch_ctm) and the captured
evaluation graphs, on the GPU.

The batched solve must reproduce each tile's single-tile solve bit for bit:
evaluations are per tile with fixed-order reductions, so sharing a launch
changes nothing.  The graph-replay path must equal direct launches bit for
bit over a whole optimizer run (every evaluation replays the graph, with the
host's uploads and synchronisations in between)."""
import numpy as np
import pytest

from conftest import litho_amd, targets
from oracle import optics as op

pytestmark = pytest.mark.gpu

OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)
_SK = {}


def socs(n, k=8):
    if (n, k) not in _SK:
        _SK[(n, k)] = op.socs_kernels(n, **OPT, num_kernels=k)
    return _SK[(n, k)]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("method,iters,chunk", [("lbfgs", 6, 0), ("lbfgs", 6, 2), ("gradient_descent", 5, 0)])
def test_batch_equals_single_tile_solves(gpu_ctx, method, iters, chunk):
    import torch
    la = litho_amd()
    tg = targets()
    sizes = [64, 80, 128]
    plans = [la.Plan.from_socs(gpu_ctx, socs(n), capacity=3) for n in sizes]
    tiles, singles = [], []
    for pi, n in enumerate(sizes):
        for j in range(3):
            t = dev(tg.manhattan_target(n, 100 * pi + j))
            out = torch.zeros_like(t)
            tiles.append((pi, t, out))
    print(f"case batch {method} iters {iters} chunk {chunk}", flush=True)
    info, used = la.batch_ctm(plans, tiles, method=method, iterations=iters, threshold=0.225, alpha=50.0,
                              chunk=chunk)
    assert all(1 <= u <= 3 for u in used)
    for (pi, t, out), inf in zip(tiles, info):
        ref = torch.zeros_like(t)
        r = plans[pi].ctm_optimize(t, ref, method=method, iterations=iters, threshold=0.225, alpha=50.0)
        assert r["evaluations"] == inf["evaluations"]
        assert r["final_pe"] == inf["final_pe"]
        assert torch.equal(out, ref)
    print(f"batch {method}: chunks used {used}, evaluations {[i['evaluations'] for i in info]}")
    for p in plans:
        p.close()


@pytest.mark.parametrize("method", ["lbfgs", "gradient_descent"])
def test_graph_replay_equals_direct_launches(method):
    import torch
    la = litho_amd()
    n = 128
    runs = []
    for graphs in (True, False):
        ctx = la.Context(0)
        ctx.set_graphs(graphs)
        plan = la.Plan.from_socs(ctx, socs(n), capacity=1)
        t = dev(targets().manhattan_target(n, 5))
        out = torch.zeros_like(t)
        th = torch.zeros_like(t)
        r = plan.ctm_optimize(t, out, out_theta=th, method=method, iterations=8, threshold=0.225, alpha=50.0)
        runs.append((r, out.clone(), th.clone()))
        plan.close()
        ctx.close()
    (ra, ma, ta), (rb, mb, tb) = runs
    assert ra["evaluations"] == rb["evaluations"] and ra["evaluations"] >= 8
    assert np.array_equal(ra["history"], rb["history"])
    assert torch.equal(ma, mb) and torch.equal(ta, tb)


def test_batch_rejects_bad_plan_index(gpu_ctx):
    import torch
    la = litho_amd()
    plan = la.Plan.from_socs(gpu_ctx, socs(64), capacity=1)
    t = torch.zeros((64, 64), dtype=torch.float32, device="cuda")
    with pytest.raises(la.LithoError):
        la.batch_ctm([plan], [(1, t, t)])
    plan.close()


def test_batch_across_contexts_equals_single():
    """Plans on distinct contexts run concurrently (one host thread each)."""
    import torch
    la = litho_amd()
    sizes = [64, 96]
    ctxs = [la.Context(0) for _ in sizes]
    plans = [la.Plan.from_socs(c, socs(n), capacity=2) for c, n in zip(ctxs, sizes)]
    tiles = []
    for i in range(6):
        t = dev(targets().manhattan_target(sizes[i % 2], 90 + i))
        tiles.append((i % 2, t, torch.zeros_like(t)))
    info, used = la.batch_ctm(plans, tiles, method="lbfgs", iterations=5, threshold=0.225, alpha=50.0, chunk=2)
    torch.cuda.synchronize()
    for (pi, t, out), inf in zip(tiles, info):
        ref = torch.zeros_like(t)
        r = plans[pi].ctm_optimize(t, ref, method="lbfgs", iterations=5, threshold=0.225, alpha=50.0)
        assert r["evaluations"] == inf["evaluations"] and torch.equal(out, ref)
    for p in plans:
        p.close()
    for c in ctxs:
        c.close()


def test_band_limited_batch_equals_single_tile(gpu_ctx):
    """At n = 640, K = 24 a single tile splits its power spectra and its
    correlation adjoint over 8 k groups to fill the chip, where a batch of 6
    tiles would split them differently; the aerial images, gradients and PE of
    the batched evaluation must still equal the single-tile ones bit for bit."""
    import torch
    from test_gpu_socs import window_kernels
    la = litho_amd()
    n, k, b = 640, 24, 6
    plan = la.Plan.from_socs(gpu_ctx, window_kernels(n, k), capacity=b)
    tg = targets()
    tgt = np.stack([tg.manhattan_target(n, 40 + i) for i in range(b)]).astype(np.float32)
    rng = np.random.default_rng(3)
    mask = np.clip(tgt + 0.3 * rng.standard_normal(tgt.shape), 0, 1).astype(np.float32)
    print(f"case band-limited batch n={n} k={k} batch={b}", flush=True)
    a_b = torch.zeros((b, n, n), dtype=torch.float32, device="cuda")
    g_b = torch.zeros_like(a_b)
    pe_b, _ = plan.pe_gradient(dev(mask), dev(tgt), grad_mask=g_b, aerial=a_b, threshold=0.225, alpha=50.0, batch=b)
    for i in range(b):
        a1 = torch.zeros((1, n, n), dtype=torch.float32, device="cuda")
        g1 = torch.zeros_like(a1)
        pe1, _ = plan.pe_gradient(dev(mask[i]), dev(tgt[i]), grad_mask=g1, aerial=a1, threshold=0.225, alpha=50.0)
        assert torch.equal(a1[0], a_b[i])
        assert torch.equal(g1[0], g_b[i])
        assert pe1[0] == pe_b[i]
    plan.close()
