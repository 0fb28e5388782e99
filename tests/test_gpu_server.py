"""The continuous-batching CTM server (litho_server_*) on the GPU.

Every served request must equal the same request solved alone
(litho_ctm_optimize) bit for bit, whatever it shared its batches with, under
slot pressure (more requests than slots) and under a budget that admits one
request at a time."""
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


def alone(plan, t, method, iters):
    import torch
    out = torch.zeros_like(t)
    r = plan.ctm_optimize(t, out, method=method, iterations=iters, threshold=0.225, alpha=50.0)
    return r, out


@pytest.mark.parametrize("method,iters,budget_requests", [("gradient_descent", 6, None), ("lbfgs", 5, None),
                                                          ("gradient_descent", 4, 1)])
def test_served_equals_alone(method, iters, budget_requests):
    import torch
    la = litho_amd()
    ctx = la.Context(0)
    sizes = [64, 96, 128]
    plans = [la.Plan.from_socs(ctx, socs(n), capacity=2) for n in sizes]
    reqs = []
    for i in range(10):
        pi = i % 3
        t = dev(targets().manhattan_target(sizes[pi], 50 + i))
        reqs.append((pi, t, torch.zeros_like(t)))
    # reference results first (the server owns the context while it runs)
    refs = [alone(plans[pi], t, method, iters) for pi, t, _ in reqs]
    torch.cuda.synchronize()
    srv = la.Server(plans, method=method, iterations=iters, threshold=0.225, alpha=50.0)
    if budget_requests is not None:
        b = max(srv.request_bytes(i) for i in range(3))
        srv.close()
        srv = la.Server(plans, method=method, iterations=iters, threshold=0.225, alpha=50.0,
                        memory_budget=float(b * budget_requests))
    print(f"case server {method} iters {iters} budget {budget_requests}", flush=True)
    tickets = [srv.submit(pi, t, out) for pi, t, out in reqs[:6]]
    first = srv.wait(tickets[0])
    tickets += [srv.submit(pi, t, out) for pi, t, out in reqs[6:]]
    results = [first] + [srv.wait(tk) for tk in tickets[1:]]
    srv.close()
    for (pi, t, out), r, (ref, ref_out) in zip(reqs, results, refs):
        assert r["evaluations"] == ref["evaluations"]
        assert r["final_pe"] == ref["final_pe"]
        assert torch.equal(out, ref_out)
        assert r["queued_ms"] >= 0.0 and r["service_ms"] > 0.0
    print(f"server {method}: {len(reqs)} requests, evaluations {[r['evaluations'] for r in results]}")
    for p in plans:
        p.close()
    ctx.close()


def test_server_rejects_budget_below_one_request():
    la = litho_amd()
    ctx = la.Context(0)
    plan = la.Plan.from_socs(ctx, socs(64), capacity=1)
    with pytest.raises(la.LithoError):
        la.Server([plan], memory_budget=1024.0)
    plan.close()
    ctx.close()


def test_server_wait_timeout_and_bad_ticket():
    import torch
    la = litho_amd()
    ctx = la.Context(0)
    plan = la.Plan.from_socs(ctx, socs(64), capacity=1)
    srv = la.Server([plan], iterations=3)
    t = dev(targets().manhattan_target(64, 1))
    out = torch.zeros_like(t)
    tk = srv.submit(0, t, out)
    assert srv.wait(tk + 1000, timeout=0.01) is None        # unknown ticket: never done
    assert srv.wait(tk) is not None
    with pytest.raises(la.LithoError):
        srv.submit(5, t, out)
    srv.close()
    plan.close()
    ctx.close()


def test_server_across_contexts_equals_alone():
    """Plans on distinct contexts: one worker per context, results unchanged."""
    import torch
    la = litho_amd()
    sizes = [64, 96]
    ctxs = [la.Context(0) for _ in sizes]
    plans = [la.Plan.from_socs(c, socs(n), capacity=2) for c, n in zip(ctxs, sizes)]
    reqs = []
    for i in range(6):
        t = dev(targets().manhattan_target(sizes[i % 2], 70 + i))
        reqs.append((i % 2, t, torch.zeros_like(t)))
    refs = [alone(plans[pi], t, "gradient_descent", 5) for pi, t, _ in reqs]
    torch.cuda.synchronize()
    srv = la.Server(plans, iterations=5, threshold=0.225, alpha=50.0)
    tickets = [srv.submit(pi, t, out) for pi, t, out in reqs]
    results = [srv.wait(tk) for tk in tickets]
    srv.close()
    for (pi, t, out), r, (ref, ref_out) in zip(reqs, results, refs):
        assert r["final_pe"] == ref["final_pe"]
        assert torch.equal(out, ref_out)
    for p in plans:
        p.close()
    for c in ctxs:
        c.close()
