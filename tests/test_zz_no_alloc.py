"""litho_ls_reinit_sdf_noalloc: reinit_sdf on the static per-GPU scratch, with
no device allocation, against litho_ls_reinit_sdf (bit for bit), over two
calls in a row (the second smaller, so it runs on what the first left in the
shared scratch), and its size limit."""
import numpy as np
import pytest

from conftest import litho_amd

pytestmark = pytest.mark.gpu


def _phi(b, n, m, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:n, 0:m]
    out = []
    for i in range(b):
        cy, cx, r = rng.uniform(0.3 * n, 0.7 * n), rng.uniform(0.3 * m, 0.7 * m), rng.uniform(0.15, 0.3) * min(n, m)
        phi = r - np.hypot(y - cy, x - cx) + 0.5 * np.sin(x / 3.0 + i) * np.cos(y / 5.0)
        phi[rng.random((n, m)) < 0.01] *= -1.0
        out.append(phi.astype(np.float32))
    return np.stack(out)


def test_noalloc_reinit_equals_allocating_entry_point():
    import torch
    la = litho_amd()
    ctx = la.Context(0)
    for b, n, m, seed in [(4, 640, 640, 1), (2, 96, 72, 2)]:
        print(f"case reinit_sdf_noalloc batch {b} {n}x{m}", flush=True)
        phi = torch.from_numpy(_phi(b, n, m, seed)).cuda()
        want = torch.zeros_like(phi)
        got = torch.full_like(phi, 7.0)
        ctx.reinit_sdf(phi, want, dx=0.5, dy=0.5)
        ctx.reinit_sdf_noalloc(phi, got, dx=0.5, dy=0.5)
        torch.cuda.synchronize()
        ctx.synchronize()
        bad = int(torch.count_nonzero(got != want).item())
        print(f"reinit_sdf_noalloc batch {b} {n}x{m}: {bad} mismatches", flush=True)
        assert bad == 0
    ctx.close()


def test_noalloc_rejects_sizes_above_the_scratch():
    import torch
    la = litho_amd()
    ctx = la.Context(0)
    phi = torch.zeros((5, 1024, 1024), dtype=torch.float32, device="cuda")   # 5 Mi pixels > 4 Mi
    with pytest.raises(la.LithoError, match="scratch"):
        ctx.reinit_sdf_noalloc(phi, phi)
    ctx.close()
