"""The window-pruned row and column transforms against the unpruned kernels of
the same tile size, on the GPU.

Setting LITHO_ROW_PATH / LITHO_COLUMN_PATH to "generic" when a plan is
created selects the unpruned kernels (full transforms with zero-filled
spectra); LITHO_COLUMN_PATH = "fused" selects the fused column kernel with
pruned length-N transforms instead of the default band-limited column path
(length-L correlations).  Each path is checked against the FP64 oracle in
test_gpu_socs.py; here the combinations must agree with each other to FP32
rounding: aerial image relative L-inf <= 1e-5, gradient relative L2 <= 1e-5,
PE relative <= 1e-6.
"""
import numpy as np
import pytest

from conftest import litho_amd, targets
from oracle import ctm as oc
from oracle import imaging as im
from test_gpu_socs import dev, empty, host, rel_l2, rel_linf, window_kernels

pytestmark = pytest.mark.gpu

PATHS = [("pruned", "default"), ("generic", "default"), ("pruned", "fused"), ("pruned", "generic"),
         ("generic", "generic")]


def run(gpu_ctx, monkeypatch, s, mask, tgt, rows, cols):
    la = litho_amd()
    n = s.n
    # the transform adjoint on every path: the comparison is at FP32 rounding
    # (the matrix-core adjoint has its own test, test_gpu_mfma.py)
    monkeypatch.setenv("LITHO_COLUMN_ADJOINT", "fft")
    for var, val in (("LITHO_ROW_PATH", rows), ("LITHO_COLUMN_PATH", cols)):
        if val in ("generic", "fused"):
            monkeypatch.setenv(var, val)
        else:
            monkeypatch.delenv(var, raising=False)
    plan = la.Plan.from_socs(gpu_ctx, s)
    monkeypatch.delenv("LITHO_ROW_PATH", raising=False)
    monkeypatch.delenv("LITHO_COLUMN_PATH", raising=False)
    monkeypatch.delenv("LITHO_COLUMN_ADJOINT", raising=False)
    a, w, g = empty(n), empty(n), empty(n)
    plan.forward(dev(mask), a, w, threshold=0.225, alpha=50.0)
    pe, _ = plan.pe_gradient(dev(mask), dev(tgt), grad_mask=g, threshold=0.225, alpha=50.0)
    plan.close()
    return host(a)[0], host(g)[0], pe[0]


@pytest.mark.parametrize("n,k", [(384, 6), (640, 6), (1024, 6), (1152, 6), (2048, 4)])
def test_pruned_transforms_equal_unpruned(gpu_ctx, monkeypatch, n, k):
    s = window_kernels(n, k)
    tgt = targets().manhattan_target(n, 3)
    mask = im.mask_from_theta(oc.initialize_theta(tgt) + 0.3 * np.random.default_rng(n).standard_normal((n, n)))
    mask = mask.astype(np.float32)
    out = {}
    for rows, cols in PATHS:
        print(f"case paths n={n} k={k} rows={rows} columns={cols}", flush=True)
        out[(rows, cols)] = run(gpu_ctx, monkeypatch, s, mask, tgt, rows, cols)
    a0, g0, pe0 = out[("pruned", "default")]
    assert np.max(a0) > 0.05 and np.linalg.norm(g0) > 0.0
    for key, (a, g, pe) in out.items():
        ea, eg = rel_linf(a, a0), rel_l2(g, g0)
        print(f"paths n={n} {key}: aerial rel Linf {ea:.2e}, gradient rel L2 {eg:.2e}, PE {pe:.6f} vs {pe0:.6f}")
        assert ea <= 1e-5
        assert eg <= 1e-5
        assert abs(pe - pe0) <= 1e-6 * max(abs(pe0), 1.0)
