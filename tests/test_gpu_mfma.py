"""The band-limited adjoint as a Toeplitz GEMM on the matrix cores
(corr_mfma.hip) against the FP64 oracle and against the length-L transform
adjoint, on the GPU.

The GEMM runs in bf16 pairs with three products per term (about 16
significant bits per operand, FP32 accumulation), so its results carry the
gradient tolerance, not FP32 rounding: gradient relative L2 <= 1e-4 against
the oracle and against the transform adjoint, PE relative <= 1e-5; the forward
image does not pass through it (aerial equal bit for bit between the two
adjoints).  Window sizes cover one and two waves per column and one to three
GEMM column tiles (K = 4, 8, 24).
"""
import numpy as np
import pytest

from conftest import litho_amd, targets
from oracle import ctm as oc
from oracle import imaging as im
from test_gpu_socs import dev, empty, host, rel_l2, window_kernels

pytestmark = pytest.mark.gpu


def run(gpu_ctx, monkeypatch, s, mask, tgt, adjoint, batch=1):
    la = litho_amd()
    monkeypatch.setenv("LITHO_COLUMN_ADJOINT", adjoint)
    plan = la.Plan.from_socs(gpu_ctx, s, capacity=batch)
    monkeypatch.delenv("LITHO_COLUMN_ADJOINT", raising=False)
    path = plan.column_path()
    n = s.n
    a, w, g = empty(n, batch), empty(n, batch), empty(n, batch)
    m = dev(np.broadcast_to(mask, (batch, n, n)))
    t = dev(np.broadcast_to(tgt, (batch, n, n)))
    plan.forward(m, a, w, threshold=0.225, alpha=50.0, batch=batch)
    pe, _ = plan.pe_gradient(m, t, grad_mask=g, threshold=0.225, alpha=50.0, batch=batch)
    plan.close()
    return path, host(a), host(g), list(pe)


@pytest.mark.parametrize("n,k", [(384, 24), (640, 8), (896, 24), (1152, 4), (2048, 24)])
def test_mfma_adjoint_matches_oracle_and_transforms(gpu_ctx, monkeypatch, n, k):
    s = window_kernels(n, k)
    tgt = targets().manhattan_target(n, 7)
    mask = im.mask_from_theta(oc.initialize_theta(tgt) + 0.3 * np.random.default_rng(n).standard_normal((n, n)))
    mask = mask.astype(np.float32)
    print(f"case mfma adjoint n={n} k={k}", flush=True)
    pm, am, gm, pem = run(gpu_ctx, monkeypatch, s, mask, tgt, "mfma")
    pf, af, gf, pef = run(gpu_ctx, monkeypatch, s, mask, tgt, "fft")
    assert pm["corr_len"] > 0 and pm["adjoint"] == "mfma" and pf["adjoint"] == "fft"
    cache = im.Cache.from_socs(s)
    _, rw, fields = im.forward(cache, mask.astype(np.float64), 0.225, 50.0)
    rg = im.pe_gradient(cache, rw, tgt.astype(np.float64), fields, 50.0)
    rpe = im.pe_loss(rw, tgt)
    e_or = rel_l2(gm[0], rg)
    e_ff = rel_l2(gm[0], gf[0])
    e_fft_or = rel_l2(gf[0], rg)
    print(f"n={n} k={k} window {s.kernels.shape[1]}: gradient rel L2 mfma-oracle {e_or:.2e}, "
          f"mfma-fft {e_ff:.2e} (fft-oracle {e_fft_or:.2e}); PE {pem[0]:.6f} vs {rpe:.6f}")
    assert np.array_equal(am, af)
    assert e_or <= 1e-4 and e_ff <= 1e-4
    assert abs(pem[0] - rpe) <= 1e-5 * max(rpe, 1.0)


def test_mfma_adjoint_batch_equals_single(gpu_ctx, monkeypatch):
    """Every column's GEMM reads only its own tile: a batch of 3 equal tiles
    gives each tile the single-tile gradient bit for bit."""
    n, k = 640, 24
    s = window_kernels(n, k)
    tgt = targets().manhattan_target(n, 11)
    mask = im.mask_from_theta(oc.initialize_theta(tgt)).astype(np.float32)
    print(f"case mfma adjoint batch n={n} k={k}", flush=True)
    _, _, g1, pe1 = run(gpu_ctx, monkeypatch, s, mask, tgt, "mfma", batch=1)
    _, _, g3, pe3 = run(gpu_ctx, monkeypatch, s, mask, tgt, "mfma", batch=3)
    for b in range(3):
        assert np.array_equal(g3[b], g1[0])
        assert pe3[b] == pe1[0]


def test_default_adjoint_choice(gpu_ctx):
    """Plans run the length-L transform adjoint unless LITHO_COLUMN_ADJOINT =
    mfma asks for the GEMM: its bf16 operands move fixed-step trajectories
    (the 512^2 GD run of test_gpu_ctm.py) past their tolerance."""
    la = litho_amd()
    for n in (640, 1152):
        s = window_kernels(n, 4)
        plan = la.Plan.from_socs(gpu_ctx, s)
        got = plan.column_path()
        plan.close()
        print(f"n={n} window {s.kernels.shape[1]}: {got}")
        assert got["corr_len"] > 0 and got["adjoint"] == "fft"
