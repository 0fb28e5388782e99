"""Device binarize_advanced (litho_binarize_advanced) against the oracle
restatement of Imbinarize::binarize_advanced (imbinarize.cpp:84-190): the
adaptive threshold equal in FP64 and all six masks bit-exact."""
import numpy as np
import pytest
from scipy import ndimage

from conftest import litho_amd, targets
from oracle import binarize as ob

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _gray(h, w, seed, sigma=2.5, plateau=False):
    rng = np.random.default_rng(seed)
    g = ndimage.gaussian_filter(rng.random((h, w)), sigma)
    g = (g - g.min()) / (g.max() - g.min())
    if plateau:                                   # saturated regions: plateaus of exact 1.0 and 0.0
        g = np.clip(1.4 * g - 0.2, 0.0, 1.0)
    return g.astype(np.float32)


def _run(ctx, gray, tgt, args):
    import torch
    h, w = gray.shape
    outs = {k: torch.full((h, w), 3.0, device="cuda") for k in
            ("final", "initial", "filtered", "sraf_initial", "sraf_final", "keep_out")}
    thr, npk = ctx.binarize_advanced(_dev(gray), _dev(tgt), outs["final"], *args, initial=outs["initial"],
                                     filtered=outs["filtered"], sraf_initial=outs["sraf_initial"],
                                     sraf_final=outs["sraf_final"], keep_out=outs["keep_out"])
    return thr, npk, {k: v.cpu().numpy() for k, v in outs.items()}


@pytest.mark.parametrize("h,w,seed,args,plateau", [
    (257, 300, 1, (0.1, 0.2, 50, 15), False),
    (96, 64, 2, (0.125, 0.1, 5, 5), True),
    (512, 512, 3, (0.3, 0.5, 20, 3), False),
    (2048, 2048, 4, (0.125, 0.1, 5, 5), True),
    (128, 128, 5, (0.1, 0.2, 5, 0), False),
])
def test_binarize_matches_oracle(gpu_ctx, h, w, seed, args, plateau):
    print(f"case binarize {h}x{w} seed {seed} args {args}", flush=True)
    gray = _gray(h, w, seed, plateau=plateau)
    tgt = targets().manhattan_target(w, seed, height=h)
    thr, npk, got = _run(gpu_ctx, gray, tgt, args)
    ref = ob.binarize_advanced(gray.astype(np.float64), tgt, *args)
    print(f"binarize {h}x{w}: peaks {npk} vs {ref['num_peaks']}, threshold {thr!r} vs {ref['threshold']!r}, "
          f"final pixels {int(got['final'].sum())}, sraf {int(got['sraf_final'].sum())}")
    assert npk == ref["num_peaks"]
    assert thr == ref["threshold"]
    for k in ("initial", "filtered", "sraf_initial", "sraf_final", "keep_out", "final"):
        bad = int(np.count_nonzero(got[k] != ref[k]))
        assert bad == 0, (k, bad)


def test_binarize_ctm_gray_mask(gpu_ctx):
    """The full-chip demo's hand-off (0.125, 0.1, 5, 5) on a real CTM gray mask."""
    import torch
    from oracle import optics as op
    la = litho_amd()
    n = 256
    sk = op.socs_kernels(n, 4.0, 193.0, 1.35, 0.6, 0.9, num_kernels=24)
    plan = la.Plan.from_socs(gpu_ctx, sk)
    tgt = targets().manhattan_target(n, 77)
    mask = torch.zeros((n, n), device="cuda")
    plan.ctm_optimize(_dev(tgt), mask, method="gradient_descent", iterations=15, learning_rate=0.9,
                      threshold=0.225, alpha=50.0)
    plan.close()
    gray = mask.cpu().numpy()
    thr, npk, got = _run(gpu_ctx, gray, tgt, (0.125, 0.1, 5, 5))
    ref = ob.binarize_advanced(gray.astype(np.float64), tgt, 0.125, 0.1, 5, 5)
    print(f"binarize CTM mask: peaks {npk}, threshold {thr:.6f}, final {int(got['final'].sum())} px "
          f"(target {int(tgt.sum())}), mismatches {int(np.count_nonzero(got['final'] != ref['final']))}")
    assert thr == ref["threshold"] and npk == ref["num_peaks"]
    for k in ("initial", "filtered", "sraf_initial", "sraf_final", "keep_out", "final"):
        assert np.array_equal(got[k], ref[k]), k


def test_binarize_no_peak_and_arguments(gpu_ctx):
    import torch
    la = litho_amd()
    gray = np.full((40, 50), 0.05, np.float32)
    tgt = targets().manhattan_target(50, 1, height=40)
    thr, npk, got = _run(gpu_ctx, gray, tgt, (0.1, 0.2, 5, 5))
    assert npk == 0 and thr == 0.0
    assert all(not v.any() for v in got.values())
    out = torch.zeros((40, 50), device="cuda")
    with pytest.raises(la.LithoError):
        gpu_ctx.binarize_advanced(_dev(gray), _dev(tgt), out, min_spacing_pixels=-1)
