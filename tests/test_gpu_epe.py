"""Device EPE loss and gradient (litho_ep_set / litho_epe_gradient) against
the FP64 oracle restatement of Loss::evaluate_sparse_epe and
Gradient::epe_gradient, alone and in the PE_EPE combination.
Tolerances: losses relative <= 1e-4 (the sampled I - threshold is a
difference of FP32 images), gradient relative L2 <= 1e-4."""
import numpy as np
import pytest

from conftest import litho_amd, targets
from oracle import epe as oe
from oracle import imaging as im
from test_gpu_socs import dev, empty, host, rel_l2, socs

pytestmark = pytest.mark.gpu


def edge_points(tgt, seed, step=7):
    """Points on the boundaries of the target's rectangles (half a pixel
    outside the last inside pixel), every `step` pixels, random weights with
    some inactive points."""
    rng = np.random.default_rng(seed)
    pts = []
    h, w = tgt.shape
    for y in range(1, h - 1, step):
        row = tgt[y]
        for x in np.flatnonzero(np.diff(row) != 0):
            pts.append((y + rng.uniform(-0.3, 0.3), x + 0.5))
    for x in range(1, w - 1, step):
        col = tgt[:, x]
        for y in np.flatnonzero(np.diff(col) != 0):
            pts.append((y + 0.5, x + rng.uniform(-0.3, 0.3)))
    pts = np.array(pts, dtype=np.float64)
    wts = rng.uniform(0.5, 2.0, len(pts))
    wts[rng.random(len(pts)) < 0.1] = 0.0
    return pts, wts


@pytest.mark.parametrize("n,k,with_pe", [(128, 12, False), (128, 12, True), (512, 24, True)])
def test_epe_matches_oracle(gpu_ctx, n, k, with_pe):
    print(f"case epe n={n} k={k} pe={with_pe}", flush=True)
    la = litho_amd()
    s = socs(n, k)
    cache = im.Cache.from_socs(s)
    plan = la.Plan.from_socs(gpu_ctx, s)
    tgt = targets().manhattan_target(n, 600 + n)
    rng = np.random.default_rng(n)
    mask = np.clip(tgt + 0.2 * rng.standard_normal(tgt.shape), 0, 1).astype(np.float32)
    pts, wts = edge_points(tgt, n)
    eps = la.EpSet(plan, pts, wts)
    g = empty(n)
    a = empty(n)
    pe_w, epe_w = (0.7, 1.3) if with_pe else (1.0, 1.0)
    r = eps.gradient(dev(mask), target=dev(tgt) if with_pe else None, grad_mask=g[0], aerial=a[0], pe_weight=pe_w,
                     epe_weight=epe_w, threshold=0.225, alpha=50.0, pixel_size_nm=4.0)
    gref, loss, absnm, act, aerial, wafer, fields = oe.epe_gradient(cache, mask.astype(np.float64), pts, wts, 0.225,
                                                                     50.0, 4.0)
    gref = epe_w * gref
    if with_pe:
        gref = gref + pe_w * im.pe_gradient(cache, wafer, tgt.astype(np.float64), fields, 50.0)
        assert abs(r["pe_loss"] - im.pe_loss(wafer, tgt)) <= 1e-5 * im.pe_loss(wafer, tgt)
    eg = rel_l2(host(g)[0], gref)
    print(f"epe n={n}: {len(pts)} points, active {r['active_count']}/{act}; loss {r['epe_loss']:.6f} vs {loss:.6f}, "
          f"abs nm {r['epe_abs_sum_nm']:.4f} vs {absnm:.4f}; gradient rel L2 {eg:.2e}")
    assert r["active_count"] == act
    assert abs(r["epe_loss"] - loss) <= 1e-4 * loss
    assert abs(r["epe_abs_sum_nm"] - absnm) <= 1e-4 * absnm
    assert eg <= 1e-4
    # a second evaluation on the same set gives the same bits
    g2 = empty(n)
    r2 = eps.gradient(dev(mask), target=dev(tgt) if with_pe else None, grad_mask=g2[0], pe_weight=pe_w,
                      epe_weight=epe_w, threshold=0.225, alpha=50.0, pixel_size_nm=4.0)
    assert r2 == r and bool((g2 == g).all())
    eps.close()
    plan.close()


def test_epe_empty_set_and_arguments(gpu_ctx):
    la = litho_amd()
    s = socs(64, 8)
    plan = la.Plan.from_socs(gpu_ctx, s)
    tgt = targets().manhattan_target(64, 3)
    eps = la.EpSet(plan, np.zeros((0, 2)), np.zeros(0))
    g = empty(64)
    r = eps.gradient(dev(tgt), target=None, grad_mask=g[0])
    assert r["active_count"] == 0 and r["epe_loss"] == 0.0
    assert float(g.abs().max()) == 0.0
    with pytest.raises(la.LithoError):
        la.EpSet(plan, np.array([[1.0, np.nan]]), np.array([1.0]))
    with pytest.raises(la.LithoError):
        eps.gradient(dev(tgt), alpha=0.0)
    eps.close()
    plan.close()
