"""Pins of the oracle's partition, Micro gradient exchange and full-chip
CTM / LSM (no GPU).  Ports of test_micro_gradient_exchange.cpp,
test_full_chip_ctm.cpp and test_full_chip_lsm.cpp checks (PE modes)."""
import numpy as np
import pytest

from oracle import ctm as oc
from oracle import fullchip as fc
from oracle import imaging as im
from oracle import levelset as ls


def test_partition_geometry():
    meta = fc.build(1024, 1024, 4, 4, 128, 30)    # partition.h example: 16 Micros, tile 512
    assert meta.tile_size == 512 and len(meta.tiles) == 16
    t = meta.tiles[5]
    assert t.core_global == fc.Rect(256, 512, 256, 512)
    assert t.core_local == fc.Rect(128, 384, 128, 384)
    assert t.update_local == fc.Rect(98, 414, 98, 414)
    assert [nb.direction for nb in t.neighbors] == [fc.LEFT, fc.RIGHT, fc.TOP, fc.BOTTOM]
    # uneven split, odd tile rounded up to even
    m2 = fc.build(25, 23, 2, 3, 3, 2)
    assert m2.tile_size == 13 + 6 + 1
    assert m2.tiles[0].core_global == fc.Rect(0, 12, 0, 7)
    assert sum(t.core_global.rows() * t.core_global.cols() for t in m2.tiles) == 25 * 23


def test_extract_assemble_roundtrip_periodic():
    meta = fc.build(24, 24, 2, 2, 4, 3)
    g = np.arange(24 * 24, dtype=float).reshape(24, 24)
    locs = fc.extract_all(g, meta)
    assert np.array_equal(fc.assemble_global(locs, meta), g)
    t0 = meta.tiles[0]
    assert locs[0][0, 0] == g[(t0.tile_global.y0) % 24, (t0.tile_global.x0) % 24]   # periodic halo


def synthetic_case(meta):
    h, w = meta.height, meta.width
    y, x = np.mgrid[0:h, 0:w]
    mask = 0.45 + 0.15 * np.sin(0.31 * x + 0.17 * y)
    target = (((x // 4 + y // 3) % 3) == 0).astype(float)
    return mask, target


def test_exchange_slices_and_loss():
    meta = fc.build(24, 24, 2, 2, 4, 3)
    cache = im.Cache.synthetic(meta.tile_size)
    mask, target = synthetic_case(meta)
    res = fc.exchange_compute(meta, cache, fc.extract_all(mask, meta), target, fc.Options(), 0.25, 8.0)
    assert res.loss == sum(r.loss for r in res.locals)
    # Core-restricted loss and finite difference of tile 0's Core loss
    t = meta.tiles[0]
    lm = fc.extract_local_tile(mask, t)
    lt = fc.extract_local_tile(target, t)
    c = t.core_local

    def loss_at(m):
        _, w, _ = im.forward(cache, m, 0.25, 8.0)
        return float(((w[c.sl()] - lt[c.sl()]) ** 2).sum())

    assert abs(loss_at(lm) - res.locals[0].loss) < 1e-12
    g = res.locals[0].raw_mask_gradient
    d = g / np.linalg.norm(g)
    h = 1e-5
    num = (loss_at(lm + h * d) - loss_at(lm - h * d)) / (2 * h)
    assert abs(num - np.linalg.norm(g)) < 1e-6 * max(1.0, np.linalg.norm(g))
    # a Halo pixel keeps its gradient
    y, x = c.y0 - 1, c.x0 + 1
    p = np.zeros_like(lm)
    p[y, x] = 1.0
    num_p = (loss_at(lm + h * p) - loss_at(lm - h * p)) / (2 * h)
    assert abs(g[y, x]) > 1e-8 and abs(num_p - g[y, x]) < 1e-6 * max(1.0, abs(g[y, x]))


def options_case():
    meta = fc.build(24, 24, 2, 2, 4, 3)
    cache = im.Cache.synthetic(meta.tile_size)
    target = np.zeros((24, 24))
    target[6:18, 6:18] = 1.0
    y, x = np.mgrid[0:24, 0:24]
    mask = 0.35 + 0.1 * np.sin(0.29 * x) + 0.08 * np.cos(0.19 * y)
    masks = fc.extract_all(mask, meta)
    for i in range(len(masks)):
        masks[i] = masks[i] + 0.025 * (i + 1)
    return meta, cache, mask, masks, target


@pytest.mark.parametrize("loss_type", ["PE", "PEPenalty"])
def test_cept_joint_derivative(loss_type):
    # test_micro_gradient_exchange.cpp:229-257: two-sided CEPT gives the total-loss derivative
    meta, cache, mask, masks, target = options_case()
    o = fc.Options(loss_type=loss_type, enable_cept=True, cept_weight=0.7, pe_weight=0.6)
    consistent = fc.exchange_compute(meta, cache, fc.extract_all(mask, meta), target, o, 0.25, 8.0)
    assert consistent.cept_loss == 0.0
    res = fc.exchange_compute(meta, cache, masks, target, o, 0.25, 8.0)
    assert res.cept_loss > 0.0
    if loss_type == "PEPenalty":
        assert res.locals[0].bg_penalty_raw > 0.0
    for r in res.locals:
        assert abs(r.data_loss - (0.6 * r.pe_loss + r.penalty_loss)) < 1e-12
    norm = np.sqrt(sum(float((r.mask_gradient ** 2).sum()) for r in res.locals))
    h = 1e-5
    plus = [m + h * r.mask_gradient / norm for m, r in zip(masks, res.locals)]
    minus = [m - h * r.mask_gradient / norm for m, r in zip(masks, res.locals)]
    num = (fc.exchange_compute(meta, cache, plus, target, o, 0.25, 8.0).loss -
           fc.exchange_compute(meta, cache, minus, target, o, 0.25, 8.0).loss) / (2 * h)
    assert abs(num - norm) < 2e-5 * max(1.0, norm)


@pytest.mark.parametrize("mode", ["Average", "Linear"])
def test_fusion_weights(mode):
    # test_micro_gradient_exchange.cpp:350-401: fusion only changes Guard pixels, per-direction weights
    meta, cache, mask, masks, target = options_case()
    base = fc.exchange_compute(meta, cache, masks, target, fc.Options(), 0.25, 8.0)
    o = fc.Options(enable_grad_fusion=True, fusion_weight_owner=0.3, fusion_weight_neighbor=0.7, fusion_mode=mode)
    res = fc.exchange_compute(meta, cache, masks, target, o, 0.25, 8.0)
    assert res.loss == base.loss
    change = 0.0
    for t in meta.tiles:
        r = res.locals[t.id]
        assert np.array_equal(r.raw_mask_gradient, base.locals[t.id].raw_mask_gradient)
        c = t.core_local
        assert np.array_equal(r.mask_gradient[c.sl()], r.raw_mask_gradient[c.sl()])
        change += float(((r.mask_gradient - r.raw_mask_gradient) ** 2).sum())
        for nb in t.neighbors:
            g = nb.my_guard_local
            recv = res.locals[nb.id].raw_mask_gradient[nb.neighbor_overlap_local.sl()]
            for yy in range(g.rows()):
                for xx in range(g.cols()):
                    dist = {fc.LEFT: c.x0 - (g.x0 + xx), fc.RIGHT: g.x0 + xx - c.x1 + 1,
                            fc.TOP: c.y0 - (g.y0 + yy), fc.BOTTOM: g.y0 + yy - c.y1 + 1}[nb.direction]
                    s = 1.0 if mode == "Average" else (dist - 1) / (t.guard - 1)
                    nw = 0.7 * s
                    raw = r.raw_mask_gradient[g.y0 + yy, g.x0 + xx]
                    exp = (0.3 * raw + nw * recv[yy, xx]) / (0.3 + nw)
                    assert abs(r.mask_gradient[g.y0 + yy, g.x0 + xx] - exp) < 1e-12
    assert change > 1e-8


def test_full_chip_ctm_single_tile_equals_gd():
    # test_full_chip_ctm.cpp:50-97: a 1x1 partition without halo is the standalone GD trajectory
    n = 24
    meta = fc.build(n, n, 1, 1, 0, 0)
    cache = im.Cache.synthetic(meta.tile_size)
    target = np.zeros((n, n))
    target[6:18, 8:16] = 1.0
    res = fc.full_chip_ctm(meta, cache, target, target, fc.Options(), 5, 0.9, 0.25, 8.0)
    best, hist, theta = oc.gradient_descent(cache, target, 0.25, 8.0, 6, 0.9, 0.0)
    h = [x[1] for x in res["history"]]
    assert np.allclose(h, hist, rtol=1e-11, atol=0)


def test_full_chip_lsm_single_tile_matches_manual_rk2():
    n = 24
    meta = fc.build(n, n, 1, 1, 0, 0)
    cache = im.Cache.synthetic(meta.tile_size)
    target = np.zeros((n, n))
    target[6:18, 8:16] = 1.0
    lo = fc.LSMOptions(iterations=3, cfl=0.5, reinit_interval=2)
    res = fc.full_chip_lsm(meta, cache, target, target, fc.Options(), lo, 0.25, 8.0)
    # manual trajectory
    o = fc.Options(loss_region="CoreGuard", cept_guard_only=True, collect_core_metrics=True)
    phi = ls.reinit_sdf(target - 0.5, 1.0, 1.0)
    t = meta.tiles[0]
    losses = []
    for it in range(0, 4):
        mask = ls.render_binary(phi)
        ev = fc.exchange_compute(meta, cache, [mask], target, o, 0.25, 8.0)
        losses.append(ev.loss)
        if it == 3:
            break
        phi, _, _ = fc.lsm_update_tile(phi, ev.locals[0].mask_gradient, t, n, lo, it + 1)
    assert np.allclose([x[1] for x in res["history"]], losses, rtol=1e-10, atol=0)
    assert np.allclose(res["final_phi"][0], phi, rtol=0, atol=1e-11)


def test_mask_continuity_reference_numbers():
    """The reference's own known answers (tests/test_mask_continuity.cpp:13-39)."""
    meta = fc.build(12, 16, 2, 2, 2, 1)
    target = np.zeros((12, 16))
    target[:, 8:] = 1.0
    locals_ = fc.extract_all(target, meta)
    r = fc.evaluate_mask_continuity(meta, locals_, target)
    assert r["overlap_samples"] == 56 and r["seam_samples"] == 28
    assert r["overlap_sse"] == 0.0 and r["seam_reference_error_mae"] == 0.0
    assert abs(r["seam_jump_mae"] - 12.0 / 28.0) < 1e-14
    assert r["flat_seam_samples"] == 16 and r["foreground_seam_samples"] == 8
    target[:] = 0.0
    for t in meta.tiles:
        locals_[t.id][:] = t.col
    b = fc.evaluate_mask_continuity(meta, locals_, target)
    assert b["overlap_sse"] == 24.0 and b["overlap_max_abs"] == 1.0
    assert abs(b["overlap_rmse"] - np.sqrt(24.0 / 56.0)) < 1e-14
    assert abs(b["seam_reference_error_mae"] - 12.0 / 28.0) < 1e-14
    locals_ = fc.extract_all(target, meta)
    g = meta.tiles[0].neighbors[0].my_guard_local
    locals_[0][g.sl()] = 0.8
    c = fc.evaluate_mask_continuity(meta, locals_, target)
    assert abs(c["overlap_sse"] - 3.84) < 1e-14
    assert c["seam_jump_mae"] == 0.0
