"""GPU parity of the full-chip Micro path against the FP64 oracle, through the C ABI.

Tolerances: periodic gather and Core scatter are exact; exchange losses
relative <= 1e-5 per Micro (FP32 fields, FP64 sums); mask gradients relative
L2 <= 1e-4 per Micro; full-chip histories relative <= 1e-4 (CTM) and the
binary LSM masks are compared pixel for pixel, with flips allowed only where
the oracle's phi lies within 1e-4 of the threshold.
"""
import numpy as np
import pytest

from conftest import litho_amd
from oracle import fullchip as fc
from oracle import levelset as ls
from oracle import imaging as im
from oracle import optics as op

pytestmark = pytest.mark.gpu

OPT = dict(pixel_nm=4.0, wavelength_nm=193.0, na=1.35, sigma_in=0.6, sigma_out=0.9)
THR, ALPHA = 0.225, 50.0
_SK = {}


def socs(n, k=8):
    if (n, k) not in _SK:
        _SK[(n, k)] = op.socs_kernels(n, **OPT, num_kernels=k)
    return _SK[(n, k)]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    return t.detach().cpu().double().numpy()


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def chip_target(h, w, seed):
    rng = np.random.default_rng(seed)
    t = np.zeros((h, w))
    for _ in range(h * w // 300):
        a, b = rng.integers(6, 16, size=2)
        y0, x0 = rng.integers(0, h - a), rng.integers(0, w - b)
        t[y0:y0 + a, x0:x0 + b] = 1.0
    return t


def hist_err(got, ref):
    """Largest error per history entry, scaled by max(|entry|, 1e-3 x the column's largest magnitude)."""
    scale = np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref), axis=0, keepdims=True))
    return float(np.max(np.abs(got - ref) / np.maximum(scale, 1e-12)))


def rel_l2(a, ref):
    d = np.linalg.norm(ref)
    return float(np.linalg.norm(a - ref) / d) if d > 0 else float(np.linalg.norm(a))


GEOM = (96, 96, 2, 2, 16, 8)     # chip 96 x 96, 2 x 2 Micros, halo 16, guard 8 -> tile 80


def setup(capacity):
    la = litho_amd()
    ts, tiles = la.partition_build(*GEOM)
    meta = fc.build(*GEOM)
    sk = socs(ts)
    ctx = la.Context(0)
    plan = la.Plan.from_socs(ctx, sk, capacity=capacity)
    return la, ctx, plan, ts, tiles, meta, im.Cache.from_socs(sk)


def test_extract_assemble_exact(gpu_ctx):
    import torch
    la = litho_amd()
    ts, tiles = la.partition_build(100, 90, 3, 2, 12, 6)
    meta = fc.build(100, 90, 3, 2, 12, 6)
    g = f32(np.random.default_rng(1).standard_normal((100, 90)))
    loc = torch.zeros((len(tiles), ts, ts), dtype=torch.float32, device="cuda")
    gpu_ctx.partition_extract(dev(g), tiles, ts, loc)
    ref = fc.extract_all(g, meta)
    assert np.array_equal(host(loc), np.stack(ref))
    out = torch.zeros((100, 90), dtype=torch.float32, device="cuda")
    gpu_ctx.partition_assemble(loc, tiles, ts, out)
    assert np.array_equal(host(out), fc.assemble_global(ref, meta))


@pytest.mark.parametrize("capacity", [4, 3])
def test_micro_imaging_matches_oracle(capacity):
    """MicroImaging::compute: the chunked batch of Micros (capacity 3 splits the
    four) against the oracle's per-Micro imaging; aerial <= 1e-5 relative L-inf
    (the forward tolerance), wafer <= 1e-4 absolute."""
    import torch
    la, ctx, plan, ts, tiles, meta, cache = setup(capacity)
    mask = f32(np.clip(chip_target(96, 96, 7) + 0.2 * np.random.default_rng(2).standard_normal((96, 96)), 0, 1))
    print(f"case micro_imaging capacity {capacity}", flush=True)
    ga, gw = torch.zeros((96, 96), dtype=torch.float32, device="cuda"), torch.zeros((96, 96), dtype=torch.float32,
                                                                                       device="cuda")
    loc = torch.zeros((len(tiles), ts, ts), dtype=torch.float32, device="cuda")
    plan.micro_imaging(tiles, dev(mask), ga, gw, local_aerial=loc, threshold=THR, alpha=ALPHA)
    ra, rw, rla, _ = fc.micro_imaging(meta, cache, mask, THR, ALPHA)
    ea = np.max(np.abs(host(ga) - ra)) / np.max(np.abs(ra))
    print(f"micro_imaging: aerial rel Linf {ea:.2e}, wafer abs {np.max(np.abs(host(gw) - rw)):.2e}")
    assert ea <= 1e-5
    assert np.max(np.abs(host(gw) - rw)) <= 1e-4
    assert np.max(np.abs(host(loc) - np.stack(rla))) <= 1e-5 * np.max(np.abs(ra))
    plan.close()
    ctx.close()


def test_compare_micro_imaging_matches_oracle():
    """compare_micro_imaging: target Micros from periodic extracts, optimized
    Micros as given (their Guards differ from the neighbours' Cores), both
    Core-assembled; images at the forward tolerances, PE within 1e-5."""
    import torch
    la, ctx, plan, ts, tiles, meta, cache = setup(3)
    target = chip_target(96, 96, 7)
    rng = np.random.default_rng(5)
    local = [f32(np.clip(m + 0.1 * rng.standard_normal(m.shape), 0, 1)) for m in fc.extract_all(target, meta)]
    print("case compare_micro_imaging", flush=True)
    z = lambda: torch.zeros((96, 96), dtype=torch.float32, device="cuda")
    om, ta, tw, oa, ow = z(), z(), z(), z(), z()
    loc = torch.zeros((len(tiles), ts, ts), dtype=torch.float32, device="cuda")
    tpe, ope = plan.compare_micro_imaging(tiles, dev(target), dev(np.stack(local)), threshold=THR, alpha=ALPHA,
                                          optimized_mask=om, target_aerial=ta, target_wafer=tw, optimized_aerial=oa,
                                          optimized_wafer=ow, optimized_local_aerial=loc)
    ref = fc.compare_micro_imaging(meta, cache, target, local, THR, ALPHA)
    print(f"compare_micro_imaging: PE target {tpe:.6f} vs {ref['target_pe']:.6f}, optimized {ope:.6f} vs "
          f"{ref['optimized_pe']:.6f}")
    assert abs(tpe - ref["target_pe"]) <= 1e-5 * ref["target_pe"]
    assert abs(ope - ref["optimized_pe"]) <= 1e-5 * ref["optimized_pe"]
    assert np.array_equal(host(om), ref["optimized_mask"].astype(np.float32).astype(np.float64))
    for got, want in [(ta, ref["target_aerial"]), (oa, ref["optimized_aerial"])]:
        assert np.max(np.abs(host(got) - want)) <= 1e-5 * np.max(np.abs(want))
    for got, want in [(tw, ref["target_wafer"]), (ow, ref["optimized_wafer"])]:
        assert np.max(np.abs(host(got) - want)) <= 1e-4
    assert np.max(np.abs(host(loc) - np.stack(ref["optimized_local_aerial"]))) <= 1e-5 * np.max(
        np.abs(ref["optimized_aerial"]))
    # without optional outputs the PE is the same
    tpe2, ope2 = plan.compare_micro_imaging(tiles, dev(target), dev(np.stack(local)), threshold=THR, alpha=ALPHA)
    assert (tpe2, ope2) == (tpe, ope)
    plan.close()
    ctx.close()


def test_mask_continuity_reference_numbers(gpu_ctx):
    """evaluate_mask_continuity on the device: the reference's known answers
    (tests/test_mask_continuity.cpp:13-39) and a random case against the oracle."""
    import torch
    la = litho_amd()
    ts, tiles = la.partition_build(12, 16, 2, 2, 2, 1)
    meta = fc.build(12, 16, 2, 2, 2, 1)
    print("case mask continuity", flush=True)
    target = np.zeros((12, 16))
    target[:, 8:] = 1.0
    locals_ = fc.extract_all(target, meta)
    r = gpu_ctx.mask_continuity(tiles, ts, dev(np.stack(locals_)), dev(target))
    assert r["overlap_samples"] == 56 and r["seam_samples"] == 28
    assert r["overlap_sse"] == 0.0 and r["seam_reference_error_mae"] == 0.0
    assert abs(r["seam_jump_mae"] - 12.0 / 28.0) < 1e-14
    assert r["flat_seam_samples"] == 16 and r["foreground_seam_samples"] == 8
    target[:] = 0.0
    for t in meta.tiles:
        locals_[t.id][:] = t.col
    b = gpu_ctx.mask_continuity(tiles, ts, dev(np.stack(locals_)), dev(target))
    assert b["overlap_sse"] == 24.0 and b["overlap_max_abs"] == 1.0
    assert abs(b["overlap_rmse"] - np.sqrt(24.0 / 56.0)) < 1e-14
    assert abs(b["seam_reference_error_mae"] - 12.0 / 28.0) < 1e-14
    locals_ = fc.extract_all(target, meta)
    g = meta.tiles[0].neighbors[0].my_guard_local
    locals_[0][g.sl()] = 0.8
    c = gpu_ctx.mask_continuity(tiles, ts, dev(np.stack(locals_)), dev(target))
    assert abs(c["overlap_sse"] - 6 * float(np.float32(0.8)) ** 2) < 1e-12 and c["seam_jump_mae"] == 0.0
    # random masks on a 3 x 2 partition with guard 6
    ts2, tiles2 = la.partition_build(100, 90, 3, 2, 12, 6)
    meta2 = fc.build(100, 90, 3, 2, 12, 6)
    rng = np.random.default_rng(4)
    tgt2 = (rng.random((100, 90)) < 0.4).astype(np.float64)
    loc2 = [f32(rng.random((ts2, ts2))) for _ in range(len(tiles2))]
    got = gpu_ctx.mask_continuity(tiles2, ts2, dev(np.stack(loc2)), dev(tgt2))
    ref = fc.evaluate_mask_continuity(meta2, loc2, tgt2)
    for k, v in ref.items():
        assert abs(got[k] - v) <= 1e-12 * max(1.0, abs(v)), (k, got[k], v)


CASES = {
    "pe_core": dict(),
    "penalty_coreguard_cept2_avg": dict(loss_type="PEPenalty", loss_region="CoreGuard", enable_cept=True,
                                        enable_grad_fusion=True, fusion_mode="Average"),
    "pe_cept_guard_linear_metrics": dict(enable_cept=True, cept_guard_only=True, enable_grad_fusion=True,
                                         fusion_mode="Linear", fusion_weight_owner=0.3, fusion_weight_neighbor=0.9,
                                         collect_core_metrics=True),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("capacity", [4, 3])
def test_tile_exchange_matches_oracle(case, capacity):
    import torch
    la, ctx, plan, ts, tiles, meta, cache = setup(capacity)
    kw = CASES[case]
    target = chip_target(96, 96, 7)
    rng = np.random.default_rng(11)
    glob_mask = f32(np.clip(target + 0.2 * rng.standard_normal(target.shape), 0, 1))
    # independent local masks: the Guard of each Micro differs from its neighbour's Core
    local_masks = [f32(np.clip(m + 0.05 * rng.standard_normal(m.shape), 0, 1)) for m in fc.extract_all(glob_mask, meta)]
    local_targets = fc.extract_all(target, meta)
    o = fc.Options(**{k: v for k, v in kw.items()})
    print(f"case exchange {case} capacity {capacity}", flush=True)
    ref = fc.exchange_compute(meta, cache, local_masks, target, o, THR, ALPHA)
    d_masks = dev(np.stack(local_masks))
    d_targets = dev(np.stack(local_targets))
    mg = torch.zeros_like(d_masks)
    raw = torch.zeros_like(d_masks)
    opts = la.exchange_options(threshold=THR, alpha=ALPHA, **kw)
    per, tot = plan.tile_exchange(tiles, d_masks, d_targets, opts, mg, raw)
    for name, a, b in [("loss", tot[0], ref.loss), ("data", tot[1], ref.data_loss), ("cept", tot[2], ref.cept_loss)]:
        assert abs(a - b) <= 1e-5 * max(abs(b), 1e-12), (name, a, b)
    g_mg, g_raw = host(mg), host(raw)
    worst = 0.0
    for t in range(len(tiles)):
        r = ref.locals[t]
        for name in ("loss", "data_loss", "pe_loss", "penalty_loss", "bg_penalty_raw", "cept_loss", "core_pe_loss",
                     "core_bg_penalty_raw", "core_pe_penalty_loss", "core_data_loss"):
            a, b = getattr(per[t], name), getattr(r, name)
            assert abs(a - b) <= 1e-5 * max(abs(b), 1e-9), (t, name, a, b)
        e1 = rel_l2(g_raw[t], r.raw_mask_gradient)
        e2 = rel_l2(g_mg[t], r.mask_gradient)
        worst = max(worst, e1, e2)
        assert e1 <= 1e-4 and e2 <= 1e-4, (t, e1, e2)
    print(f"exchange {case}: loss {tot[0]:.6f} vs {ref.loss:.6f}, worst gradient rel L2 {worst:.2e}")
    plan.close()
    ctx.close()


def test_fullchip_ctm_matches_oracle():
    import torch
    la, ctx, plan, ts, tiles, meta, cache = setup(4)
    target = chip_target(96, 96, 3)
    kw = dict(enable_cept=True, enable_grad_fusion=True)
    o = fc.Options(**kw)
    iters, lr = 4, 0.9
    print("case fullchip_ctm 2x2 micros, 4 iterations", flush=True)
    ref = fc.full_chip_ctm(meta, cache, target, target, o, iters, lr, THR, ALPHA)
    gm = torch.zeros((96, 96), dtype=torch.float32, device="cuda")
    loc = torch.zeros((len(tiles), ts, ts), dtype=torch.float32, device="cuda")
    r = plan.fullchip_ctm(tiles, dev(target), dev(target), la.exchange_options(threshold=THR, alpha=ALPHA, **kw),
                          iters, lr, gm, loc)
    hist_ref = np.array([h[1:] for h in ref["history"]])
    err = hist_err(r["history"], hist_ref)
    print(f"fullchip_ctm: best {r['best_iteration']} {r['best_loss']:.6f} vs {ref['best_iteration']} "
          f"{ref['best_loss']:.6f}; history rel err {err:.2e}")
    assert r["best_iteration"] == ref["best_iteration"]
    assert err <= 1e-4
    assert np.max(np.abs(host(gm) - ref["global_mask"])) <= 1e-4
    assert np.max(np.abs(host(loc) - np.stack(ref["local_masks"]))) <= 1e-4
    # the history rows' breakdown (full_chip_ctm.cpp:84-104): data = PE + penalty contributions,
    # loss = data + weighted CEPT; and they reach history.csv
    for row in r["rows"]:
        assert row.data_loss == pytest.approx(row.pe_contribution + row.penalty_contribution, rel=1e-12)
        assert row.loss == pytest.approx(row.data_loss + row.cept_contribution, rel=1e-12)
        assert row.pe_loss_raw == pytest.approx(row.pe_contribution, rel=1e-12)     # pe_weight 1
    import tempfile, csv, os
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "history.csv")
        la.write_history_csv(path, "CTM", list(r["rows"]))
        rec = list(csv.DictReader(open(path)))
        assert len(rec) == iters + 1 and float(rec[-1]["loss"]) == r["rows"][-1].loss
    plan.close()
    ctx.close()


@pytest.mark.parametrize("mask_type", ["Binary", "Gray"])
def test_fullchip_lsm_matches_oracle(mask_type):
    import torch
    la, ctx, plan, ts, tiles, meta, cache = setup(4)
    target = chip_target(96, 96, 5)
    kw = dict(loss_type="PEPenalty", enable_cept=True, enable_grad_fusion=True)
    o = fc.Options(**kw)
    lo = fc.LSMOptions(iterations=4, dx=1.0, dy=1.0, cfl=0.5, curvature_weight=0.01, reinit_interval=2,
                       mask_type=mask_type, render_scale=4)
    print(f"case fullchip_lsm 2x2 micros, 4 iterations, reinit every 2, {mask_type} render", flush=True)
    ref = fc.full_chip_lsm(meta, cache, target, target, o, lo, THR, ALPHA)
    gm = torch.zeros((96, 96), dtype=torch.float32, device="cuda")
    gp = torch.zeros_like(gm)
    r = plan.fullchip_lsm(tiles, dev(target), dev(target), la.exchange_options(threshold=THR, alpha=ALPHA, **kw),
                          lo.iterations, gm, gp, dx=lo.dx, dy=lo.dy, cfl=lo.cfl,
                          curvature_weight=lo.curvature_weight, reinit_interval=lo.reinit_interval,
                          mask_type=mask_type, render_scale=lo.render_scale)
    hist_ref = np.array([h[1:] for h in ref["history"]])
    err = hist_err(r["history"], hist_ref)
    if mask_type == "Binary":
        near = np.abs(ref["global_phi"]) < 1e-4
    else:   # a pixel whose coverage can change: some subpixel of its interpolated phi within 1e-4 of zero
        g = ref["global_phi"]
        near = (np.abs(ls.upscale_linear(g, 4)) < 1e-4).reshape(g.shape[0], 4, g.shape[1], 4).any(axis=(1, 3))
    mism = int(np.count_nonzero((host(gm) != ref["global_mask"]) & ~near))
    print(f"fullchip_lsm: best {r['best_iteration']} {r['best_loss']:.6f} vs {ref['best_iteration']} "
          f"{ref['best_loss']:.6f}; history rel err {err:.2e}; mask mismatches {mism}")
    assert r["best_iteration"] == ref["best_iteration"]
    assert err <= 1e-4
    assert mism == 0
    assert np.max(np.abs(host(gp) - ref["global_phi"])) <= 1e-4 * np.max(np.abs(ref["global_phi"]))
    # rows: PEPenalty contributions and the Core metrics the LSM collects (full_chip_lsm.cpp:152-182)
    for row in r["rows"]:
        assert row.data_loss == pytest.approx(row.pe_contribution + row.penalty_contribution, rel=1e-12)
        assert row.loss == pytest.approx(row.data_loss + row.cept_contribution, rel=1e-12)
        assert row.core_pe > 0.0 and row.core_data_loss > 0.0 and row.core_pe_penalty > 0.0
    plan.close()
    ctx.close()


def test_exchange_rejects_mismatched_tile_size():
    import torch
    la = litho_amd()
    ts, tiles = la.partition_build(*GEOM)
    sk = socs(64)
    ctx = la.Context(0)
    plan = la.Plan.from_socs(ctx, sk, capacity=4)
    x = torch.zeros((4, ts, ts), dtype=torch.float32, device="cuda")
    with pytest.raises(la.LithoError):
        plan.tile_exchange(tiles, x, x, la.exchange_options(), x)
    plan.close()
    ctx.close()
