"""Full-chip Micro tiling, gradient exchange, full-chip CTM and LSM, FP64 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates source/utils/partition.cpp, source/utils/micro_gradient_exchange.cpp
(PE / PEPenalty data losses, guard-only and two-sided CEPT, Average / Linear
guard fusion; the EPE modes are out of scope), source/optimizer/full_chip_ctm.cpp
and source/optimizer/full_chip_lsm.cpp.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import levelset as ls
from .ctm import initialize_theta
from .imaging import (Cache, background_penalty_loss, forward, pe_gradient,
                      pe_gradient_with_penalty, pe_loss)

LEFT, RIGHT, TOP, BOTTOM = 0, 1, 2, 3


@dataclass(frozen=True)
class Rect:
    """Half-open [y0, y1) x [x0, x1) (partition.h:10-24)."""
    y0: int
    y1: int
    x0: int
    x1: int

    def rows(self):
        return self.y1 - self.y0

    def cols(self):
        return self.x1 - self.x0

    def shifted(self, dy, dx):
        return Rect(self.y0 + dy, self.y1 + dy, self.x0 + dx, self.x1 + dx)

    def expanded(self, w):
        return Rect(self.y0 - w, self.y1 + w, self.x0 - w, self.x1 + w)

    @staticmethod
    def intersection(a, b):
        return Rect(max(a.y0, b.y0), min(a.y1, b.y1), max(a.x0, b.x0), min(a.x1, b.x1))

    def sl(self):
        return (slice(self.y0, self.y1), slice(self.x0, self.x1))


@dataclass
class Neighbor:
    id: int
    direction: int
    guard_global: Rect
    my_guard_local: Rect
    neighbor_overlap_local: Rect
    overlap_global: Rect
    my_overlap_local: Rect
    neighbor_guard_local: Rect


@dataclass
class Tile:
    id: int
    row: int
    col: int
    halo: int
    guard: int
    core_global: Rect
    tile_global: Rect
    core_local: Rect
    update_local: Rect
    neighbors: list = field(default_factory=list)


@dataclass
class Meta:
    height: int
    width: int
    grid_rows: int
    grid_cols: int
    halo: int
    guard: int
    tile_size: int
    tiles: list


def build(height, width, grid_rows, grid_cols, halo, guard) -> Meta:
    """Partition::build (partition.cpp:21-80)."""
    max_r = (height + grid_rows - 1) // grid_rows
    max_c = (width + grid_cols - 1) // grid_cols
    ts = max(max_r, max_c) + 2 * halo
    ts += ts % 2
    tiles = []
    for r in range(grid_rows):
        for c in range(grid_cols):
            core = Rect(r * height // grid_rows, (r + 1) * height // grid_rows,
                        c * width // grid_cols, (c + 1) * width // grid_cols)
            oy, ox = core.y0 - halo, core.x0 - halo
            tg = Rect(oy, oy + ts, ox, ox + ts)
            cl = core.shifted(-oy, -ox)
            tiles.append(Tile(r * grid_cols + c, r, c, halo, guard, core, tg, cl, cl.expanded(guard)))
    dr = (0, 0, -1, 1)
    dc = (-1, 1, 0, 0)
    for t in tiles:
        for i in range(4):
            r, c = t.row + dr[i], t.col + dc[i]
            if r < 0 or r >= grid_rows or c < 0 or c >= grid_cols:
                continue
            nb = tiles[r * grid_cols + c]
            gg = Rect.intersection(t.core_global.expanded(guard), nb.core_global)
            og = Rect.intersection(t.core_global, nb.core_global.expanded(guard))
            t.neighbors.append(Neighbor(nb.id, i, gg, gg.shifted(-t.tile_global.y0, -t.tile_global.x0),
                                        gg.shifted(-nb.tile_global.y0, -nb.tile_global.x0), og,
                                        og.shifted(-t.tile_global.y0, -t.tile_global.x0),
                                        og.shifted(-nb.tile_global.y0, -nb.tile_global.x0)))
    return Meta(height, width, grid_rows, grid_cols, halo, guard, ts, tiles)


def extract_local_tile(glob: np.ndarray, tile: Tile) -> np.ndarray:
    """Periodic halo gather ((g % n) + n) % n (partition.cpp:83-95)."""
    h, w = glob.shape
    ys = (np.arange(tile.tile_global.y0, tile.tile_global.y1)) % h
    xs = (np.arange(tile.tile_global.x0, tile.tile_global.x1)) % w
    return glob[ys][:, xs].copy()


def extract_all(glob, meta: Meta):
    return [extract_local_tile(glob, t) for t in meta.tiles]


def assemble_global(locals_, meta: Meta) -> np.ndarray:
    """Core scatter (partition.cpp:105-113)."""
    out = np.zeros((meta.height, meta.width))
    for t in meta.tiles:
        out[t.core_global.sl()] = locals_[t.id][t.core_local.sl()]
    return out


def evaluate_mask_continuity(meta: Meta, local_masks, target) -> dict:
    """evaluate_mask_continuity (mask_continuity.cpp:8-61): Guard-vs-neighbour
    Core differences of the local masks (every directed neighbour pair) and
    the jumps across the assembled Cores' right / bottom seams, split by the
    target's own change across the seam."""
    r = dict(overlap_samples=0, overlap_sse=0.0, overlap_rmse=0.0, overlap_mae=0.0, overlap_max_abs=0.0,
             seam_samples=0, flat_seam_samples=0, foreground_seam_samples=0, seam_jump_mae=0.0,
             seam_reference_error_mae=0.0, flat_seam_jump_mae=0.0, foreground_seam_jump_mae=0.0)
    g = assemble_global(local_masks, meta)
    t = np.asarray(target, dtype=np.float64)
    overlap_abs = 0.0

    def seam(ay, ax, by, bx):
        jump = g[by, bx] - g[ay, ax]
        ref = t[by, bx] - t[ay, ax]
        r["seam_samples"] += 1
        r["seam_jump_mae"] += abs(jump)
        r["seam_reference_error_mae"] += abs(jump - ref)
        if ref == 0.0:
            r["flat_seam_samples"] += 1
            r["flat_seam_jump_mae"] += abs(jump)
        if t[ay, ax] > 0.5 and t[by, bx] > 0.5:
            r["foreground_seam_samples"] += 1
            r["foreground_seam_jump_mae"] += abs(jump)

    for tile in meta.tiles:
        for e in tile.neighbors:
            a, b = e.my_guard_local, e.neighbor_overlap_local
            d = np.asarray(local_masks[tile.id], np.float64)[a.sl()] - np.asarray(local_masks[e.id], np.float64)[b.sl()]
            r["overlap_samples"] += d.size
            r["overlap_sse"] += float(np.sum(d * d))
            overlap_abs += float(np.sum(np.abs(d)))
            if d.size:
                r["overlap_max_abs"] = max(r["overlap_max_abs"], float(np.max(np.abs(d))))
            nb = meta.tiles[e.id]
            if e.direction == RIGHT:
                x = nb.core_global.x0
                for y in range(tile.core_global.y0, tile.core_global.y1):
                    seam(y, x - 1, y, x)
            elif e.direction == BOTTOM:
                y = nb.core_global.y0
                for x in range(tile.core_global.x0, tile.core_global.x1):
                    seam(y - 1, x, y, x)
    if r["overlap_samples"]:
        r["overlap_rmse"] = float(np.sqrt(r["overlap_sse"] / r["overlap_samples"]))
        r["overlap_mae"] = overlap_abs / r["overlap_samples"]
    if r["seam_samples"]:
        r["seam_jump_mae"] /= r["seam_samples"]
        r["seam_reference_error_mae"] /= r["seam_samples"]
    if r["flat_seam_samples"]:
        r["flat_seam_jump_mae"] /= r["flat_seam_samples"]
    if r["foreground_seam_samples"]:
        r["foreground_seam_jump_mae"] /= r["foreground_seam_samples"]
    return r


def update_window(tile: Tile, size: int) -> np.ndarray:
    """1 inside Core + Guard (partition.cpp:115-120)."""
    w = np.zeros((size, size))
    w[tile.update_local.sl()] = 1.0
    return w


# ----------------------------------------------------------------- exchange
@dataclass
class Options:
    """MicroGradientOptions (micro_gradient_exchange.h:19-40), PE modes only."""
    loss_type: str = "PE"            # "PE" | "PEPenalty"
    loss_region: str = "Core"        # "Core" | "CoreGuard"
    pe_weight: float = 1.0
    penalty_weight: float = 2.0
    penalty_threshold: float = 0.01
    enable_cept: bool = False
    cept_guard_only: bool = False
    cept_weight: float = 1.1
    enable_grad_fusion: bool = False
    fusion_mode: str = "Average"     # "Average" | "Linear"
    fusion_weight_owner: float = 0.5
    fusion_weight_neighbor: float = 0.5
    collect_core_metrics: bool = False


def loss_weights(o: Options):
    """micro_loss_weights (micro_gradient_exchange.cpp:95-99): (pe, penalty)."""
    return o.pe_weight, (o.penalty_weight if o.loss_type == "PEPenalty" else 0.0)


def loss_windows(local_target, region: Rect, o: Options):
    """build_micro_loss_windows, PE / PEPenalty (micro_gradient_exchange.cpp:65-93)."""
    h, w = local_target.shape
    pe = np.zeros((h, w))
    pen = np.zeros((h, w))
    t = local_target[region.sl()]
    pen[region.sl()] = (t < 0.5)
    pe[region.sl()] = (t > 0.5) if o.loss_type == "PEPenalty" else 1.0
    return pe, pen


def linear_strength(direction, row, col, rows, cols):
    """micro_gradient_exchange.cpp:52-61."""
    if direction == LEFT:
        return 1.0 if cols == 1 else 1.0 - col / (cols - 1)
    if direction == RIGHT:
        return 1.0 if cols == 1 else col / (cols - 1)
    if direction == TOP:
        return 1.0 if rows == 1 else 1.0 - row / (rows - 1)
    return 1.0 if rows == 1 else row / (rows - 1)


@dataclass
class LocalResult:
    loss: float = 0.0
    data_loss: float = 0.0
    pe_loss: float = 0.0
    penalty_loss: float = 0.0
    bg_penalty_raw: float = 0.0
    cept_loss: float = 0.0
    core_pe_loss: float = 0.0
    core_bg_penalty_raw: float = 0.0
    core_pe_penalty_loss: float = 0.0
    core_data_loss: float = 0.0
    raw_mask_gradient: np.ndarray = None
    cept_mask_gradient: np.ndarray = None
    mask_gradient: np.ndarray = None
    aerial: np.ndarray = None
    wafer: np.ndarray = None


@dataclass
class ExchangeResult:
    loss: float
    data_loss: float
    cept_loss: float
    locals: list


def exchange_compute(meta: Meta, cache: Cache, local_masks, global_target, o: Options,
                     threshold: float, alpha: float) -> ExchangeResult:
    """MicroGradientExchange::compute (micro_gradient_exchange.cpp:134-334)."""
    pe_w, pen_w = loss_weights(o)
    res = []
    for t in meta.tiles:                                                  # phase 1 :175-248
        lt = extract_local_tile(global_target, t)
        aerial, wafer, fields = forward(cache, local_masks[t.id], threshold, alpha)
        region = t.update_local if o.loss_region == "CoreGuard" else t.core_local
        pe_win, pen_win = loss_windows(lt, region, o)
        r = LocalResult()
        r.aerial, r.wafer = aerial, wafer
        r.pe_loss = pe_loss(wafer[region.sl()], lt[region.sl()], pe_win[region.sl()])
        if o.loss_type == "PEPenalty" or o.collect_core_metrics:
            r.bg_penalty_raw = background_penalty_loss(wafer[region.sl()], lt[region.sl()],
                                                       o.penalty_threshold, pen_win[region.sl()])
        r.penalty_loss = pen_w * r.bg_penalty_raw
        n = meta.tile_size
        r.raw_mask_gradient = np.zeros((n, n))
        r.cept_mask_gradient = np.zeros((n, n))
        if pen_w != 0.0:
            r.raw_mask_gradient = pe_gradient_with_penalty(cache, wafer, lt, fields, alpha, pen_w,
                                                           o.penalty_threshold, pe_win, pe_w, pen_win)
        elif pe_w != 0.0:
            r.raw_mask_gradient = pe_w * pe_gradient(cache, wafer, lt, fields, alpha, pe_win)
        r.data_loss = pe_w * r.pe_loss + r.penalty_loss
        if o.collect_core_metrics:
            c = t.core_local
            r.core_pe_loss = pe_loss(wafer[c.sl()], lt[c.sl()], pe_win[c.sl()])
            r.core_bg_penalty_raw = background_penalty_loss(wafer[c.sl()], lt[c.sl()],
                                                            o.penalty_threshold, pen_win[c.sl()])
            # micro_gradient_exchange.cpp:231-243
            main = (lt[c.sl()] > 0.5).astype(np.float64)
            r.core_pe_penalty_loss = (o.pe_weight * pe_loss(wafer[c.sl()], lt[c.sl()], main)
                                      + o.penalty_weight * r.core_bg_penalty_raw)
            r.core_data_loss = pe_w * r.core_pe_loss + pen_w * r.core_bg_penalty_raw
        r.loss = r.data_loss
        r.mask_gradient = r.raw_mask_gradient.copy()
        res.append(r)
    for t in meta.tiles:                                                  # phase 2 :253-326
        r = res[t.id]
        for nb in t.neighbors:
            ng = res[nb.id].raw_mask_gradient
            gs = nb.neighbor_overlap_local
            guard_grad = ng[gs.sl()]
            g = nb.my_guard_local
            if o.enable_cept:
                nmask = local_masks[nb.id]
                diff = local_masks[t.id][g.sl()] - nmask[gs.sl()]
                r.cept_loss += float((diff * diff).sum())
                r.cept_mask_gradient[g.sl()] += 2.0 * diff
                if not o.cept_guard_only:
                    ov = nb.my_overlap_local
                    os_ = nb.neighbor_guard_local
                    odiff = local_masks[t.id][ov.sl()] - nmask[os_.sl()]
                    r.cept_mask_gradient[ov.sl()] += 2.0 * odiff
            if o.enable_grad_fusion:
                rows, cols = g.rows(), g.cols()
                for y in range(rows):
                    for x in range(cols):
                        s = 1.0 if o.fusion_mode == "Average" else linear_strength(nb.direction, y, x, rows, cols)
                        nw = o.fusion_weight_neighbor * s
                        if nw == 0.0:
                            continue
                        lv = r.raw_mask_gradient[g.y0 + y, g.x0 + x]
                        r.mask_gradient[g.y0 + y, g.x0 + x] = (
                            (o.fusion_weight_owner * lv + nw * guard_grad[y, x]) / (o.fusion_weight_owner + nw))
        if o.enable_cept:
            r.mask_gradient = r.mask_gradient + o.cept_weight * r.cept_mask_gradient
            r.loss += o.cept_weight * r.cept_loss
    loss = data = cept = 0.0
    for r in res:                                                         # :327-332
        loss += r.loss
        data += r.data_loss
        cept += r.cept_loss
    return ExchangeResult(loss, data, cept, res)


# -------------------------------------------------------------- full-chip CTM
def full_chip_ctm(meta: Meta, cache: Cache, initial_pattern, global_target, o: Options,
                  iterations: int, lr: float, threshold: float, alpha: float):
    """FullChipCTMOptimizer::optimize (full_chip_ctm.cpp:25-152).

    Returns dict(best_iteration, best_loss, global_mask, local_masks, local_theta, history)."""
    o = Options(**{**o.__dict__, "cept_guard_only": True})
    theta = extract_all(initialize_theta(initial_pattern), meta)
    masks = [0.5 * (1.0 + np.cos(th)) for th in theta]
    ev = exchange_compute(meta, cache, masks, global_target, o, threshold, alpha)
    best = {"best_loss": np.inf}
    history = []

    def record(it):
        history.append((it, ev.loss, ev.data_loss, ev.cept_loss))
        if ev.loss < best["best_loss"]:
            best.update(best_iteration=it, best_loss=ev.loss,
                        local_masks=[m.copy() for m in masks], local_theta=[t.copy() for t in theta])

    record(0)
    for it in range(1, iterations + 1):
        for t in meta.tiles:                                              # :133-141
            w = t.update_local
            g = ev.locals[t.id].mask_gradient[w.sl()]
            th = theta[t.id]
            th[w.sl()] = th[w.sl()] - lr * (-0.5 * g * np.sin(th[w.sl()]))
            masks[t.id] = 0.5 * (1.0 + np.cos(th))
        ev = exchange_compute(meta, cache, masks, global_target, o, threshold, alpha)
        record(it)
    best["global_mask"] = assemble_global(best["local_masks"], meta)
    best["history"] = history
    best["final_local_masks"] = masks
    return best


# -------------------------------------------------------------- full-chip LSM
@dataclass
class LSMOptions:
    iterations: int = 20
    dx: float = 1.0
    dy: float = 1.0
    cfl: float = 0.7
    curvature_weight: float = 0.01
    reinit_interval: int = 20
    mask_type: str = "Binary"   # or "Gray" (full_chip_lsm.h:14-15)
    render_scale: int = 4


def evolution_term(phi, grad, lo: LSMOptions):
    """full_chip_lsm.cpp:26-32: Normal = kappa - delta, dt by CFL."""
    delta, h1, h2 = ls.evolve_normal_weno_godunov(phi, grad, lo.dy, lo.dx)
    kappa = ls.evolve_kappa(phi, lo.dx, lo.dy, lo.curvature_weight)
    dt = ls.get_dt_normal_kappa(lo.cfl, lo.dx, lo.dy, h1, h2, lo.curvature_weight)
    return kappa - delta, dt


def micro_imaging(meta: Meta, cache: Cache, global_mask, threshold: float, alpha: float):
    """MicroImaging::compute (mirco_imaging.cpp:9-27): every Micro imaged from
    its periodic extract of the global mask, the Cores assembled.
    Returns (global_aerial, global_wafer, local_aerial, local_wafer)."""
    la, lw = [], []
    for t in meta.tiles:
        a, w, _ = forward(cache, extract_local_tile(global_mask, t), threshold, alpha)
        la.append(a)
        lw.append(w)
    return assemble_global(la, meta), assemble_global(lw, meta), la, lw


def compare_micro_imaging(meta: Meta, cache: Cache, global_target, optimized_local_masks, threshold: float,
                          alpha: float) -> dict:
    """compare_micro_imaging (full_chip_imaging_compare.cpp:19-71): the
    target's Micros from periodic extracts, the optimized Micros as given,
    both Core-assembled; PE of each against the global target."""
    ta, tw, _, _ = micro_imaging(meta, cache, global_target, threshold, alpha)
    oa_l, ow_l = [], []
    for m in optimized_local_masks:
        a, w, _ = forward(cache, m, threshold, alpha)
        oa_l.append(a)
        ow_l.append(w)
    oa, ow = assemble_global(oa_l, meta), assemble_global(ow_l, meta)
    t = np.asarray(global_target, dtype=np.float64)
    return dict(optimized_mask=assemble_global(optimized_local_masks, meta), target_aerial=ta, target_wafer=tw,
                optimized_aerial=oa, optimized_wafer=ow, optimized_local_aerial=oa_l,
                target_pe=float(np.sum((tw - t) ** 2)), optimized_pe=float(np.sum((ow - t) ** 2)))


def lsm_update_tile(phi, mask_gradient, tile: Tile, size: int, lo: LSMOptions, iteration: int):
    """One RK2 update of one Micro (full_chip_lsm.cpp:215-241), then the
    render of lo.mask_type.  Returns (next_phi, next_mask, dt)."""
    w = tile.update_local
    grad = mask_gradient * update_window(tile, size)
    n1, dt = evolution_term(phi, grad, lo)
    phi1 = phi.copy()
    phi1[w.sl()] += dt * n1[w.sl()]
    n2, _ = evolution_term(phi1, grad, lo)
    nxt = phi.copy()
    nxt[w.sl()] = 0.5 * phi[w.sl()] + 0.5 * (phi1[w.sl()] + dt * n2[w.sl()])
    if lo.reinit_interval > 0 and iteration % lo.reinit_interval == 0:
        signs = (nxt >= 0.0).astype(np.float64) - 0.5
        sdf = ls.reinit_sdf(signs, lo.dx, lo.dy)
        nxt[w.sl()] = sdf[w.sl()]
    return nxt, ls.render_levelset_mask(nxt, lo.mask_type, lo.render_scale), dt


def full_chip_lsm(meta: Meta, cache: Cache, initial_binary, global_target, o: Options, lo: LSMOptions,
                  threshold: float, alpha: float):
    """FullChipLSMOptimizer::optimize (full_chip_lsm.cpp:39-255), render of lo.mask_type."""
    o = Options(**{**o.__dict__, "loss_region": "CoreGuard", "cept_guard_only": True,
                   "collect_core_metrics": True})
    phi, masks = [], []
    for t in meta.tiles:                                                  # :141-146
        m = extract_local_tile(initial_binary, t)
        p = ls.reinit_sdf(m - 0.5, lo.dx, lo.dy)
        phi.append(p)
        masks.append(ls.render_levelset_mask(p, lo.mask_type, lo.render_scale))
    ev = exchange_compute(meta, cache, masks, global_target, o, threshold, alpha)
    best = {"best_loss": np.inf}
    history = []

    def record(it, mn, mx):
        history.append((it, ev.loss, ev.data_loss, ev.cept_loss, mn, mx))
        if ev.loss < best["best_loss"]:
            best.update(best_iteration=it, best_loss=ev.loss, local_phi=[p.copy() for p in phi],
                        local_masks=[m.copy() for m in masks])

    record(0, 0.0, 0.0)
    for it in range(1, lo.iterations + 1):
        nphi, nmask, steps = [], [], []
        for t in meta.tiles:
            p2, m2, dt = lsm_update_tile(phi[t.id], ev.locals[t.id].mask_gradient, t, meta.tile_size, lo, it)
            nphi.append(p2)
            nmask.append(m2)
            steps.append(dt)
        phi, masks = nphi, nmask
        ev = exchange_compute(meta, cache, masks, global_target, o, threshold, alpha)
        record(it, min(steps), max(steps))
    best["global_phi"] = assemble_global(best["local_phi"], meta)
    best["global_mask"] = assemble_global(best["local_masks"], meta)
    best["history"] = history
    best["final_phi"] = phi
    return best
