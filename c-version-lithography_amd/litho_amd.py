"""ctypes binding of liblitho_amd.so (the C ABI declared in include/litho_amd.h).

The library is the product; this module only declares its entry points and
offers thin helpers that pass torch tensors' device pointers through.  It never
falls back to anything: a missing library or a failing call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liblitho_amd.so"

OK, EINVAL, ENOMEM, EDEVICE, EUNSUPPORTED, ERUNTIME, ETIMEDOUT = 0, 1, 2, 3, 4, 5, 6


_P_ = C.c_void_p


class LithoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[litho status {code}] {msg}")
        self.code = code


class CtmOptions(C.Structure):
    _fields_ = [("method", C.c_int), ("iterations", C.c_int), ("learning_rate", C.c_double),
                ("lbfgs_history", C.c_int), ("gradient_tolerance", C.c_double),
                ("threshold", C.c_double), ("alpha", C.c_double)]


_MASK_TYPES = {"Binary": 0, "Gray": 1}


class LsmOptions(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("cfl", C.c_double), ("curvature_weight", C.c_double),
                ("mask_type", C.c_int), ("render_scale", C.c_int)]

    def __init__(self, dx=1.0, dy=1.0, cfl=0.7, curvature_weight=0.01, mask_type=0, render_scale=4):
        super().__init__(dx, dy, cfl, curvature_weight, mask_type, render_scale)


class CtmInfo(C.Structure):
    _fields_ = [("evaluations", C.c_int), ("iterations", C.c_int), ("final_pe", C.c_double),
                ("final_grad_norm", C.c_double), ("status", C.c_int)]


class ServerOptions(C.Structure):
    _fields_ = [("ctm", CtmOptions), ("memory_budget_bytes", C.c_double)]


class ServerResult(C.Structure):
    _fields_ = [("info", CtmInfo), ("queued_ms", C.c_double), ("service_ms", C.c_double)]


class Rect(C.Structure):
    _fields_ = [("y0", C.c_int), ("y1", C.c_int), ("x0", C.c_int), ("x1", C.c_int)]

    def tuple(self):
        return (self.y0, self.y1, self.x0, self.x1)


class Neighbor(C.Structure):
    _fields_ = [("id", C.c_int), ("direction", C.c_int), ("guard_global", Rect), ("my_guard_local", Rect),
                ("neighbor_overlap_local", Rect), ("overlap_global", Rect), ("my_overlap_local", Rect),
                ("neighbor_guard_local", Rect)]


class Tile(C.Structure):
    _fields_ = [("id", C.c_int), ("row", C.c_int), ("col", C.c_int), ("halo", C.c_int), ("guard", C.c_int),
                ("core_global", Rect), ("tile_global", Rect), ("core_local", Rect), ("update_local", Rect),
                ("num_neighbors", C.c_int), ("neighbors", Neighbor * 4)]


class ExchangeOptions(C.Structure):
    """litho_exchange_options; defaults follow MicroGradientOptions."""
    _fields_ = [("loss_type", C.c_int), ("loss_region", C.c_int), ("pe_weight", C.c_double),
                ("penalty_weight", C.c_double), ("penalty_threshold", C.c_double), ("collect_core_metrics", C.c_int),
                ("enable_cept", C.c_int), ("cept_guard_only", C.c_int), ("cept_weight", C.c_double),
                ("enable_grad_fusion", C.c_int), ("fusion_mode", C.c_int), ("fusion_weight_owner", C.c_double),
                ("fusion_weight_neighbor", C.c_double), ("threshold", C.c_double), ("alpha", C.c_double)]


def exchange_options(loss_type="PE", loss_region="Core", pe_weight=1.0, penalty_weight=2.0, penalty_threshold=0.01,
                     collect_core_metrics=False, enable_cept=False, cept_guard_only=False, cept_weight=1.1,
                     enable_grad_fusion=False, fusion_mode="Average", fusion_weight_owner=0.5,
                     fusion_weight_neighbor=0.5, threshold=0.25, alpha=50.0) -> ExchangeOptions:
    return ExchangeOptions({"PE": 0, "PEPenalty": 1}[loss_type], {"Core": 0, "CoreGuard": 1}[loss_region],
                           pe_weight, penalty_weight, penalty_threshold, int(collect_core_metrics), int(enable_cept),
                           int(cept_guard_only), cept_weight, int(enable_grad_fusion),
                           {"Average": 0, "Linear": 1}[fusion_mode], fusion_weight_owner, fusion_weight_neighbor,
                           threshold, alpha)


class TileLoss(C.Structure):
    _fields_ = [("loss", C.c_double), ("data_loss", C.c_double), ("pe_loss", C.c_double),
                ("penalty_loss", C.c_double), ("bg_penalty_raw", C.c_double), ("cept_loss", C.c_double),
                ("core_pe_loss", C.c_double), ("core_bg_penalty_raw", C.c_double),
                ("core_pe_penalty_loss", C.c_double), ("core_data_loss", C.c_double)]


class BinarizeInfo(C.Structure):
    _fields_ = [("threshold", C.c_double), ("num_peaks", C.c_longlong)]


class MaskContinuity(C.Structure):
    _fields_ = [("overlap_samples", C.c_longlong), ("overlap_sse", C.c_double), ("overlap_rmse", C.c_double),
                ("overlap_mae", C.c_double), ("overlap_max_abs", C.c_double), ("seam_samples", C.c_longlong),
                ("flat_seam_samples", C.c_longlong), ("foreground_seam_samples", C.c_longlong),
                ("seam_jump_mae", C.c_double), ("seam_reference_error_mae", C.c_double),
                ("flat_seam_jump_mae", C.c_double), ("foreground_seam_jump_mae", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EpeResult(C.Structure):
    _fields_ = [("pe_loss", C.c_double), ("epe_loss", C.c_double), ("epe_abs_sum_nm", C.c_double),
                ("active_count", C.c_int)]


class FullChipRow(C.Structure):
    _fields_ = [("iteration", C.c_int), ("loss", C.c_double), ("data_loss", C.c_double), ("cept_loss", C.c_double),
                ("min_dt", C.c_double), ("max_dt", C.c_double), ("pe_contribution", C.c_double),
                ("penalty_contribution", C.c_double), ("epe_contribution", C.c_double),
                ("cept_contribution", C.c_double), ("pe_loss_raw", C.c_double), ("bg_penalty_raw", C.c_double),
                ("epe_loss_raw", C.c_double), ("epe_abs_sum_nm", C.c_double), ("epe_point_count", C.c_int),
                ("core_pe", C.c_double), ("core_bg_penalty_raw", C.c_double), ("core_pe_penalty", C.c_double),
                ("core_epe_loss", C.c_double), ("core_epe_abs_sum_nm", C.c_double), ("core_ep_count", C.c_int),
                ("core_epe_mean", C.c_double), ("core_data_loss", C.c_double), ("epe_mean", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchTile(C.Structure):
    _fields_ = [("plan_index", C.c_int), ("target", _P_), ("theta0", _P_), ("out_mask", _P_), ("out_theta", _P_)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes); the declaration list of include/litho_amd.h
SIGNATURES = {
    "litho_last_error": (C.c_char_p, []),
    "litho_supported_size": (_I, [_I]),
    "litho_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "litho_ctx_destroy": (_I, [_P]),
    "litho_ctx_synchronize": (_I, [_P]),
    "litho_ctx_stream": (_P, [_P]),
    "litho_ctx_set_graphs": (_I, [_P, _I]),
    "litho_ctx_time_pair_kernel": (_I, [_P, _I]),
    "litho_ctx_pair_kernel_time": (_I, [_P, C.POINTER(C.c_longlong), _DP]),
    "litho_ctx_graph_counts": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "litho_socs_window": (_I, [_I, _D, _D, _D, _D, _D, _IP, _IP, _IP, _IP, _IP, _DP]),
    "litho_socs_last_eigen_gap": (_D, []),
    "litho_socs_precompute": (_I, [_P, _I, _D, _D, _D, _D, _D, _I, _DP, _DP, _DP]),
    "litho_socs_precompute_zernike": (_I, [_P, _I, _D, _D, _D, _D, _D, _IP, _DP, _I, _I, _DP, _DP, _DP]),
    "litho_plan_create": (_I, [_P, _I, _I, _I, _I, _I, _I, _DP, _DP, _D, _I, C.POINTER(_P)]),
    "litho_plan_destroy": (_I, [_P]),
    "litho_plan_device_bytes": (C.c_size_t, [_P]),
    "litho_plan_column_path": (C.c_int, [_P, _IP, _IP, _IP]),
    "litho_socs_forward": (_I, [_P, _I, _P, _P, _P, _D, _D]),
    "litho_pe_gradient": (_I, [_P, _I, _P, _P, _P, _P, _D, _D, _D, _D, _D, _P, _P, _P, _DP, _DP]),
    "litho_ctm_eval": (_I, [_P, _I, _P, _P, _D, _D, _P, _P, _DP, _DP]),
    "litho_ctm_optimize": (_I, [_P, _P, _P, C.POINTER(CtmOptions), _P, _P, _DP, _I, C.POINTER(CtmInfo)]),
    "litho_batch_ctm": (_I, [C.POINTER(_P), _I, C.POINTER(BatchTile), _I, C.POINTER(CtmOptions), _I,
                             C.POINTER(CtmInfo), _IP]),
    "litho_server_create": (_I, [C.POINTER(_P), _I, C.POINTER(ServerOptions), C.POINTER(_P)]),
    "litho_server_destroy": (_I, [_P]),
    "litho_server_submit": (_I, [_P, _I, _P, _P, _P, _P, C.POINTER(C.c_longlong)]),
    "litho_server_wait": (_I, [_P, C.c_longlong, _D, C.POINTER(ServerResult)]),
    "litho_server_request_bytes": (C.c_size_t, [_P, _I]),
    "litho_initialize_theta": (_I, [_P, _P, _P, C.c_size_t]),
    "litho_profile_ctm_eval": (_I, [_P, _I, _P, _P, _D, _D, _I, _DP]),
    "litho_profile_lbfgs_pair": (_I, [_P, C.c_size_t, _I, _I, _DP]),
    "litho_profile_stream": (_I, [_P, C.c_size_t, _I, _I, _I, _DP]),
    "litho_lbfgs_direction": (_I, [_P, C.c_size_t, _I, _I, _P, _P, _P, _DP, _IP]),
    "litho_ls_terms": (_I, [_P, _I, _I, _I, _P, _P, _D, _D, _D, _P, _P]),
    "litho_ls_dt": (_D, [_D, _D, _D, _D, _D, _D]),
    "litho_ls_reinit_sdf": (_I, [_P, _I, _I, _I, _P, _D, _D, _P]),
    "litho_ls_reinit_sdf_noalloc": (_I, [_P, _I, _I, _I, _P, _D, _D, _P]),
    "litho_compare_micro_imaging": (_I, [_P, C.POINTER(Tile), _I, _I, _I, _P, _P, _D, _D, _P, _P, _P, _P, _P, _P,
                                         _DP, _DP]),
    "litho_evaluate_mask_continuity": (_I, [_P, C.POINTER(Tile), _I, _I, _I, _I, _P, _P, C.POINTER(MaskContinuity)]),
    "litho_ep_set_create": (_I, [_P, _I, _DP, _DP, C.POINTER(_P)]),
    "litho_ep_set_destroy": (_I, [_P]),
    "litho_epe_gradient": (_I, [_P, _P, _P, _P, _D, _D, _D, _D, _D, _P, _P, _P, C.POINTER(EpeResult)]),
    "litho_binarize_advanced": (_I, [_P, _I, _I, _P, _P, _D, _D, _I, _I, _P, _P, _P, _P, _P, _P,
                                     C.POINTER(BinarizeInfo)]),
    "litho_render_levelset_mask": (_I, [_P, _I, _I, _I, _P, _I, _I, _P]),
    "litho_micro_imaging": (_I, [_P, C.POINTER(Tile), _I, _I, _I, _P, _D, _D, _P, _P, _P, _P]),
    "litho_lsm_step": (_I, [_P, _I, _I, _P, _P, _IP, C.POINTER(LsmOptions), _I, _P, _P, _DP]),
    "litho_partition_build": (_I, [_I, _I, _I, _I, _I, _I, _IP, C.POINTER(Tile)]),
    "litho_partition_extract": (_I, [_P, _P, _I, _I, C.POINTER(Tile), _I, _I, _P]),
    "litho_partition_assemble": (_I, [_P, _P, C.POINTER(Tile), _I, _I, _I, _I, _P]),
    "litho_tile_exchange": (_I, [_P, C.POINTER(Tile), _I, _P, _P, C.POINTER(ExchangeOptions), _P, _P, _P,
                                 C.POINTER(TileLoss), _DP]),
    "litho_fullchip_ctm": (_I, [_P, C.POINTER(Tile), _I, _I, _I, _P, _P, C.POINTER(ExchangeOptions), _I, _D, _P, _P,
                                C.POINTER(FullChipRow), _IP, _DP]),
    "litho_fullchip_lsm": (_I, [_P, C.POINTER(Tile), _I, _I, _I, _P, _P, C.POINTER(ExchangeOptions),
                                C.POINTER(LsmOptions), _I, _I, _P, _P, C.POINTER(FullChipRow), _IP, _DP]),
    "litho_save_txt": (_I, [C.c_char_p, _DP, _I, _I, _I]),
    "litho_save_txt_int": (_I, [C.c_char_p, _DP, _I, _I]),
    "litho_load_txt": (_I, [C.c_char_p, _DP, C.c_size_t, _IP, _IP]),
    "litho_save_matrix_txt": (_I, [C.c_char_p, _DP, _I, _I]),
    "litho_write_history_csv": (_I, [C.c_char_p, _I, C.POINTER(FullChipRow), _I]),
    "litho_write_summary_csv": (_I, [C.c_char_p, _I, _I, _D, _I, _D, _D]),
    "litho_write_imaging_metrics_csv": (_I, [C.c_char_p, _I, _D, _D]),
    "litho_write_sampling_csv": (_I, [C.c_char_p, _I, _I, _I, _I, _D, _I, _I, _I, _I]),
}

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load the native library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != OK:
        msg = load().litho_last_error().decode(errors="replace")
        raise LithoError(code, msg)


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def _dp(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_DP)


class Context:
    def __init__(self, device: int = 0):
        self.lib = load()
        h = _P()
        check(self.lib.litho_ctx_create(device, C.byref(h)))
        self.h = h
        self._plans = weakref.WeakSet()   # plans made on this context: closed before it

    def close(self):
        if self.h:
            for p in list(self._plans):
                p.close()
            check(self.lib.litho_ctx_destroy(self.h))
            self.h = None

    def synchronize(self):
        check(self.lib.litho_ctx_synchronize(self.h))

    def set_graphs(self, enable: bool):
        check(self.lib.litho_ctx_set_graphs(self.h, 1 if enable else 0))

    def time_pair_kernel(self, enable: bool):
        """Start (or stop) HIP-event timing of the L-BFGS pair kernel in this context's solves; resets totals."""
        check(self.lib.litho_ctx_time_pair_kernel(self.h, 1 if enable else 0))

    def pair_kernel_time(self):
        """(launches, total ms) of the pair-kernel launches timed since time_pair_kernel(True)."""
        n = C.c_longlong(0)
        ms = C.c_double(0.0)
        check(self.lib.litho_ctx_pair_kernel_time(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def graph_counts(self):
        """Evaluation graphs built on this context: dict(captured, failed, direct)."""
        v = [C.c_longlong(0) for _ in range(3)]
        check(self.lib.litho_ctx_graph_counts(self.h, *[C.byref(x) for x in v]))
        return dict(captured=v[0].value, failed=v[1].value, direct=v[2].value)

    def profile_lbfgs_pair(self, n, m=10, repeats=20):
        """HIP-event time (ms) of one L-BFGS pair-kernel launch over n floats with m pairs."""
        ms = C.c_double(0.0)
        check(self.lib.litho_profile_lbfgs_pair(self.h, n, m, repeats, C.byref(ms)))
        return ms.value

    def profile_stream(self, n, reads, writes, repeats=20):
        """HIP-event time (ms) of a plain kernel reading `reads` and writing `writes` vectors of n floats."""
        ms = C.c_double(0.0)
        check(self.lib.litho_profile_stream(self.h, n, reads, writes, repeats, C.byref(ms)))
        return ms.value

    def lbfgs_direction(self, x_points, g_points, d_out, m=10):
        """litho_lbfgs_direction: x_points, g_points (P, n) float32 device tensors; d_out (n,).
        Returns (g . d, pairs kept)."""
        p, n = x_points.shape
        gd = C.c_double(0.0)
        kept = C.c_int(0)
        check(self.lib.litho_lbfgs_direction(self.h, n, m, p, ptr(x_points), ptr(g_points), ptr(d_out),
                                             C.byref(gd), C.byref(kept)))
        return gd.value, kept.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def socs_precompute(self, n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out, k, singular_values=False,
                        aberrations=None):
        """Device SOCS precompute; returns dict like oracle SocsKernels fields.  aberrations:
        {fringe Zernike index: coefficient in waves} (pupil wavefront), or None."""
        win = socs_window(n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out)
        kern = np.zeros((k, win["rows"], win["cols"], 2), dtype=np.float64)
        w = np.zeros(k, dtype=np.float64)
        sv = np.zeros(win["num_points"], dtype=np.float64) if singular_values else None
        if aberrations:
            idx = np.array(list(aberrations.keys()), dtype=np.int32)
            cf = np.array(list(aberrations.values()), dtype=np.float64)
            check(self.lib.litho_socs_precompute_zernike(self.h, n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out,
                                                         idx.ctypes.data_as(_IP), _dp(cf), len(idx), k, _dp(kern),
                                                         _dp(w), _dp(sv)))
        else:
            check(self.lib.litho_socs_precompute(self.h, n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out, k,
                                                 _dp(kern), _dp(w), _dp(sv)))
        out = dict(win)
        out.update(kernels=kern[..., 0] + 1j * kern[..., 1], weights=w, singular_values=sv,
                   eigen_gap=self.lib.litho_socs_last_eigen_gap())
        return out


    # ------------------------------------------------------------ level set
    def ls_terms(self, phi, vn=None, delta=None, kappa=None, dx=1.0, dy=1.0, curvature_weight=0.01):
        """evolve_normal_WENO_godunov delta and evolve_kappa on (batch, rows, cols) float32 tensors."""
        b, r, c = _shape3(phi)
        check(self.lib.litho_ls_terms(self.h, b, r, c, ptr(phi), ptr(vn), dx, dy, curvature_weight, ptr(delta),
                                      ptr(kappa)))

    def reinit_sdf(self, phi, sdf, dx=1.0, dy=1.0):
        b, r, c = _shape3(phi)
        check(self.lib.litho_ls_reinit_sdf(self.h, b, r, c, ptr(phi), dx, dy, ptr(sdf)))

    def reinit_sdf_noalloc(self, phi, sdf, dx=1.0, dy=1.0):
        """litho_ls_reinit_sdf_noalloc: reinit_sdf on the static per-GPU scratch (no device allocation)."""
        b, r, c = _shape3(phi)
        check(self.lib.litho_ls_reinit_sdf_noalloc(self.h, b, r, c, ptr(phi), dx, dy, ptr(sdf)))

    def binarize_advanced(self, gray, target, final_mask, peak_min_intensity=0.1, quantile_val=0.2,
                          area_threshold=50, min_spacing_pixels=15, initial=None, filtered=None, sraf_initial=None,
                          sraf_final=None, keep_out=None):
        """Imbinarize::binarize_advanced on device (H, W) float32 tensors; returns (threshold, num_peaks)."""
        h, w = gray.shape
        info = BinarizeInfo()
        check(self.lib.litho_binarize_advanced(self.h, h, w, ptr(gray), ptr(target), peak_min_intensity,
                                               quantile_val, area_threshold, min_spacing_pixels, ptr(final_mask),
                                               ptr(initial), ptr(filtered), ptr(sraf_initial), ptr(sraf_final),
                                               ptr(keep_out), C.byref(info)))
        return info.threshold, info.num_peaks

    def render_levelset_mask(self, phi, mask_out, mask_type="Binary", scale=4):
        """render_levelset_mask of (batch, rows, cols) tiles; mask_type "Binary" or "Gray"."""
        b, r, c = _shape3(phi)
        check(self.lib.litho_render_levelset_mask(self.h, b, r, c, ptr(phi), _MASK_TYPES[mask_type], scale,
                                                  ptr(mask_out)))

    def lsm_step(self, phi, grad, rects, phi_out, mask_out=None, dx=1.0, dy=1.0, cfl=0.7, curvature_weight=0.01,
                 reinit=False, mask_type="Binary", render_scale=4):
        """One RK2 update of (batch, n, n) tiles; rects: (batch, 4) y0, y1, x0, x1.  Returns dt per tile."""
        b, r, c = _shape3(phi)
        if r != c:
            raise ValueError("lsm_step tiles are square")
        rc = np.ascontiguousarray(rects, dtype=np.int32).reshape(b, 4)
        dt = np.zeros(b)
        o = LsmOptions(dx, dy, cfl, curvature_weight, _MASK_TYPES[mask_type], render_scale)
        check(self.lib.litho_lsm_step(self.h, b, r, ptr(phi), ptr(grad), rc.ctypes.data_as(_IP), C.byref(o),
                                      1 if reinit else 0, ptr(phi_out), ptr(mask_out), _dp(dt)))
        return dt


    # ------------------------------------------------------------ partition
    def partition_extract(self, global_, tiles, tile_size, out):
        h, w = global_.shape
        check(self.lib.litho_partition_extract(self.h, ptr(global_), h, w, tiles, len(tiles), tile_size, ptr(out)))

    def mask_continuity(self, tiles, tile_size, local_masks, target):
        """evaluate_mask_continuity of device Micros against the device global target; returns a dict."""
        h, w = target.shape
        m = MaskContinuity()
        check(self.lib.litho_evaluate_mask_continuity(self.h, tiles, len(tiles), tile_size, h, w, ptr(local_masks),
                                                      ptr(target), C.byref(m)))
        return m.as_dict()

    def partition_assemble(self, local, tiles, tile_size, out):
        h, w = out.shape
        check(self.lib.litho_partition_assemble(self.h, ptr(local), tiles, len(tiles), tile_size, h, w, ptr(out)))


def batch_ctm(plans, tiles, method="lbfgs", iterations=30, learning_rate=0.9, history=10, tol=0.0,
              threshold=0.25, alpha=50.0, chunk=0):
    """litho_batch_ctm.  tiles: list of (plan_index, target, out_mask[, theta0[, out_theta]]) torch tensors.
    Returns (list of info dicts, chunk used per plan)."""
    lib = load()
    hs = (_P * len(plans))(*[p.h for p in plans])
    arr = (BatchTile * len(tiles))()
    for i, t in enumerate(tiles):
        pi, tg, om = t[0], t[1], t[2]
        t0 = t[3] if len(t) > 3 else None
        ot = t[4] if len(t) > 4 else None
        arr[i] = BatchTile(pi, ptr(tg), ptr(t0), ptr(om), ptr(ot))
    o = CtmOptions({"gradient_descent": 0, "lbfgs": 1}[method], iterations, learning_rate, history, tol,
                   threshold, alpha)
    info = (CtmInfo * len(tiles))()
    used = (C.c_int * len(plans))()
    check(lib.litho_batch_ctm(hs, len(plans), arr, len(tiles), C.byref(o), chunk, info, used))
    out = [dict(evaluations=r.evaluations, iterations=r.iterations, final_pe=r.final_pe,
                final_grad_norm=r.final_grad_norm, status=r.status) for r in info]
    return out, list(used)


class Server:
    """litho_server: continuous-batching CTM over several plans (tile sizes)."""

    def __init__(self, plans, method="gradient_descent", iterations=20, learning_rate=0.9, history=10, tol=0.0,
                 threshold=0.25, alpha=50.0, memory_budget=64.0 * 2**30):
        self.lib = load()
        self.plans = list(plans)
        hs = (_P * len(plans))(*[p.h for p in plans])
        o = ServerOptions(CtmOptions({"gradient_descent": 0, "lbfgs": 1}[method], iterations, learning_rate,
                                     history, tol, threshold, alpha), memory_budget)
        h = _P()
        check(self.lib.litho_server_create(hs, len(plans), C.byref(o), C.byref(h)))
        self.h = h

    def submit(self, plan_index, target, out_mask, theta0=None, out_theta=None) -> int:
        t = C.c_longlong()
        check(self.lib.litho_server_submit(self.h, plan_index, ptr(target), ptr(theta0), ptr(out_mask),
                                           ptr(out_theta), C.byref(t)))
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """dict(evaluations, ..., queued_ms, service_ms), or None on timeout."""
        r = ServerResult()
        code = self.lib.litho_server_wait(self.h, ticket, timeout, C.byref(r))
        if code == ETIMEDOUT:
            return None
        check(code)
        i = r.info
        return dict(evaluations=i.evaluations, iterations=i.iterations, final_pe=i.final_pe,
                    final_grad_norm=i.final_grad_norm, status=i.status, queued_ms=r.queued_ms,
                    service_ms=r.service_ms)

    def request_bytes(self, plan_index) -> int:
        return int(self.lib.litho_server_request_bytes(self.h, plan_index))

    def close(self):
        if self.h:
            check(self.lib.litho_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def partition_build(height, width, grid_rows, grid_cols, halo, guard):
    """Partition::build through the library: (tile_size, ctypes array of Tile)."""
    lib = load()
    tiles = (Tile * (grid_rows * grid_cols))()
    ts = C.c_int()
    check(lib.litho_partition_build(height, width, grid_rows, grid_cols, halo, guard, C.byref(ts), tiles))
    return ts.value, tiles


def _shape3(t):
    if t.dim() == 2:
        return 1, t.shape[0], t.shape[1]
    if t.dim() == 3:
        return t.shape[0], t.shape[1], t.shape[2]
    raise ValueError("expected (rows, cols) or (batch, rows, cols)")


def ls_dt(cfl, dx, dy, max_h1, max_h2, curvature_weight) -> float:
    return float(load().litho_ls_dt(cfl, dx, dy, max_h1, max_h2, curvature_weight))


def socs_window(n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out) -> dict:
    lib = load()
    r, c, y0, x0, npts = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    ws = C.c_double()
    check(lib.litho_socs_window(n, pixel_nm, wavelength_nm, na, sigma_in, sigma_out, C.byref(r), C.byref(c),
                                C.byref(y0), C.byref(x0), C.byref(npts), C.byref(ws)))
    return dict(rows=r.value, cols=c.value, y0=y0.value, x0=x0.value, num_points=npts.value,
                weight_sum=ws.value)


class Plan:
    """litho_plan: SOCS kernels + workspace for `capacity` tiles of size n."""

    def __init__(self, ctx: Context, n: int, kernels: np.ndarray, weights, weight_sum: float,
                 win_y0: int, win_x0: int, capacity: int = 1):
        self.ctx = ctx
        self.lib = ctx.lib
        kern = np.ascontiguousarray(kernels, dtype=np.complex128)
        if kern.ndim != 3:
            raise ValueError("kernels must be (K, Wr, Wc)")
        k, wr, wc = kern.shape
        kd = np.ascontiguousarray(np.stack([kern.real, kern.imag], axis=-1))
        w = np.ascontiguousarray(weights, dtype=np.float64)
        h = _P()
        check(self.lib.litho_plan_create(ctx.h, n, k, wr, wc, win_y0, win_x0, _dp(kd), _dp(w),
                                         float(weight_sum), capacity, C.byref(h)))
        self.h = h
        ctx._plans.add(self)
        self.n, self.k, self.capacity = n, k, capacity

    @classmethod
    def from_socs(cls, ctx, sk, capacity=1):
        """From an oracle-style object with n, kernels, weights, weight_sum, win_y0, win_x0."""
        return cls(ctx, sk.n, sk.kernels, sk.weights, sk.weight_sum, sk.win_y0, sk.win_x0, capacity)

    def close(self):
        if self.h:
            check(self.lib.litho_plan_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self) -> int:
        return int(self.lib.litho_plan_device_bytes(self.h))

    def column_path(self) -> dict:
        """litho_plan_column_path: dict(transforms="direct_dft" | "fft", corr_len=L or 0,
        adjoint="mfma" | "fft")."""
        g, L, m = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.litho_plan_column_path(self.h, C.byref(g), C.byref(L), C.byref(m)))
        return dict(transforms="direct_dft" if g.value else "fft", corr_len=L.value,
                    adjoint="mfma" if m.value else "fft")

    def forward(self, mask, aerial=None, wafer=None, threshold=0.25, alpha=50.0, batch=1):
        check(self.lib.litho_socs_forward(self.h, batch, ptr(mask), ptr(aerial), ptr(wafer), threshold, alpha))

    # ------------------------------------------------------------ full chip
    def tile_exchange(self, tiles, masks, targets, options, mask_gradient, raw_gradient=None, wafer=None):
        """MicroGradientExchange::compute on device Micros; returns (per-tile TileLoss list, totals)."""
        per = (TileLoss * len(tiles))()
        tot = np.zeros(3)
        check(self.lib.litho_tile_exchange(self.h, tiles, len(tiles), ptr(masks), ptr(targets), C.byref(options),
                                           ptr(mask_gradient), ptr(raw_gradient), ptr(wafer), per, _dp(tot)))
        return list(per), tot

    def compare_micro_imaging(self, tiles, global_target, optimized_local_masks, threshold=0.25, alpha=50.0,
                              optimized_mask=None, target_aerial=None, target_wafer=None, optimized_aerial=None,
                              optimized_wafer=None, optimized_local_aerial=None):
        """compare_micro_imaging; returns (target_pe, optimized_pe)."""
        h, w = global_target.shape
        tpe, ope = C.c_double(0.0), C.c_double(0.0)
        check(self.lib.litho_compare_micro_imaging(self.h, tiles, len(tiles), h, w, ptr(global_target),
                                                   ptr(optimized_local_masks), threshold, alpha, ptr(optimized_mask),
                                                   ptr(target_aerial), ptr(target_wafer), ptr(optimized_aerial),
                                                   ptr(optimized_wafer), ptr(optimized_local_aerial), C.byref(tpe),
                                                   C.byref(ope)))
        return tpe.value, ope.value

    def fullchip_ctm(self, tiles, initial_pattern, target, options, iterations, learning_rate, global_mask_out,
                     local_masks_out=None):
        h, w = target.shape
        rows = (FullChipRow * (iterations + 1))()
        bi, bl = C.c_int(), C.c_double()
        check(self.lib.litho_fullchip_ctm(self.h, tiles, len(tiles), h, w, ptr(initial_pattern), ptr(target),
                                          C.byref(options), iterations, learning_rate, ptr(global_mask_out),
                                          ptr(local_masks_out), rows, C.byref(bi), C.byref(bl)))
        hist = np.array([[r.loss, r.data_loss, r.cept_loss] for r in rows])
        return dict(best_iteration=bi.value, best_loss=bl.value, history=hist, rows=rows)

    def micro_imaging(self, tiles, global_mask, global_aerial=None, global_wafer=None, local_aerial=None,
                      local_wafer=None, threshold=0.25, alpha=50.0):
        """MicroImaging::compute over the Micros `tiles` of a (height, width) device mask."""
        h, w = global_mask.shape
        check(self.lib.litho_micro_imaging(self.h, tiles, len(tiles), h, w, ptr(global_mask), threshold, alpha,
                                           ptr(global_aerial), ptr(global_wafer), ptr(local_aerial),
                                           ptr(local_wafer)))

    def fullchip_lsm(self, tiles, initial_binary, target, options, iterations, global_mask_out, global_phi_out=None,
                     dx=1.0, dy=1.0, cfl=0.7, curvature_weight=0.01, reinit_interval=20, mask_type="Binary",
                     render_scale=4):
        h, w = target.shape
        rows = (FullChipRow * (iterations + 1))()
        bi, bl = C.c_int(), C.c_double()
        lo = LsmOptions(dx, dy, cfl, curvature_weight, _MASK_TYPES[mask_type], render_scale)
        check(self.lib.litho_fullchip_lsm(self.h, tiles, len(tiles), h, w, ptr(initial_binary), ptr(target),
                                          C.byref(options), C.byref(lo), iterations, reinit_interval,
                                          ptr(global_mask_out), ptr(global_phi_out), rows, C.byref(bi),
                                          C.byref(bl)))
        hist = np.array([[r.loss, r.data_loss, r.cept_loss, r.min_dt, r.max_dt] for r in rows])
        return dict(best_iteration=bi.value, best_loss=bl.value, history=hist, rows=rows)

    def pe_gradient(self, mask, target, grad_mask=None, aerial=None, wafer=None, pe_window=None,
                    penalty_window=None, pe_weight=1.0, penalty=0.0, penalty_threshold=0.01,
                    threshold=0.25, alpha=50.0, batch=1):
        pe = np.zeros(batch)
        bg = np.zeros(batch)
        check(self.lib.litho_pe_gradient(self.h, batch, ptr(mask), ptr(target), ptr(pe_window),
                                         ptr(penalty_window), pe_weight, penalty, penalty_threshold,
                                         threshold, alpha, ptr(aerial), ptr(wafer), ptr(grad_mask),
                                         _dp(pe), _dp(bg)))
        return pe, bg

    def ctm_eval(self, theta, target, mask=None, grad_theta=None, threshold=0.25, alpha=50.0, batch=1):
        pe = np.zeros(batch)
        gn = np.zeros(batch)
        check(self.lib.litho_ctm_eval(self.h, batch, ptr(theta), ptr(target), threshold, alpha, ptr(mask),
                                      ptr(grad_theta), _dp(pe), _dp(gn)))
        return pe, gn

    def profile_eval(self, theta, target, threshold=0.25, alpha=50.0, repeats=10, batch=1):
        """HIP-event stage times (ms): forward half, column image, column adjoint
        (the correlation adjoint kernel alone), adjoint tail, total."""
        out = np.zeros(5)
        check(self.lib.litho_profile_ctm_eval(self.h, batch, ptr(theta), ptr(target), threshold, alpha, repeats,
                                              _dp(out)))
        return out

    def ctm_optimize(self, target, out_mask, theta0=None, out_theta=None, method="lbfgs", iterations=50,
                     learning_rate=0.9, history=10, tol=0.0, threshold=0.25, alpha=50.0,
                     history_capacity=4096):
        o = CtmOptions(1 if method == "lbfgs" else 0, iterations, learning_rate, history, tol, threshold, alpha)
        info = CtmInfo()
        hist = np.zeros(history_capacity)
        check(self.lib.litho_ctm_optimize(self.h, ptr(target), ptr(theta0), C.byref(o), ptr(out_mask),
                                          ptr(out_theta), _dp(hist), history_capacity, C.byref(info)))
        return dict(evaluations=info.evaluations, iterations=info.iterations, final_pe=info.final_pe,
                    final_grad_norm=info.final_grad_norm, status=info.status,
                    history=hist[: min(info.evaluations, history_capacity)].copy())


def supported_size(n: int) -> bool:
    return bool(load().litho_supported_size(n))


class EpSet:
    """litho_ep_set: edge points (count, 2) (y, x) and weights of one plan's tile size."""

    def __init__(self, plan, points_yx, weights):
        self.lib = plan.lib
        self.plan = plan
        p = np.ascontiguousarray(points_yx, dtype=np.float64).reshape(-1, 2)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        h = _P()
        check(self.lib.litho_ep_set_create(plan.h, len(w), _dp(p), _dp(w), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(self.lib.litho_ep_set_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gradient(self, mask, target=None, grad_mask=None, aerial=None, wafer=None, pe_window=None, pe_weight=1.0,
                 epe_weight=1.0, threshold=0.25, alpha=50.0, pixel_size_nm=4.0):
        r = EpeResult()
        check(self.lib.litho_epe_gradient(self.h, ptr(mask), ptr(target), ptr(pe_window), pe_weight, epe_weight,
                                          threshold, alpha, pixel_size_nm, ptr(grad_mask), ptr(aerial), ptr(wafer),
                                          C.byref(r)))
        return dict(pe_loss=r.pe_loss, epe_loss=r.epe_loss, epe_abs_sum_nm=r.epe_abs_sum_nm,
                    active_count=r.active_count)


# ------------------------------------------------------------------ on-disk formats
def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def save_txt(path, matrix, precision=4):
    """SaveTxt::save_mat of a 2-D array."""
    m = np.ascontiguousarray(matrix, dtype=np.float64)
    check(load().litho_save_txt(_path(path), _dp(m), m.shape[0], m.shape[1], precision))


def save_txt_int(path, matrix):
    m = np.ascontiguousarray(matrix, dtype=np.float64)
    check(load().litho_save_txt_int(_path(path), _dp(m), m.shape[0], m.shape[1]))


def load_txt(path) -> np.ndarray:
    """SaveTxt::load_txt: the matrix of a TXT file (rows x cols float64)."""
    lib = load()
    r, c = C.c_int(), C.c_int()
    check(lib.litho_load_txt(_path(path), None, 0, C.byref(r), C.byref(c)))
    out = np.zeros((r.value, c.value))
    check(lib.litho_load_txt(_path(path), _dp(out), out.size, C.byref(r), C.byref(c)))
    return out


def save_matrix_txt(path, matrix):
    """The full-chip demos' "# rows= cols=" matrix TXT (precision 17, aligned)."""
    m = np.ascontiguousarray(matrix, dtype=np.float64)
    check(load().litho_save_matrix_txt(_path(path), _dp(m), m.shape[0], m.shape[1]))


def write_history_csv(path, stage, rows):
    """history.csv of the full-chip CTM ("CTM") or LSM ("LSM") demo from FullChipRow rows."""
    arr = (FullChipRow * len(rows))(*rows)
    check(load().litho_write_history_csv(_path(path), {"CTM": 0, "LSM": 1}[stage], arr, len(rows)))


def write_summary_csv(path, stage, best_iteration, best_loss, final_iteration, final_loss, seconds=0.0):
    check(load().litho_write_summary_csv(_path(path), {"CTM": 0, "LSM": 1}[stage], best_iteration, best_loss,
                                         final_iteration, final_loss, seconds))


def write_imaging_metrics_csv(path, stage, target_pe, optimized_pe):
    check(load().litho_write_imaging_metrics_csv(_path(path), {"CTM": 0, "LSM": 1}[stage], target_pe, optimized_pe))


def write_sampling_csv(path, grid_size, source_grid_size, source_points, kernels, pixel_size_nm, micro_core,
                       image_rows, image_cols, micro_count):
    check(load().litho_write_sampling_csv(_path(path), grid_size, source_grid_size, source_points, kernels,
                                          pixel_size_nm, 1 if micro_core else 0, image_rows, image_cols, micro_count))


def initialize_theta(ctx: Context, pattern, theta) -> None:
    check(ctx.lib.litho_initialize_theta(ctx.h, ptr(pattern), ptr(theta), pattern.numel()))
