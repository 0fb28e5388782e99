"""The native library loads and exports every entry point include/litho_amd.h
declares (CPU only: no compute call is made)."""
import ctypes
import re
from pathlib import Path

from conftest import ROOT, litho_amd

HEADER = ROOT / "include" / "litho_amd.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(litho_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "litho_socs_forward" in syms and "litho_ctm_optimize" in syms
    assert len(syms) >= 15


def test_library_exports_every_declared_symbol():
    la = litho_amd()
    lib = ctypes.CDLL(str(la.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    la = litho_amd()
    assert sorted(la.SIGNATURES) == declared_symbols()


def test_host_only_entry_points():
    la = litho_amd()
    assert la.supported_size(2048) and la.supported_size(640) and not la.supported_size(2050)
    w = la.socs_window(512, 4.0, 193.0, 1.35, 0.6, 0.9)
    assert (w["rows"], w["cols"], w["num_points"]) == (55, 55, 380)
    assert abs(w["weight_sum"] - 299.64) < 1e-9
    w = la.socs_window(2048, 4.0, 193.0, 1.35, 0.6, 0.9)
    assert (w["rows"], w["cols"], w["num_points"], w["y0"]) == (219, 219, 5008, 915)
    # the host restatement matches the oracle's window and point set exactly
    from oracle import optics as op
    for n in (64, 65, 384, 640, 896, 1152, 2176):
        r, c, npts, s, y0, x0 = op.source_summary(n, 4.0, 193.0, 1.35, 0.6, 0.9)
        w = la.socs_window(n, 4.0, 193.0, 1.35, 0.6, 0.9)
        assert (w["rows"], w["cols"], w["num_points"], w["y0"], w["x0"]) == (r, c, npts, y0, x0)
        assert abs(w["weight_sum"] - s) <= 1e-12 * s


def test_error_reporting_without_gpu():
    la = litho_amd()
    try:
        la.socs_window(1, 4.0, 193.0, 1.35, 0.6, 0.9)
    except la.LithoError as e:
        assert e.code == la.EINVAL and "Grid" in str(e)
    else:
        raise AssertionError("size 1 must be rejected")


def test_partition_build_matches_oracle():
    """Partition::build (host-only entry point) equals the oracle's restatement."""
    la = litho_amd()
    from oracle import fullchip as fc
    for args in [(96, 96, 2, 2, 16, 8), (100, 90, 3, 2, 12, 12), (8192, 8192, 16, 16, 64, 40), (64, 64, 1, 1, 8, 4)]:
        ts, tiles = la.partition_build(*args)
        meta = fc.build(*args)
        assert ts == meta.tile_size
        assert len(tiles) == len(meta.tiles)
        for t, m in zip(tiles, meta.tiles):
            assert (t.id, t.row, t.col, t.halo, t.guard) == (m.id, m.row, m.col, m.halo, m.guard)
            for name in ("core_global", "tile_global", "core_local", "update_local"):
                r = getattr(m, name)
                assert getattr(t, name).tuple() == (r.y0, r.y1, r.x0, r.x1), name
            assert t.num_neighbors == len(m.neighbors)
            for j, nb in enumerate(m.neighbors):
                e = t.neighbors[j]
                assert (e.id, e.direction) == (nb.id, nb.direction)
                for name in ("guard_global", "my_guard_local", "neighbor_overlap_local", "overlap_global",
                             "my_overlap_local", "neighbor_guard_local"):
                    r = getattr(nb, name)
                    assert getattr(e, name).tuple() == (r.y0, r.y1, r.x0, r.x1), name


def test_partition_build_rejects_bad_grid():
    la = litho_amd()
    for args in [(0, 96, 2, 2, 16, 8), (96, 96, 0, 2, 16, 8), (96, 96, 2, 2, 8, 16)]:
        try:
            la.partition_build(*args)
        except la.LithoError as e:
            assert e.code == la.EINVAL
        else:
            raise AssertionError(f"{args} must be rejected")
