"""The on-disk formats of the reference's callers through the C ABI (host
only, no GPU): SaveTxt save_mat / save_mat_int / load_txt, the full-chip
matrix TXT, history / summary / metrics / sampling CSV.  Expected texts are
written out by hand from the reference's stream settings."""
import csv

import numpy as np
import pytest

from conftest import litho_amd


def test_save_txt_fixed_precision(tmp_path):
    la = litho_amd()
    p = tmp_path / "m.txt"
    la.save_txt(p, np.array([[0.5, -1.25, 2.0], [1 / 3, 0.0, 10.0]]), precision=4)
    assert p.read_text() == "# rows=2 cols=3\n0.5000 -1.2500 2.0000\n0.3333 0.0000 10.0000\n"
    la.save_txt_int(p, np.array([[0.4, 0.6, -1.6], [2.5, 3.49, 7.0]]))
    # std::round: half away from zero
    assert p.read_text() == "# rows=2 cols=3\n0 1 -2\n3 3 7\n"


def test_load_txt_round_trip_and_errors(tmp_path):
    la = litho_amd()
    rng = np.random.default_rng(0)
    m = rng.standard_normal((7, 5))
    p = tmp_path / "a.txt"
    la.save_txt(p, m, precision=17)
    assert np.array_equal(la.load_txt(p), m)          # 17 decimals in fixed notation reproduce these doubles
    q = tmp_path / "b.txt"
    q.write_text("# comment\n\n 1 2 3\n# more\n4 5 6\n\t\n")
    assert la.load_txt(q).tolist() == [[1, 2, 3], [4, 5, 6]]
    q.write_text("1 2 3\n4 5\n")
    with pytest.raises(la.LithoError) as e:
        la.load_txt(q)
    assert e.value.code == la.ERUNTIME and "inconsistent column count at row 2" in str(e.value)
    q.write_text("# only a comment\n")
    with pytest.raises(la.LithoError):
        la.load_txt(q)
    with pytest.raises(la.LithoError):
        la.load_txt(tmp_path / "missing.txt")


def test_matrix_txt_eigen_stream_format(tmp_path):
    la = litho_amd()
    p = tmp_path / "g.txt"
    la.save_matrix_txt(p, np.array([[0.1, 1.0], [-2.5, 3.0]]))
    w = len("0.10000000000000001")
    want = ("# rows=2 cols=2\n" + "0.10000000000000001" + " " + "1".rjust(w) + "\n" + "-2.5".rjust(w) + " " +
            "3".rjust(w) + "\n")
    assert p.read_text() == want
    m = np.random.default_rng(1).random((4, 6))
    la.save_matrix_txt(p, m)
    assert np.array_equal(la.load_txt(p), m)


CTM_HEADER = ("iteration,loss,data_loss,cept_loss,pe_contribution,penalty_contribution,epe_contribution,"
              "cept_contribution,pe_loss_raw,bg_penalty_raw,epe_squared_pixel,epe_abs_sum_nm,epe_point_count")
LSM_HEADER = ("iteration,loss,data_loss,cept_loss,min_dt,max_dt,pe_contribution,penalty_contribution,epe_contribution,"
              "cept_contribution,core_pe,core_bg_penalty_raw,core_pe_penalty,core_epe_squared_pixel,"
              "core_epe_abs_sum_nm,core_ep_count,core_epe_mean_nm,core_data_loss,total_pe,total_bg_penalty_raw,"
              "total_epe_squared_pixel,total_epe_abs_sum_nm,total_ep_count,epe_mean_nm")


def test_history_and_summary_csv(tmp_path):
    la = litho_amd()
    rows = []
    for i in range(3):
        r = la.FullChipRow()
        r.iteration, r.loss, r.data_loss, r.cept_loss = i, 10.0 / (i + 1), 9.0 / (i + 1), 0.1 * i
        r.min_dt, r.max_dt, r.pe_contribution, r.core_pe, r.pe_loss_raw = 0.25, 0.5, 8.0, 1.0 / 3.0, 8.5
        rows.append(r)
    p = tmp_path / "history.csv"
    la.write_history_csv(p, "CTM", rows)
    lines = p.read_text().splitlines()
    assert lines[0] == CTM_HEADER and len(lines) == 4
    rec = list(csv.DictReader(open(p)))
    assert [int(x["iteration"]) for x in rec] == [0, 1, 2]
    assert float(rec[2]["loss"]) == 10.0 / 3 and rec[1]["pe_loss_raw"] == "8.5"
    la.write_history_csv(p, "LSM", rows)
    rec = list(csv.DictReader(open(p)))
    assert open(p).readline().strip() == LSM_HEADER
    assert float(rec[0]["core_pe"]) == 1.0 / 3.0 and rec[0]["min_dt"] == "0.25" and rec[0]["total_pe"] == "8.5"
    s = tmp_path / "summary.csv"
    la.write_summary_csv(s, "CTM", 2, 3.25, 50, 3.5)
    assert s.read_text() == "best_iteration,best_loss,final_iteration,final_loss\n2,3.25,50,3.5\n"
    la.write_summary_csv(s, "LSM", 1, 0.1, 20, 0.2, 12.5)
    assert s.read_text() == ("best_iteration,best_loss,final_iteration,final_loss,seconds\n"
                             "1,0.10000000000000001,20,0.20000000000000001,12.5\n")


def test_imaging_comparison_csv(tmp_path):
    la = litho_amd()
    m = tmp_path / "metrics.csv"
    la.write_imaging_metrics_csv(m, "LSM", 12.5, 3.0)
    assert m.read_text() == "case,pe\ntarget,12.5\noptimized_lsm,3\n"
    s = tmp_path / "sampling.csv"
    la.write_sampling_csv(s, 640, 640, 552, 24, 4.0, True, 8192, 8192, 256)
    assert s.read_text() == ("grid_size,source_grid_size,source_points,kernels,pixel_size_nm,imaging_mode,image_rows,"
                             "image_cols,micro_count\n640,640,552,24,4,micro_core,8192,8192,256\n")
