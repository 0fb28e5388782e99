// On-disk formats of the reference's callers, host only:
// * SaveTxt::save_mat / save_mat_int / load_txt (save_txt.cpp:23-105);
// * the "# rows= cols=" matrix TXT of the full-chip demos and of
//   save_full_chip_imaging_comparison (full_chip_imaging_compare.cpp:76-80):
//   an Eigen matrix streamed at precision 17, every coefficient padded on the
//   left to the widest one's length;
// * history.csv / summary.csv (demo_full_chip_ctm.cpp:112-128,
//   demo_full_chip_lsm.cpp:170-201) and metrics.csv / sampling.csv
//   (full_chip_imaging_compare.cpp:94-101).
// Text is written with C++ streams, the reference's own formatting.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "io.hpp"

namespace litho {

namespace {

std::ofstream open_out(const char* path) {
  if (path == nullptr) throw Error(1, "null path");
  std::ofstream f(path);
  if (!f) throw Error(5, std::string("cannot open ") + path + " for writing");
  return f;
}

void check_matrix(const double* m, int rows, int cols) {
  if (rows < 0 || cols < 0) throw Error(1, "negative matrix size");
  if (m == nullptr && rows > 0 && cols > 0) throw Error(1, "null matrix");
}

void finish(std::ofstream& f, const char* path) {
  f.flush();
  if (!f) throw Error(5, std::string("write failed: ") + path);
}

}  // namespace

void save_txt(const char* path, const double* m, int rows, int cols, int precision) {
  check_matrix(m, rows, cols);
  if (precision < 0) throw Error(1, "negative precision");
  std::ofstream f = open_out(path);
  f << std::fixed << std::setprecision(precision);
  f << "# rows=" << rows << " cols=" << cols << "\n";
  for (int r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) {
      f << m[(size_t)r * cols + c];
      if (c < cols - 1) f << " ";
    }
    f << "\n";
  }
  finish(f, path);
}

void save_txt_int(const char* path, const double* m, int rows, int cols) {
  check_matrix(m, rows, cols);
  std::ofstream f = open_out(path);
  f << "# rows=" << rows << " cols=" << cols << "\n";
  for (int r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) {
      f << static_cast<int>(std::round(m[(size_t)r * cols + c]));
      if (c < cols - 1) f << " ";
    }
    f << "\n";
  }
  finish(f, path);
}

void load_txt(const char* path, std::vector<double>& out, int* rows, int* cols) {
  if (path == nullptr) throw Error(1, "null path");
  std::ifstream f(path);
  if (!f) throw Error(5, std::string("load_txt: cannot open ") + path);
  std::string line;
  size_t nc = 0, nr = 0;
  out.clear();
  while (std::getline(f, line)) {
    const size_t p = line.find_first_not_of(" \t\r\n");
    if (p == std::string::npos || line[p] == '#') continue;
    std::istringstream iss(line);
    std::vector<double> vals;
    double v;
    while (iss >> v) vals.push_back(v);
    if (vals.empty()) continue;
    if (nc == 0)
      nc = vals.size();
    else if (vals.size() != nc)
      throw Error(5, "load_txt: inconsistent column count at row " + std::to_string(nr + 1) + " in " + path);
    out.insert(out.end(), vals.begin(), vals.end());
    ++nr;
  }
  if (nr == 0) throw Error(5, std::string("load_txt: empty matrix in ") + path);
  *rows = (int)nr;
  *cols = (int)nc;
}

void save_matrix_txt(const char* path, const double* m, int rows, int cols) {
  check_matrix(m, rows, cols);
  std::ofstream f = open_out(path);
  f << "# rows=" << rows << " cols=" << cols << '\n' << std::setprecision(17);
  // the widest coefficient at the stream's settings sets every field's width
  std::streamsize width = 0;
  for (size_t i = 0; i < (size_t)rows * cols; ++i) {
    std::ostringstream ss;
    ss.copyfmt(f);
    ss << m[i];
    width = std::max<std::streamsize>(width, (std::streamsize)ss.str().size());
  }
  for (int r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) {
      if (c > 0) f << ' ';
      f.width(width);
      f << m[(size_t)r * cols + c];
    }
    if (r < rows - 1) f << '\n';
  }
  f << '\n';
  finish(f, path);
}

void write_history_csv(const char* path, int stage, const litho_fullchip_row* rows, int count) {
  if (stage != 0 && stage != 1) throw Error(1, "history stage must be CTM (0) or LSM (1)");
  if (count < 0 || (count > 0 && rows == nullptr)) throw Error(1, "null history rows");
  std::ofstream f = open_out(path);
  if (stage == 0) {
    f << "iteration,loss,data_loss,cept_loss,pe_contribution,penalty_contribution,epe_contribution,"
         "cept_contribution,pe_loss_raw,bg_penalty_raw,epe_squared_pixel,epe_abs_sum_nm,epe_point_count\n"
      << std::setprecision(17);
    for (int i = 0; i < count; ++i) {
      const litho_fullchip_row& r = rows[i];
      f << r.iteration << ',' << r.loss << ',' << r.data_loss << ',' << r.cept_loss << ',' << r.pe_contribution << ','
        << r.penalty_contribution << ',' << r.epe_contribution << ',' << r.cept_contribution << ',' << r.pe_loss_raw
        << ',' << r.bg_penalty_raw << ',' << r.epe_loss_raw << ',' << r.epe_abs_sum_nm << ',' << r.epe_point_count
        << '\n';
    }
  } else {
    f << "iteration,loss,data_loss,cept_loss,min_dt,max_dt,pe_contribution,penalty_contribution,epe_contribution,"
         "cept_contribution,core_pe,core_bg_penalty_raw,core_pe_penalty,core_epe_squared_pixel,core_epe_abs_sum_nm,"
         "core_ep_count,core_epe_mean_nm,core_data_loss,total_pe,total_bg_penalty_raw,total_epe_squared_pixel,"
         "total_epe_abs_sum_nm,total_ep_count,epe_mean_nm\n"
      << std::setprecision(17);
    for (int i = 0; i < count; ++i) {
      const litho_fullchip_row& r = rows[i];
      f << r.iteration << ',' << r.loss << ',' << r.data_loss << ',' << r.cept_loss << ',' << r.min_dt << ','
        << r.max_dt << ',' << r.pe_contribution << ',' << r.penalty_contribution << ',' << r.epe_contribution << ','
        << r.cept_contribution << ',' << r.core_pe << ',' << r.core_bg_penalty_raw << ',' << r.core_pe_penalty << ','
        << r.core_epe_loss << ',' << r.core_epe_abs_sum_nm << ',' << r.core_ep_count << ',' << r.core_epe_mean << ','
        << r.core_data_loss << ',' << r.pe_loss_raw << ',' << r.bg_penalty_raw << ',' << r.epe_loss_raw << ','
        << r.epe_abs_sum_nm << ',' << r.epe_point_count << ',' << r.epe_mean << '\n';
    }
  }
  finish(f, path);
}

void write_summary_csv(const char* path, int stage, int best_iteration, double best_loss, int final_iteration,
                       double final_loss, double seconds) {
  if (stage != 0 && stage != 1) throw Error(1, "summary stage must be CTM (0) or LSM (1)");
  std::ofstream f = open_out(path);
  f << (stage == 0 ? "best_iteration,best_loss,final_iteration,final_loss\n"
                   : "best_iteration,best_loss,final_iteration,final_loss,seconds\n")
    << std::setprecision(17);
  f << best_iteration << ',' << best_loss << ',' << final_iteration << ',' << final_loss;
  if (stage == 1) f << ',' << seconds;
  f << '\n';
  finish(f, path);
}

void write_imaging_metrics_csv(const char* path, int stage, double target_pe, double optimized_pe) {
  if (stage != 0 && stage != 1) throw Error(1, "metrics stage must be CTM (0) or LSM (1)");
  std::ofstream f = open_out(path);
  f << "case,pe\n" << std::setprecision(17);
  f << "target," << target_pe << '\n' << (stage == 1 ? "optimized_lsm" : "optimized_gray") << ',' << optimized_pe
    << '\n';
  finish(f, path);
}

void write_sampling_csv(const char* path, int grid_size, int source_grid_size, int source_points, int kernels,
                        double pixel_size_nm, int micro_core, int image_rows, int image_cols, int micro_count) {
  std::ofstream f = open_out(path);
  f << "grid_size,source_grid_size,source_points,kernels,pixel_size_nm,imaging_mode,image_rows,image_cols,"
       "micro_count\n";
  f << grid_size << ',' << source_grid_size << ',' << source_points << ',' << kernels << ',' << pixel_size_nm << ','
    << (micro_core ? "micro_core" : "full_frame") << ',' << image_rows << ',' << image_cols << ',' << micro_count
    << '\n';
  finish(f, path);
}

}  // namespace litho
