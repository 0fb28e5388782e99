// Host-side text formats of the reference's callers (io.cpp).
#pragma once
#include <vector>

#include "../../include/litho_amd.h"
#include "plan.hpp"

namespace litho {

void save_txt(const char* path, const double* m, int rows, int cols, int precision);
void save_txt_int(const char* path, const double* m, int rows, int cols);
void load_txt(const char* path, std::vector<double>& out, int* rows, int* cols);
void save_matrix_txt(const char* path, const double* m, int rows, int cols);
void write_history_csv(const char* path, int stage, const litho_fullchip_row* rows, int count);
void write_summary_csv(const char* path, int stage, int best_iteration, double best_loss, int final_iteration,
                       double final_loss, double seconds);
void write_imaging_metrics_csv(const char* path, int stage, double target_pe, double optimized_pe);
void write_sampling_csv(const char* path, int grid_size, int source_grid_size, int source_points, int kernels,
                        double pixel_size_nm, int micro_core, int image_rows, int image_cols, int micro_count);

}  // namespace litho
