// Optical model (FP64 host code) and the one-time SOCS kernel precompute.
//
// Restates Grid (grid.cpp:4-48), Pupil with fringe Zernike aberrations
// (pupil.cpp:5-46, zernike.cpp:11-53),
// Source::compute_source_map (source.cpp:114-158),
// LithoPrepare::_extract_source_points (litho_prepare.cpp:65-102) and the
// compact SOCS window (litho_prepare.cpp:238-256) with the same IEEE operation
// order, so the window and the source-point set match the reference exactly.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

namespace litho {

struct OpticsParams {
  int n;
  double pixel_nm, wavelength_nm, na, sigma_in, sigma_out;
  // fringe Zernike terms {index 1..11, coefficient in waves} of the pupil
  // wavefront, summed in the given order; zero coefficients are skipped
  std::vector<std::pair<int, double>> zernike;
};

// Fringe Zernike Z_index(f, g) on normalised pupil coordinates
// (zernike.cpp:11-43); 0 for an index outside 1..11.
double zernike_fringe(int index, double f, double g);

struct SourceSet {
  std::vector<double> f_norm;          // centred frequency axis / (NA / lambda), length n
  std::vector<double> fs, gs, ws;      // source points (fx, fy, weight)
  double weight_sum = 0.0;
};

struct Window {
  int y0, x0, rows, cols;
};

SourceSet make_source_set(const OpticsParams& op);
Window socs_window(const OpticsParams& op, const SourceSet& src);

// Device precompute (rocBLAS zherk + rocSOLVER zheevx); outputs as in
// litho_socs_precompute (include/litho_amd.h).  eigen_gap_out (nullable):
// (lambda_K - lambda_{K+1}) / lambda_K, 1 when K is the number of source
// points; a gap at rounding level (a split degenerate eigenvalue) throws.
void socs_precompute_device(const OpticsParams& op, int k, hipStream_t stream, double* kernels_out,
                            double* weights_out, double* singular_values_out, double* eigen_gap_out);

}  // namespace litho
