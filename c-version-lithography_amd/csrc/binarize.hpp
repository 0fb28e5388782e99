// CTM -> binary hand-off on the device: Imbinarize::binarize_advanced
// (imbinarize.cpp:84-190) — local-peak quantile threshold, 8-connected
// component area filter, SRAF extraction and minimum-spacing keep-out.
#pragma once
#include "plan.hpp"

namespace litho {

struct BinarizeOutputs {
  float* final_mask = nullptr;        // required
  float* initial_mask = nullptr;      // the rest nullable
  float* filtered_mask = nullptr;
  float* sraf_initial = nullptr;
  float* sraf_final = nullptr;
  float* keep_out = nullptr;
};

struct BinarizeInfo {
  double threshold = 0.0;             // adaptive threshold (0 when no peak)
  long long num_peaks = 0;
};

// gray, target: device height x width float32 (row-major).  Outputs are 0/1
// float32 device arrays of the same shape.  Synchronises the context stream.
BinarizeInfo binarize_advanced(Context& ctx, int height, int width, const float* gray, const float* target,
                               double peak_min_intensity, double quantile_val, int area_threshold,
                               int min_spacing_pixels, const BinarizeOutputs& out);

// Host restatement of OpenCV's getStructuringElement(MORPH_ELLIPSE, (2r+1)^2):
// half-width dx of row r + dy for dy = -r .. r (-1 for an empty row).
void ellipse_half_widths(int r, int* dx_out);

}  // namespace litho
