// FP64 host restatement of the optics set-up (see optics.hpp for citations).
#include "optics.hpp"

#include <algorithm>
#include <cmath>

#include "plan.hpp"

namespace litho {

namespace {

// Eigen::VectorXd::LinSpaced for floating scalars
std::vector<double> linspaced(int num, double low, double high) {
  std::vector<double> out(num);
  if (num == 1) {
    out[0] = high;
    return out;
  }
  const double step = (high - low) / (double)(num - 1);
  if (std::abs(high) < std::abs(low)) {
    for (int i = 0; i < num; ++i) out[i] = (i == 0) ? low : high - (double)(num - 1 - i) * step;
  } else {
    for (int i = 0; i < num; ++i) out[i] = (i == num - 1) ? high : low + (double)i * step;
  }
  return out;
}

}  // namespace

double zernike_fringe(int index, double f, double g) {
  const double s3 = std::sqrt(3.0), s5 = std::sqrt(5.0), s6 = std::sqrt(6.0), s8 = std::sqrt(8.0);
  const double r2 = f * f + g * g;
  switch (index) {
    case 1: return 1.0;                                        // piston
    case 2: return 2.0 * f;                                    // tilt x
    case 3: return 2.0 * g;                                    // tilt y
    case 4: return s3 * (2.0 * r2 - 1.0);                      // defocus
    case 5: return s6 * (2.0 * f * g);                         // astigmatism 45
    case 6: return s6 * (f * f - g * g);                       // astigmatism 0
    case 7: return s8 * (3.0 * r2 - 2.0) * g;                  // coma y
    case 8: return s8 * (3.0 * r2 - 2.0) * f;                  // coma x
    case 9: return s8 * (3.0 * f * f * g - g * g * g);         // trefoil y
    case 10: return s8 * (f * f * f - 3.0 * f * g * g);        // trefoil x
    case 11: return s5 * (6.0 * r2 * r2 - 6.0 * r2 + 1.0);     // spherical
    default: return 0.0;
  }
}

SourceSet make_source_set(const OpticsParams& op) {
  if (op.n < 2) throw Error(1, "Grid: grid size must be at least 2");
  if (!(op.pixel_nm > 0.0) || !(op.wavelength_nm > 0.0) || !(op.na > 0.0))
    throw Error(1, "Grid: pixel size, wavelength and NA must be positive");
  SourceSet s;
  const int n = op.n;
  const double df = 1.0 / (n * op.pixel_nm);          // grid.cpp:42
  s.f_norm.resize(n);
  for (int i = 0; i < n; ++i) s.f_norm[i] = ((double)(i - n / 2) * df) / (op.na / op.wavelength_nm);
  // source.cpp:116-127 — axis points inside [-2, 2]
  std::vector<double> ff;
  for (double v : s.f_norm)
    if (v >= -2.0 && v <= 2.0) ff.push_back(v);
  const int num = (int)ff.size();
  const int up = 10;
  const int hn = num * up;
  const std::vector<double> hi = linspaced(hn, ff.front(), ff.back());
  // annulus on the fine grid, 10 x 10 mean (source.cpp:132-148), hard edges (smoothing 0)
  std::vector<double> wmap((size_t)num * num, 0.0);
  for (int i = 0; i < num; ++i) {
    for (int j = 0; j < num; ++j) {
      double sum = 0.0;
      for (int a = 0; a < up; ++a) {
        const double fy = hi[i * up + a];
        for (int b = 0; b < up; ++b) {
          const double fx = hi[j * up + b];
          const double rho = std::sqrt(fx * fx + fy * fy);
          const double outer = (op.sigma_out - rho) > 0.0 ? 1.0 : 0.0;
          const double inner = (rho - op.sigma_in) > 0.0 ? 1.0 : 0.0;
          sum += outer * inner;
        }
      }
      wmap[(size_t)i * num + j] = std::max(sum / (double)(up * up), 0.0);
    }
  }
  // litho_prepare.cpp:65-102, row-major, w > 1e-8
  for (int i = 0; i < num; ++i)
    for (int j = 0; j < num; ++j) {
      const double w = wmap[(size_t)i * num + j];
      if (w > 1e-8) {
        s.fs.push_back(ff[j]);
        s.gs.push_back(ff[i]);
        s.ws.push_back(w);
      }
    }
  if (s.fs.empty()) throw Error(5, "source_weight_map has no nonzero points");
  for (double w : s.ws) s.weight_sum += w;
  return s;
}

Window socs_window(const OpticsParams& op, const SourceSet& src) {
  const int n = op.n;
  const double dfx = src.f_norm[1] - src.f_norm[0];
  const double dfy = dfx;
  int y0 = n, x0 = n, y1 = 0, x1 = 0;
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      const double fx = src.f_norm[x], fy = src.f_norm[y];
      if (!((fx * fx + fy * fy) <= 1.0)) continue;   // pupil support (pupil.cpp:20-22)
      y0 = std::min(y0, y);
      x0 = std::min(x0, x);
      y1 = std::max(y1, y);
      x1 = std::max(x1, x);
    }
  const double gmax = *std::max_element(src.gs.begin(), src.gs.end());
  const double gmin = *std::min_element(src.gs.begin(), src.gs.end());
  const double fmax = *std::max_element(src.fs.begin(), src.fs.end());
  const double fmin = *std::min_element(src.fs.begin(), src.fs.end());
  const int oy = std::max(0, y0 - (int)std::floor(gmax / dfy) - 1);
  const int ox = std::max(0, x0 - (int)std::floor(fmax / dfx) - 1);
  const int ey = std::min(n, y1 - (int)std::floor(gmin / dfy) + 1);
  const int ex = std::min(n, x1 - (int)std::floor(fmin / dfx) + 1);
  return Window{oy, ox, ey - oy, ex - ox};
}

}  // namespace litho
