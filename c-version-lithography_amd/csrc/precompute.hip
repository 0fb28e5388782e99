// One-time SOCS kernel precompute on the device, FP64
// (LithoPrepare::_compute_socs_kernels, litho_prepare.cpp:225-385).
//
// The reference forms A = [sqrt(w_i) P(f + f_i)] over the compact window
// (W^2 rows x Ns columns) and takes a thin BDCSVD.  Here A is built by a
// kernel, G = A^H A (Ns x Ns Hermitian) by rocBLAS zherk, the K largest
// eigenpairs (sigma_k^2, v_k) of G by rocSOLVER zheevx, and the left singular
// vectors U_k = A v_k / sigma_k by zgemm.  lambda_k = sigma_k^2 as in the
// reference (litho_prepare.cpp:328).  Kernel phases are arbitrary, as they are
// for an SVD; observables (the aerial image) do not depend on them.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <string>
#include <cmath>
#include <vector>

#include "optics.hpp"
#include "plan.hpp"

namespace litho {

namespace {

struct PointShift {
  int r0, c0;
  double tr, tc, sw;
};

// A[p + i*lda], p = r*nc + c: bilinear P(f + f_i) on the window, scaled by sqrt(w_i)
__global__ void k_build_a(const double2* __restrict__ pupil, int n, const PointShift* __restrict__ pts, int ns,
                          int oy, int ox, int nr, int nc, double2* __restrict__ a, size_t lda) {
  const int i = blockIdx.y;
  if (i >= ns) return;
  const PointShift ps = pts[i];
  const int npix = nr * nc;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    const int r = p / nc, c = p - r * nc;
    const int ra = oy + r + ps.r0, ca = ox + c + ps.c0;
    auto at = [&](int rr, int cc) -> double2 {
      if (rr < 0 || rr >= n || cc < 0 || cc >= n) return make_double2(0.0, 0.0);
      return pupil[(size_t)rr * n + cc];
    };
    const double2 v00 = at(ra, ca), v01 = at(ra, ca + 1), v10 = at(ra + 1, ca), v11 = at(ra + 1, ca + 1);
    const double w00 = (1 - ps.tr) * (1 - ps.tc), w01 = (1 - ps.tr) * ps.tc, w10 = ps.tr * (1 - ps.tc),
                 w11 = ps.tr * ps.tc;
    double2 out;
    out.x = ps.sw * (w00 * v00.x + w01 * v01.x + w10 * v10.x + w11 * v11.x);
    out.y = ps.sw * (w00 * v00.y + w01 * v01.y + w10 * v10.y + w11 * v11.y);
    a[(size_t)i * lda + p] = out;
  }
}

__global__ void k_scale_columns(double2* __restrict__ u, size_t ld, int rows, const double* __restrict__ inv_sigma) {
  const int k = blockIdx.y;
  const double s = inv_sigma[k];
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < rows; p += gridDim.x * blockDim.x) {
    double2 v = u[(size_t)k * ld + p];
    v.x *= s;
    v.y *= s;
    u[(size_t)k * ld + p] = v;
  }
}

// Smallest accepted relative gap (lambda_K - lambda_{K+1}) / lambda_K.  An
// exactly degenerate pair comes back with a gap of a few FP64 ulps, far
// below this.
constexpr double kMinEigenGap = 1e-9;

void rb(rocblas_status st, const char* what) {
  if (st != rocblas_status_success) throw Error(3, std::string(what) + " failed with rocblas status " +
                                                       std::to_string((int)st));
}

}  // namespace

void socs_precompute_device(const OpticsParams& op, int k, hipStream_t stream, double* kernels_out,
                            double* weights_out, double* singular_values_out, double* eigen_gap_out) {
  const SourceSet src = make_source_set(op);
  const Window win = socs_window(op, src);
  const int n = op.n;
  const int ns = (int)src.fs.size();
  if (k <= 0 || k > ns) throw Error(1, "SOCS: kernel count must be in [1, number of source points]");
  const double dfx = src.f_norm[1] - src.f_norm[0];
  const double dfy = dfx;

  // pupil on the centred grid: [rho^2 <= 1] exp(i 2 pi W), W = sum_j c_j Z_j
  // (pupil.cpp:14-23,41-46; zernike.cpp:46-53)
  for (const auto& z : op.zernike)
    if (z.second != 0.0 && (z.first < 1 || z.first > 11 || !std::isfinite(z.second)))
      throw Error(1, "Pupil: Zernike terms must have fringe index 1..11 and a finite coefficient");
  constexpr double kTwoPi = 6.283185307179586476925286766559;
  std::vector<double2> pupil((size_t)n * n);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      const double fx = src.f_norm[x], fy = src.f_norm[y];
      if (!((fx * fx + fy * fy) <= 1.0)) {
        pupil[(size_t)y * n + x] = make_double2(0.0, 0.0);
        continue;
      }
      double w = 0.0;
      for (const auto& z : op.zernike)
        if (z.second != 0.0) w += z.second * zernike_fringe(z.first, fx, fy);
      const double ph = kTwoPi * w;
      pupil[(size_t)y * n + x] = make_double2(std::cos(ph), std::sin(ph));
    }
  // per-point integer shift and bilinear weights, same operation order as the reference
  std::vector<PointShift> pts(ns);
  for (int i = 0; i < ns; ++i) {
    const double dc = src.fs[i] / dfx, dr = src.gs[i] / dfy;
    const int c0 = (int)std::floor(dc), r0 = (int)std::floor(dr);
    pts[i] = PointShift{r0, c0, dr - r0, dc - c0, std::sqrt(src.ws[i])};
  }
  const int nr = win.rows, nc = win.cols;
  const size_t lda = (size_t)nr * nc;

  DevMem d_pupil(pupil.size() * sizeof(double2));
  DevMem d_pts(pts.size() * sizeof(PointShift));
  DevMem d_a(lda * ns * sizeof(double2));
  DevMem d_g((size_t)ns * ns * sizeof(double2));
  LITHO_HIP(hipMemcpyAsync(d_pupil.ptr, pupil.data(), d_pupil.bytes, hipMemcpyHostToDevice, stream));
  LITHO_HIP(hipMemcpyAsync(d_pts.ptr, pts.data(), d_pts.bytes, hipMemcpyHostToDevice, stream));
  {
    const int bx = (int)std::min<size_t>((lda + 255) / 256, 1024);
    hipLaunchKernelGGL(k_build_a, dim3(bx, ns), dim3(256), 0, stream, d_pupil.as<double2>(), n,
                       d_pts.as<PointShift>(), ns, win.y0, win.x0, nr, nc, d_a.as<double2>(), lda);
    LITHO_HIP(hipGetLastError());
  }
  rocblas_handle h = nullptr;
  rb(rocblas_create_handle(&h), "rocblas_create_handle");
  struct HandleGuard {
    rocblas_handle h;
    ~HandleGuard() { rocblas_destroy_handle(h); }
  } guard{h};
  rb(rocblas_set_stream(h, stream), "rocblas_set_stream");
  const double one = 1.0, zero = 0.0;
  // G = A^H A (lower triangle)
  rb(rocblas_zherk(h, rocblas_fill_lower, rocblas_operation_conjugate_transpose, ns, (int)lda, &one,
                   d_a.as<rocblas_double_complex>(), (int)lda, &zero, d_g.as<rocblas_double_complex>(), ns),
     "rocblas_zherk");
  // Top-K eigenpairs of G by rocSOLVER zheevx: Householder tridiagonalisation,
  // bisection for the K selected eigenvalues (abstol 0 = most accurate),
  // inverse iteration with cluster re-orthogonalisation for their vectors.  Z is
  // sized for all n columns.  (The divide-and-conquer zheevd returned
  // non-finite eigenvectors for n >= 228 on ROCm 7.2, so it is not used.)
  DevMem d_w((size_t)ns * sizeof(double)), d_e((size_t)ns * sizeof(double)), d_info(sizeof(rocblas_int));
  std::vector<double> ev(k), all;
  auto read_info = [&](const char* what) {
    rocblas_int info = 0;
    LITHO_HIP(hipMemcpyAsync(&info, d_info.ptr, sizeof(info), hipMemcpyDeviceToHost, stream));
    LITHO_HIP(hipStreamSynchronize(stream));
    if (info != 0) throw Error(5, std::string(what) + " did not converge (info " + std::to_string(info) + ")");
  };
  if (singular_values_out != nullptr) {
    // all eigenvalues, no vectors, on a copy
    DevMem d_g2(d_g.bytes);
    LITHO_HIP(hipMemcpyAsync(d_g2.ptr, d_g.ptr, d_g.bytes, hipMemcpyDeviceToDevice, stream));
    rb(rocsolver_zheev(h, rocblas_evect_none, rocblas_fill_lower, ns, d_g2.as<rocblas_double_complex>(), ns,
                       d_w.as<double>(), d_e.as<double>(), d_info.as<rocblas_int>()),
       "rocsolver_zheev");
    read_info("rocsolver_zheev");
    all.resize(ns);
    LITHO_HIP(hipMemcpyAsync(all.data(), d_w.ptr, ns * sizeof(double), hipMemcpyDeviceToHost, stream));
    LITHO_HIP(hipStreamSynchronize(stream));
  }
  // The top K + 1 eigenpairs (K when K = Ns): eigenvalue K + 1 gives the
  // relative gap (lambda_K - lambda_{K+1}) / lambda_K, and a truncation that
  // splits a degenerate eigenvalue (gap at rounding level, e.g. the symmetric
  // pairs of an annular source) is rejected: its kernel set would depend on
  // the solver's choice of basis inside the eigenspace.
  const int kk = k < ns ? k + 1 : k;
  const int off = kk - k;
  ev.resize(kk);
  DevMem d_zx((size_t)ns * ns * sizeof(double2));
  {
    DevMem d_nev(sizeof(rocblas_int)), d_ifail((size_t)ns * sizeof(rocblas_int));
    rb(rocsolver_zheevx(h, rocblas_evect_original, rocblas_erange_index, rocblas_fill_lower, ns,
                        d_g.as<rocblas_double_complex>(), ns, 0.0, 0.0, ns - kk + 1, ns, 0.0, d_nev.as<rocblas_int>(),
                        d_w.as<double>(), d_zx.as<rocblas_double_complex>(), ns, d_ifail.as<rocblas_int>(),
                        d_info.as<rocblas_int>()),
       "rocsolver_zheevx");
    read_info("rocsolver_zheevx");
    rocblas_int nev = 0;
    LITHO_HIP(hipMemcpyAsync(&nev, d_nev.ptr, sizeof(nev), hipMemcpyDeviceToHost, stream));
    LITHO_HIP(hipMemcpyAsync(ev.data(), d_w.ptr, kk * sizeof(double), hipMemcpyDeviceToHost, stream));
    LITHO_HIP(hipStreamSynchronize(stream));
    if (nev != kk) throw Error(5, "rocsolver_zheevx returned " + std::to_string(nev) + " eigenpairs");
  }
  // ascending: ev[off .. off + k) are the top K, ev[0] the next one when off = 1
  const double lam_k = ev[off];
  const double gap = off ? (lam_k - ev[0]) / lam_k : 1.0;
  if (eigen_gap_out != nullptr) *eigen_gap_out = gap;
  if (!(gap > kMinEigenGap))
    throw Error(1, "SOCS: K = " + std::to_string(k) + " splits a degenerate eigenvalue (relative gap " +
                       std::to_string(gap) + "); choose another K");
  const rocblas_double_complex* d_z = d_zx.as<rocblas_double_complex>() + (size_t)off * ns;
  if (singular_values_out != nullptr)
    for (int i = 0; i < ns; ++i) singular_values_out[i] = std::sqrt(std::max(all[ns - 1 - i], 0.0));
  {
    std::vector<double2> z((size_t)ns * k);
    LITHO_HIP(hipMemcpyAsync(z.data(), d_z, z.size() * sizeof(double2), hipMemcpyDeviceToHost, stream));
    LITHO_HIP(hipStreamSynchronize(stream));
    size_t bad = 0;
    for (const double2& v : z) bad += !(std::isfinite(v.x) && std::isfinite(v.y));
    if (bad) throw Error(5, "SOCS precompute: " + std::to_string(bad) + " non-finite eigenvector entries from the eigensolver");
  }
  // U = A Z, columns scaled by 1/sigma; eigenvalues come back ascending
  DevMem d_u(lda * k * sizeof(double2));
  const rocblas_double_complex c1(1.0, 0.0), c0(0.0, 0.0);
  rb(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, (int)lda, k, ns, &c1,
                   d_a.as<rocblas_double_complex>(), (int)lda, d_z, ns, &c0,
                   d_u.as<rocblas_double_complex>(), (int)lda),
     "rocblas_zgemm");
  std::vector<double> inv(k);
  for (int j = 0; j < k; ++j) {
    if (!(ev[off + j] > 0.0)) throw Error(5, "SOCS: non-positive eigenvalue among the top K");
    inv[j] = 1.0 / std::sqrt(ev[off + j]);
  }
  DevMem d_inv(k * sizeof(double));
  LITHO_HIP(hipMemcpyAsync(d_inv.ptr, inv.data(), k * sizeof(double), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_scale_columns, dim3((int)std::min<size_t>((lda + 255) / 256, 1024), k), dim3(256), 0, stream,
                     d_u.as<double2>(), lda, (int)lda, d_inv.as<double>());
  LITHO_HIP(hipGetLastError());
  std::vector<double2> u(lda * k);
  LITHO_HIP(hipMemcpyAsync(u.data(), d_u.ptr, u.size() * sizeof(double2), hipMemcpyDeviceToHost, stream));
  LITHO_HIP(hipStreamSynchronize(stream));
  {
    size_t bad = 0;
    for (const double2& v : u) bad += !(std::isfinite(v.x) && std::isfinite(v.y));
    if (bad) throw Error(5, "SOCS precompute: " + std::to_string(bad) + " non-finite kernel entries after A v / s");
  }
  // descending order: output kernel j = eigen column k-1-j
  for (int j = 0; j < k; ++j) {
    const int col = k - 1 - j;
    weights_out[j] = ev[off + col];
    for (size_t p = 0; p < lda; ++p) {
      kernels_out[2 * ((size_t)j * lda + p)] = u[(size_t)col * lda + p].x;
      kernels_out[2 * ((size_t)j * lda + p) + 1] = u[(size_t)col * lda + p].y;
    }
  }
}

}  // namespace litho
