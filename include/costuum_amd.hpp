/* costuum_amd.hpp -- C++ surface of the reference's TMatrixCalculator and
 * TMatrix over the C ABI of costuum_amd.h, header-only, for C++ callers of
 * the T-matrix path (CosTuuM.cpp, PyTMatrix.hpp:202-221,
 * testTMatrixCalculator.cpp:101-103).
 *
 *   TMatrixCalculator::calculate_TMatrix   TMatrixCalculator.hpp:64-213
 *   TMatrixCalculator::apply_orientation_distribution
 *                                          TMatrixCalculator.hpp:229-307
 *   TMatrix::get_nmax, operator()          TMatrix.hpp:758, 1028-1046
 *   TMatrix::get_scattering_coefficient,
 *            get_extinction_coefficient    TMatrix.hpp:2055-2110
 *   TMatrix::get_scattering_matrix,
 *            get_extinction_matrix         TMatrix.hpp:1891-2010
 *
 * Every TMatrix owns a library context on HIP device 0 that holds its
 * T-matrix on the device; the caller owns the returned pointers, as with the
 * reference.  Failures (no convergence within the limits, a library error)
 * print a message and abort, as the reference's ctm_error does
 * (TMatrixCalculator.hpp:138-139,183-184).  Inputs reach the library as
 * device arrays through a few small buffers of the context's device.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "costuum_amd.h"

namespace costuum_amd_shim {

[[noreturn]] inline void fail(const char *what, const ctm_context *ctx) {
  std::fprintf(stderr, "costuum_amd: %s%s%s\n", what, ctx ? ": " : "",
               ctx ? ctm_last_error(ctx) : "");
  std::abort();
}

inline void check(int rc, const char *what, const ctm_context *ctx) {
  if (rc < 0) fail(what, ctx);
}

inline void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "costuum_amd: %s: %s\n", what, hipGetErrorString(e));
    std::abort();
  }
}

// A library context and the device buffers that carry one grain's
// parameters (k copies each, for windows of candidate orders).
struct Session {
  ctm_context *ctx = nullptr;
  double *dev = nullptr;  // 5 x cap doubles, then 2 x cap outputs
  int cap = 0;
  explicit Session(int capacity) : cap(capacity) {
    check(ctm_create(&ctx, 0), "ctm_create", nullptr);
    check_hip(hipMalloc(reinterpret_cast<void **>(&dev),
                        7 * sizeof(double) * (size_t)cap),
              "hipMalloc");
  }
  ~Session() {
    if (dev) (void)hipFree(dev);
    if (ctx) (void)ctm_destroy(ctx);
  }
  Session(const Session &) = delete;
  Session &operator=(const Session &) = delete;
  // k copies of the five grain parameters on the device
  void put(double rv, double d, double lam, std::complex<double> mr, int k) {
    if (k > cap) fail("too many candidates for the session buffers", nullptr);
    std::vector<double> h(5 * (size_t)cap);
    const double v[5] = {rv, d, lam, mr.real(), mr.imag()};
    for (int a = 0; a < 5; ++a)
      for (int i = 0; i < k; ++i) h[(size_t)a * cap + i] = v[a];
    check_hip(hipMemcpy(dev, h.data(), h.size() * sizeof(double),
                        hipMemcpyHostToDevice),
              "hipMemcpy");
  }
  const double *arg(int a) const { return dev + (size_t)a * cap; }
  double *out(int j) { return dev + (size_t)(5 + j) * cap; }
};

// Row-major 4 x 4 real matrix; Z(i, j) as with the reference's Matrix.
struct Matrix4 {
  double v[16];
  double operator()(int i, int j) const { return v[4 * i + j]; }
};

}  // namespace costuum_amd_shim

class TMatrixCalculator;

// A converged (or orientation-averaged) T-matrix held by the library.
class TMatrix {
 public:
  uint_fast32_t get_nmax() const { return _nmax; }
  uint_fast32_t get_ngauss() const { return _ngauss; }

  // T^{(i1 i2)}_{n1 n2 m1 m2} (m1 = m2), the reference's element layout
  // (TMatrix.hpp:1028-1046).
  const std::complex<double> &operator()(uint_fast8_t i1, uint_fast32_t n1,
                                         uint_fast32_t m1, uint_fast8_t i2,
                                         uint_fast32_t n2,
                                         uint_fast32_t m2) const {
    if (m1 != m2 || m1 > _nmax || n1 < 1 || n1 > _nmax || n2 < 1 ||
        n2 > _nmax || i1 > 1 || i2 > 1 || n1 < m1 || n2 < m2)
      costuum_amd_shim::fail("T-matrix element out of range", nullptr);
    load();
    const size_t nm = _nmax + 1 - (m1 > 0 ? m1 : 1);
    const size_t p1 = n1 - (m1 > 0 ? m1 : 1), p2 = n2 - (m2 > 0 ? m2 : 1);
    return _host[_block[m1] + (i1 * nm + p1) * 2 * nm + i2 * nm + p2];
  }

  // Sums over the whole T-matrix (TMatrix.hpp:2055-2110), unnormalised:
  // sum_m f_m sum_{n1, n2} |T^(ij)|^2 and sum_m f_m sum_n Re T^(ij)_nn,
  // f_0 = 1, f_m = 2.
  double get_scattering_coefficient() const {
    double q = 0.;
    for (uint_fast32_t m = 0; m <= _nmax; ++m) {
      const double f = m > 0 ? 2. : 1.;
      for (uint_fast32_t n1 = std::max<uint_fast32_t>(m, 1); n1 <= _nmax; ++n1)
        for (uint_fast32_t n2 = std::max<uint_fast32_t>(m, 1); n2 <= _nmax;
             ++n2)
          for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
              q += f * std::norm((*this)(i, n1, m, j, n2, m));
    }
    return q;
  }
  double get_extinction_coefficient() const {
    double q = 0.;
    for (uint_fast32_t m = 0; m <= _nmax; ++m) {
      const double f = m > 0 ? 2. : 1.;
      for (uint_fast32_t n = std::max<uint_fast32_t>(m, 1); n <= _nmax; ++n)
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) q += f * (*this)(i, n, m, j, n, m).real();
    }
    return q;
  }

  // Z for a particle orientation (alpha, beta) and incoming/outgoing
  // directions (radians), TMatrix.hpp:1891-1952.
  costuum_amd_shim::Matrix4 get_scattering_matrix(
      double alpha, double beta, double theta_in, double phi_in,
      double theta_out, double phi_out) const {
    return query(alpha, beta, theta_in, phi_in, theta_out, phi_out, 8);
  }
  // K for a particle orientation and an incoming direction,
  // TMatrix.hpp:1972-2010.
  costuum_amd_shim::Matrix4 get_extinction_matrix(double alpha, double beta,
                                                  double theta_in,
                                                  double phi_in) const {
    return query(alpha, beta, theta_in, phi_in, theta_in, phi_in, 24);
  }

 private:
  friend class TMatrixCalculator;
  std::shared_ptr<costuum_amd_shim::Session> _s;
  uint_fast32_t _nmax = 0, _ngauss = 0;
  double _rv = 0., _d = 1., _lam = 1.;  // equal-volume radius, d, lambda
  std::complex<double> _mr;
  bool _averaged = false;
  mutable std::vector<std::complex<double>> _host;
  mutable std::vector<size_t> _block;  // start of block m in _host

  void load() const {
    if (!_host.empty()) return;
    const size_t n = ctm_tmatrix_size(_s->ctx, 0);
    std::vector<double> buf(2 * n);
    costuum_amd_shim::check(
        _averaged ? ctm_average_get(_s->ctx, 0, buf.data(), n)
                  : ctm_tmatrix_get(_s->ctx, 0, buf.data(), n),
        "reading the T-matrix", _s->ctx);
    _host.resize(n);
    for (size_t i = 0; i < n; ++i) _host[i] = {buf[2 * i], buf[2 * i + 1]};
    _block.assign(_nmax + 1, 0);
    size_t off = 0;
    for (uint_fast32_t m = 0; m <= _nmax; ++m) {
      _block[m] = off;
      const size_t nm = _nmax + 1 - (m > 0 ? m : 1);
      off += 4 * nm * nm;
    }
  }

  costuum_amd_shim::Matrix4 query(double alpha, double beta, double theta_in,
                                  double phi_in, double theta_out,
                                  double phi_out, int at) const {
    ctm_query_spec q{0, 0, alpha, beta, theta_in, phi_in, theta_out, phi_out};
    double out[40];
    costuum_amd_shim::check(ctm_query(_s->ctx, 1, &q, out), "ctm_query",
                            _s->ctx);
    costuum_amd_shim::Matrix4 r;
    for (int i = 0; i < 16; ++i) r.v[i] = out[at + i];
    return r;
  }
};

class TMatrixCalculator {
 public:
  // Two-phase convergence of TMatrixCalculator.hpp:64-213: nmax from
  // max(4, x + 4.05 x^(1/3)) upward at ngauss = ndgs nmax until both m = 0
  // sums change by at most the tolerance (then one order back), ngauss from
  // ndgs nmax + 1 upward the same way, then every m block.  Candidates are
  // evaluated 8 at a time on the GPU and walked in order here.
  // ratio_of_radii != 1: axi is the equal-surface-area radius.
  static TMatrix *calculate_TMatrix(double ratio_of_radii, double axis_ratio,
                                    double axi, double wavelength,
                                    uint_fast32_t maximum_order,
                                    double tolerance, uint_fast32_t ndgs,
                                    std::complex<double> mr,
                                    uint_fast32_t maximum_ngauss) {
    using namespace costuum_amd_shim;
    constexpr int kWindow = 8;
    if (std::fabs(ratio_of_radii - 1.) > 1.e-8)
      ratio_of_radii = ctm_rv_over_rs(axis_ratio);
    const double rv = ratio_of_radii * axi;
    const double xev = 2. * M_PI * rv / wavelength;
    auto s = std::make_shared<Session>(kWindow);
    s->put(rv, axis_ratio, wavelength, mr, kWindow);
    double old_s = 0., old_e = 0., dext = 1., dsca = 1.;
    auto going = [&](uint_fast32_t n, uint_fast32_t limit) {
      return n < limit && (dext > tolerance || dsca > tolerance);
    };
    // one window of candidates (nmax, ngauss) walked in order
    auto walk = [&](uint_fast32_t &counter, uint_fast32_t limit,
                    bool vary_order, uint_fast32_t fixed) {
      uint32_t nm[kWindow], ng[kWindow];
      double qs[kWindow], qe[kWindow];
      int k = 0;
      for (uint_fast32_t c = counter; c < limit && k < kWindow; ++c, ++k) {
        nm[k] = (uint32_t)(vary_order ? c : fixed);
        ng[k] = (uint32_t)(vary_order ? ndgs * c : c);
      }
      check(ctm_m0_sums(s->ctx, k, s->arg(0), s->arg(1), s->arg(2),
                        s->arg(3), s->arg(4), nm, ng, qs, qe),
            "ctm_m0_sums", s->ctx);
      for (int i = 0; i < k; ++i) {
        dsca = std::fabs((old_s - qs[i]) / qs[i]);
        dext = std::fabs((old_e - qe[i]) / qe[i]);
        old_s = qs[i];
        old_e = qe[i];
        ++counter;
        if (!going(counter, limit)) break;
      }
    };
    uint_fast32_t nmax =
        (uint_fast32_t)std::max(4., xev + 4.05 * std::cbrt(xev));
    while (going(nmax, maximum_order)) walk(nmax, maximum_order, true, 0);
    if (nmax == maximum_order) fail("Unable to converge!", nullptr);
    --nmax;
    uint_fast32_t ngauss = nmax * ndgs + 1;
    dext = dsca = tolerance + 1.;
    while (going(ngauss, maximum_ngauss))
      walk(ngauss, maximum_ngauss, false, nmax);
    if (ngauss == maximum_ngauss) fail("Unable to converge!", nullptr);
    --ngauss;
    return fixed(s, rv, axis_ratio, wavelength, mr, nmax, ngauss);
  }

  // The ensemble T-matrix for an orientation distribution with Legendre
  // coefficients dist.get_coefficient(n), n = 0 .. 2 nmax
  // (OrientationDistribution.hpp:251-254); the result has its own context.
  template <class Distribution>
  static TMatrix *apply_orientation_distribution(const TMatrix &T,
                                                 const Distribution &dist) {
    using namespace costuum_amd_shim;
    auto s = std::make_shared<Session>(1);
    s->put(T._rv, T._d, T._lam, T._mr, 1);
    TMatrix *A = fixed(s, T._rv, T._d, T._lam, T._mr, T._nmax, T._ngauss);
    std::vector<double> pN(2 * T._nmax + 1);
    for (size_t n = 0; n < pN.size(); ++n) pN[n] = dist.get_coefficient(n);
    check(ctm_orientation_average(s->ctx, 1, (int)pN.size(), pN.data(),
                                  nullptr),
          "ctm_orientation_average", s->ctx);
    A->_averaged = true;
    return A;
  }

 private:
  static TMatrix *fixed(const std::shared_ptr<costuum_amd_shim::Session> &s,
                        double rv, double d, double lam,
                        std::complex<double> mr, uint_fast32_t nmax,
                        uint_fast32_t ngauss) {
    using namespace costuum_amd_shim;
    const uint32_t nm = (uint32_t)nmax, ng = (uint32_t)ngauss;
    check(ctm_tmatrix_fixed(s->ctx, 1, s->arg(0), s->arg(1), s->arg(2),
                            s->arg(3), s->arg(4), &nm, &ng,
                            CTM_PRECISION_FP64, s->out(0), s->out(1)),
          "ctm_tmatrix_fixed", s->ctx);
    check(ctm_synchronize(s->ctx), "ctm_synchronize", s->ctx);
    TMatrix *T = new TMatrix();
    T->_s = s;
    T->_nmax = nmax;
    T->_ngauss = ngauss;
    T->_rv = rv;
    T->_d = d;
    T->_lam = lam;
    T->_mr = mr;
    return T;
  }
};
