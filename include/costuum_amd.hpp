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
 *   TMatrix::set_element                   TMatrix.hpp:1062-1080
 *   TMatrix::get_forward_scattering_matrix TMatrix.hpp:1109,1246,1306
 *   TMatrix::get_reduced_extinction_matrix TMatrix.hpp:2034-2048
 *   TMatrix::get_absorption_cross_sections,
 *            get_absorption_cross_section,
 *            get_average_absorption_cross_section
 *                                          TMatrix.hpp:2133-2287
 *   Matrix<T> (row-major, operator()(i, j)) Matrix.hpp:40-120
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

// The reference's floating-point type (Configuration.hpp: double).
#ifndef COSTUUM_AMD_NO_FLOAT_TYPE
typedef double float_type;
#endif

// Dense row-major matrix with the reference's accessors (Matrix.hpp).
template <typename T>
class Matrix {
 public:
  Matrix(size_t rows = 0, size_t cols = 0)
      : _rows(rows), _cols(cols), _v(rows * cols, T(0)) {}
  T &operator()(size_t i, size_t j) { return _v[i * _cols + j]; }
  const T &operator()(size_t i, size_t j) const { return _v[i * _cols + j]; }
  size_t get_number_of_rows() const { return _rows; }
  size_t get_number_of_columns() const { return _cols; }

 private:
  size_t _rows, _cols;
  std::vector<T> _v;
};

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

// 4 x 4 real matrices (Z, K) are returned as the reference's Matrix.
typedef ::Matrix<double> Matrix4;

// Device array of n doubles, freed on scope exit.
struct DeviceDoubles {
  double *p = nullptr;
  explicit DeviceDoubles(size_t n) {
    check_hip(hipMalloc(reinterpret_cast<void **>(&p),
                        std::max<size_t>(n, 1) * sizeof(double)),
              "hipMalloc");
    check_hip(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(double)),
              "hipMemset");
  }
  ~DeviceDoubles() {
    if (p) (void)hipFree(p);
  }
  DeviceDoubles(const DeviceDoubles &) = delete;
  DeviceDoubles &operator=(const DeviceDoubles &) = delete;
  std::vector<double> host(size_t n) const {
    std::vector<double> h(n);
    check_hip(hipMemcpy(h.data(), p, n * sizeof(double),
                        hipMemcpyDeviceToHost),
              "hipMemcpy");
    return h;
  }
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

  // Sets T^{(i1 i2)}_{n1 n2 m1 m2} (TMatrix.hpp:1062-1080); later queries
  // and cross sections use the new values.
  void set_element(uint_fast8_t i1, uint_fast32_t n1, uint_fast32_t m1,
                   uint_fast8_t i2, uint_fast32_t n2, uint_fast32_t m2,
                   const std::complex<double> &value) {
    const std::complex<double> &e = (*this)(i1, n1, m1, i2, n2, m2);
    const_cast<std::complex<double> &>(e) = value;
    _dirty = true;
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

  // S (2 x 2 complex) for a particle orientation (alpha, beta) and
  // incoming/outgoing directions (TMatrix.hpp:1306-1482), for the particle
  // frame (TMatrix.hpp:1246-1285) and from the sines and cosines of the
  // particle-frame angles (TMatrix.hpp:1109-1229; 1 / sin theta is not
  // needed, a zero sine being handled as the reference does).
  Matrix<std::complex<double>> get_forward_scattering_matrix(
      double alpha, double beta, double theta_in, double phi_in,
      double theta_out, double phi_out) const {
    double out[40];
    raw_query(alpha, beta, theta_in, phi_in, theta_out, phi_out, out);
    Matrix<std::complex<double>> S(2, 2);
    for (int i = 0; i < 4; ++i)
      S(i / 2, i % 2) = std::complex<double>(out[2 * i], out[2 * i + 1]);
    return S;
  }
  Matrix<std::complex<double>> get_forward_scattering_matrix(
      double theta_in, double phi_in, double theta_out,
      double phi_out) const {
    return get_forward_scattering_matrix(0., 0., theta_in, phi_in, theta_out,
                                         phi_out);
  }
  Matrix<std::complex<double>> get_forward_scattering_matrix(
      double costheta_in, double sintheta_in, double /*sintheta_in_inv*/,
      double cosphi_in, double sinphi_in, double costheta_out,
      double sintheta_out, double /*sintheta_out_inv*/, double cosphi_out,
      double sinphi_out) const {
    return get_forward_scattering_matrix(
        std::atan2(sintheta_in, costheta_in), std::atan2(sinphi_in, cosphi_in),
        std::atan2(sintheta_out, costheta_out),
        std::atan2(sinphi_out, cosphi_out));
  }

  // [C_ext, C_pol, C_cpol] = (2 pi / k) [Im(S11 + S22), Im(S11 - S22),
  // Re(S22 - S11)] for the forward direction (TMatrix.hpp:2034-2048).
  void get_reduced_extinction_matrix(double alpha, double beta,
                                     double theta_in, double phi_in,
                                     double output[3]) const {
    Matrix<std::complex<double>> S = get_forward_scattering_matrix(
        alpha, beta, theta_in, phi_in, theta_in, phi_in);
    const double pre = _lam;  // 2 pi / k
    output[0] = pre * (S(0, 0) + S(1, 1)).imag();
    output[1] = pre * (S(0, 0) - S(1, 1)).imag();
    output[2] = pre * (S(1, 1) - S(0, 0)).real();
  }

  // C_abs (result[0]) and C_abspol (result[1]) for incidence theta in the
  // particle frame, minus the scattered light over the given quadrature of
  // incoming directions (TMatrix.hpp:2133-2185).
  void get_absorption_cross_sections(
      double theta, uint_fast32_t ngauss, bool do_Cabs, bool do_Cabspol,
      const std::vector<double> &thetaGL,
      const std::vector<double> &costhetaweightsGL,
      const std::vector<double> &phiGL,
      const std::vector<double> &phiweightsGL, double *result) const {
    const size_t nq = 1 + (size_t)ngauss * ngauss;
    std::vector<ctm_query_spec> q(nq);
    q[0] = ctm_query_spec{0, 0, 0., 0., theta, 0., theta, 0.};
    for (uint_fast32_t i = 0; i < ngauss; ++i)
      for (uint_fast32_t j = 0; j < ngauss; ++j)
        q[1 + (size_t)i * ngauss + j] =
            ctm_query_spec{0, 0, 0., 0., thetaGL[i], phiGL[j], theta, 0.};
    std::vector<double> out(40 * nq);
    sync_device();
    costuum_amd_shim::check(ctm_query(_s->ctx, (int)nq, q.data(), out.data()),
                            "ctm_query", _s->ctx);
    auto S = [&](size_t k, int e) {
      return std::complex<double>(out[40 * k + 2 * e],
                                  out[40 * k + 2 * e + 1]);
    };
    const double pre = _lam;
    if (do_Cabs) result[0] = pre * (S(0, 0) + S(0, 3)).imag();
    if (do_Cabspol) result[1] = pre * (S(0, 0) - S(0, 3)).imag();
    for (uint_fast32_t i = 0; i < ngauss; ++i)
      for (uint_fast32_t j = 0; j < ngauss; ++j) {
        const size_t k = 1 + (size_t)i * ngauss + j;
        const double w = costhetaweightsGL[i] * phiweightsGL[j];
        const double a = std::norm(S(k, 0)) + std::norm(S(k, 1));
        const double b = std::norm(S(k, 2)) + std::norm(S(k, 3));
        if (do_Cabs) result[0] -= 0.5 * (a + b) * w;
        if (do_Cabspol) result[1] -= 0.5 * (a - b) * w;
      }
  }
  // C_abs, C_abspol pairs for ntheta angles with an ngauss x ngauss
  // Gauss-Legendre quadrature (TMatrix.hpp:2194-2215).
  void get_absorption_cross_sections(const double *theta,
                                     uint_fast32_t ntheta,
                                     uint_fast32_t ngauss,
                                     double *Cabs) const {
    std::vector<double> a(ntheta), p(ntheta);
    cabs(theta, ntheta, ngauss, a.data(), p.data());
    for (uint_fast32_t i = 0; i < ntheta; ++i) {
      Cabs[2 * i] = a[i];
      Cabs[2 * i + 1] = p[i];
    }
  }
  // C_abs only (TMatrix.hpp:2229-2250).
  void get_absorption_cross_section(const double *theta, uint_fast32_t ntheta,
                                    uint_fast32_t ngauss,
                                    double *Cabs) const {
    std::vector<double> p(ntheta);
    cabs(theta, ntheta, ngauss, Cabs, p.data());
  }
  // C_abs averaged over incidence directions with the same Gauss-Legendre
  // rule (TMatrix.hpp:2261-2287).
  double get_average_absorption_cross_section(uint_fast32_t ngauss) const {
    std::vector<double> x(ngauss), w(ngauss), th(ngauss), c(ngauss);
    costuum_amd_shim::check(
        ctm_gauss_legendre((uint32_t)ngauss, -1., 1., x.data(), w.data()),
        "ctm_gauss_legendre", nullptr);
    for (uint_fast32_t i = 0; i < ngauss; ++i) th[i] = std::acos(x[i]);
    get_absorption_cross_section(th.data(), ngauss, ngauss, c.data());
    double r = 0.;
    for (uint_fast32_t i = 0; i < ngauss; ++i) r += c[i] * w[i];
    return 0.5 * r;
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
  bool _dirty = false;  // set_element changed _host after the last upload
  mutable std::vector<std::complex<double>> _host;
  mutable std::vector<size_t> _block;  // start of block m in _host

  // Uploads elements changed by set_element before any device query.
  void sync_device() const {
    if (!_dirty) return;
    std::vector<double> buf(2 * _host.size());
    for (size_t i = 0; i < _host.size(); ++i) {
      buf[2 * i] = _host[i].real();
      buf[2 * i + 1] = _host[i].imag();
    }
    costuum_amd_shim::check(
        ctm_tmatrix_set(_s->ctx, 0, buf.data(), _host.size()),
        "ctm_tmatrix_set", _s->ctx);
    const_cast<TMatrix *>(this)->_dirty = false;
    const_cast<TMatrix *>(this)->_averaged = true;
  }

  void raw_query(double alpha, double beta, double theta_in, double phi_in,
                 double theta_out, double phi_out, double *out) const {
    sync_device();
    ctm_query_spec q{0, 0, alpha, beta, theta_in, phi_in, theta_out, phi_out};
    costuum_amd_shim::check(ctm_query(_s->ctx, 1, &q, out), "ctm_query",
                            _s->ctx);
  }

  // C_abs and C_abspol at ntheta angles: Q_abs, Q_abspol of the library
  // (scattered light subtracted with an ngauss x ngauss Gauss-Legendre
  // quadrature) times pi R_V^2.
  void cabs(const double *theta, uint_fast32_t ntheta, uint_fast32_t ngauss,
            double *a, double *p) const {
    if (ntheta == 0) return;
    sync_device();
    costuum_amd_shim::DeviceDoubles qa(ntheta), qp(ntheta);
    ctm_observable_spec spec{(int32_t)ntheta, theta, 0, 1, 1, (int32_t)ngauss,
                             0, 1};
    costuum_amd_shim::check(ctm_observables(_s->ctx, &spec, nullptr, nullptr,
                                            nullptr, qa.p, qp.p, nullptr),
                            "ctm_observables", _s->ctx);
    const std::vector<double> ha = qa.host(ntheta), hp = qp.host(ntheta);
    const double area = M_PI * _rv * _rv;
    for (uint_fast32_t i = 0; i < ntheta; ++i) {
      a[i] = ha[i] * area;
      p[i] = hp[i] * area;
    }
  }

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
    double out[40];
    raw_query(alpha, beta, theta_in, phi_in, theta_out, phi_out, out);
    costuum_amd_shim::Matrix4 r(4, 4);
    for (int i = 0; i < 16; ++i) r(i / 4, i % 4) = out[at + i];
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
    // the library averages the T-matrix it computes for (R_V, d, lambda,
    // m, nmax, ngauss): an ensemble or element-wise edited input has no
    // such description
    if (T._dirty || T._averaged)
      fail("apply_orientation_distribution needs a T-matrix from "
           "calculate_TMatrix (not an ensemble or edited one)",
           nullptr);
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
