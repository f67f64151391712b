// Host-only helpers of the library: Gauss-Legendre rules, the adaptive
// Gauss-Legendre quadrature, the equal-volume / equal-surface radius ratio of
// a spheroid and the Legendre coefficients of orientation distributions.
// Nothing here touches the GPU; the Python module and the C++ set-up code
// call these while building a batch.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "costuum_amd.h"
#include "ctm_host.hpp"

namespace ctm {

void gauss_legendre_host(unsigned order, double *x, double *w) {
  const unsigned odd = order % 2u, half = order / 2u + odd;
  const double inv = 1.0 / order, inv2 = inv * inv;
  for (unsigned i = 0; i < half; ++i) {
    const double t = M_PI * (i + 0.75) / (order + 0.5);
    double xi = (1. - 0.125 * inv2 + (5. / 38.) * inv * inv2 -
                 (2. / 25.) * inv2 * inv2 * (1. - (14. / 39.) / (t * t))) *
                std::cos(t);
    double tol = 1.e-16, pinv = 0., step = 0.;
    unsigned iter = 0;
    do {
      ++iter;
      if (iter > 100) tol *= 10.;
      double p_prev = 1., p_cur = xi;
      for (unsigned j = 1; j < order; ++j) {
        const double p_old = p_prev;
        p_prev = p_cur;
        p_cur = xi * p_prev + (xi * p_prev - p_old) * j / (j + 1.);
      }
      pinv = 1. / ((p_prev - xi * p_cur) * order);
      step = pinv * p_cur * (1. - xi * xi);
      xi = xi - step;
    } while (std::fabs(step) > tol * std::fabs(xi));
    const unsigned k = order - i - 1;
    x[k] = xi;
    w[k] = 2. * pinv * pinv * (1. - xi * xi);
    if (i != half - 1 || odd != 1) {
      x[i] = -x[k];
      w[i] = w[k];
    }
  }
}

void gauss_legendre_host_ab(unsigned order, double a, double b, double *x,
                            double *w) {
  gauss_legendre_host(order, x, w);
  const double half_width = 0.5 * (b - a), mid = 0.5 * (a + b);
  for (unsigned i = 0; i < order; ++i) {
    x[i] = half_width * x[i] + mid;
    w[i] *= half_width;
  }
}

namespace {

// Quadrature of f on [a, b] with Gauss-Legendre rules of increasing order
// until two successive results agree to abs_err or rel_err
// (SpecialFunctions.hpp:772-815).  Returns the last estimate; *reached
// tells whether a tolerance was met before max_order.
template <class F>
double adaptive_gl(F &&f, double a, double b, unsigned start, unsigned max,
                   double abs_err, double rel_err, bool *reached) {
  unsigned order = start;
  double previous = 0., dabs = abs_err + 1., drel = rel_err + 1.;
  std::vector<double> x, w;
  while (order < max && dabs > abs_err && drel > rel_err) {
    x.resize(order);
    w.resize(order);
    gauss_legendre_host_ab(order, a, b, x.data(), w.data());
    double q = 0.;
    for (unsigned i = 0; i < order; ++i) q += w[i] * f(x[i]);
    dabs = std::fabs(q - previous);
    drel = dabs / std::fabs(q + previous);
    previous = q;
    ++order;
  }
  if (reached) *reached = order != max;
  return previous;
}

// d^n_00(beta) = P_n(cos beta) by the m = 0 Wigner recurrence
// (SpecialFunctions.hpp:317-335), value at order n >= 1.
double legendre_by_wigner(double cosx, unsigned n) {
  double prev = 1., cur = cosx;
  for (unsigned k = 1; k < n; ++k) {
    const double kp1 = k + 1.;
    const double next = ((2. * k + 1.) * cosx * cur - k * prev) / kp1;
    prev = cur;
    cur = next;
  }
  return cur;
}

}  // namespace
}  // namespace ctm

using namespace ctm;

extern "C" {

int ctm_gauss_legendre(uint32_t order, double a, double b, double *x,
                       double *w) {
  if (order < 1 || !x || !w || !(a < b)) return CTM_ERR_ARGUMENT;
  if (a == -1. && b == 1.)
    gauss_legendre_host(order, x, w);
  else
    gauss_legendre_host_ab(order, a, b, x, w);
  return CTM_OK;
}

double ctm_rv_over_rs(double d) {
  if (d > 1.) {
    const double d2 = d * d;
    const double e = std::sqrt(1. - 1. / d2);
    const double c = std::cbrt(d2);
    return 1. / std::sqrt(0.25 * (2. * c + std::log((1. + e) / (1. - e)) /
                                              (e * c * c)));
  }
  if (d < 1.) {
    const double c = std::cbrt(d);
    const double e = std::sqrt(1. - d * d);
    return 1. / std::sqrt(0.5 * (c * c + std::asin(e) / (e * c)));
  }
  return 1.;
}

int ctm_adaptive_quadrature(ctm_real_fn f, void *arg, double a, double b,
                            uint32_t start_order, uint32_t max_order,
                            double abs_err, double rel_err, double *result) {
  if (!f || !result || start_order < 1) return CTM_ERR_ARGUMENT;
  bool ok = false;
  *result = adaptive_gl([&](double x) { return f(x, arg); }, a, b,
                        start_order, max_order, abs_err, rel_err, &ok);
  return ok ? CTM_OK : CTM_NOT_CONVERGED;
}

int ctm_orientation_coefficients(ctm_real_fn p, void *arg, uint32_t order,
                                 double abs_err, double rel_err,
                                 double *coefficients) {
  if (!coefficients) return CTM_ERR_ARGUMENT;
  // p(beta) of the distribution; NULL is the isotropic density 1/2
  auto density = [&](double beta) { return p ? p(beta, arg) : 0.5; };
  bool ok = true, reached = false;
  const double norm = adaptive_gl(
      [&](double beta) { return std::sin(beta) * density(beta); }, 0., M_PI,
      100, 1000, abs_err, rel_err, &reached);
  ok = ok && reached;
  const double scale = 1. / norm;
  coefficients[0] = 1.;
  for (uint32_t n = 1; n <= order; ++n) {
    coefficients[n] = adaptive_gl(
        [&](double beta) {
          const double c = std::cos(beta), s = std::sin(beta);
          return s * (scale * density(beta)) * legendre_by_wigner(c, n);
        },
        0., M_PI, 2 * (n + 1), 100 * (n + 1), abs_err, rel_err, &reached);
    ok = ok && reached;
  }
  return ok ? CTM_OK : CTM_NOT_CONVERGED;
}

double ctm_grain_memory_estimate(double radius, double axis_ratio,
                                 double wavelength, uint32_t min_order,
                                 uint32_t max_order, int32_t n_pairs) {
  if (!(radius > 0.) || !(axis_ratio > 0.) || !(wavelength > 0.)) return 0.;
  // order estimate from the largest semi-axis (a = R_V d^(2/3), b = R_V
  // d^(-1/3)), with a margin for the convergence windows
  const double amax = radius * std::max(std::pow(axis_ratio, 2. / 3.),
                                        std::pow(axis_ratio, -1. / 3.));
  const double x = 2. * M_PI * amax / wavelength;
  double N = std::ceil(x + 4.05 * std::cbrt(x)) + 8.;
  N = std::max(N, (double)min_order + 1.);
  N = std::min(N, std::max((double)max_order - 1., 1.));
  // sum over the m blocks of nm^2 (m = 0: N, m > 0: N + 1 - m)
  const double s2 = N * N + N * (N + 1.) * (2. * N + 1.) / 6.;
  // Q and RgQ (4 s2), T (2 s2), ensemble T (4 s2) complex values; Bessel
  // tables of the window orders (16 n^2 doubles each, summed up to N);
  // the observables' bilinear forms (4 complex values per direction pair
  // and (grain, m) block)
  return 16. * 10. * s2 + 8. * 16. * N * (N + 1.) * (2. * N + 1.) / 6. +
         64. * (N + 1.) * std::max(n_pairs, 0);
}

}  // extern "C"
