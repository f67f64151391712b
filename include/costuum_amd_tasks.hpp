/* costuum_amd_tasks.hpp -- C++ surface of the reference's task-based table
 * computation (the get_table path) over the C ABI of costuum_amd.h,
 * header-only, source-compatible with callers of TaskManager
 * (CTMModule.cpp:258-342, testTaskManager.cpp:127-235):
 *
 *   TaskManager(...), add_composition/add_size/add_wavelength
 *                                          TaskManager.hpp:371-412
 *   TaskManager::generate_tasks            TaskManager.hpp:445-455
 *   QuickSched(nthreads, log, name, progress), execute_tasks
 *                                          QuickSchedWrapper.hpp:302-389
 *   Task, Resource, Computable             QuickSchedWrapper.hpp:43-243
 *   Result, ResultKey                      Result.hpp:45-110,
 *                                          ResultKey.hpp:33-155
 *   ExtinctionCoefficientResult::get_Qext/get_Qextpol/get_Qextcpol
 *                                          ExtinctionCoefficientTask.hpp:40-126
 *   AbsorptionCoefficientResult::get_Qabs/get_Qabspol
 *                                          AbsorptionCoefficientTask.hpp:40-110
 *   ScatteringMatrixResult::get_scattering_matrix
 *                                          ScatteringMatrixTask.hpp:40-102
 *   ShapeDistribution, SingleShapeShapeDistribution,
 *   DraineHensleyShapeDistribution          ShapeDistribution.hpp:36-140,
 *                                          SingleShapeShapeDistribution.hpp,
 *                                          DraineHensleyShapeDistribution.hpp
 *   OrientationDistribution (isotropic), Mishchenko-, DavisGreenstein-,
 *   DisabledAlignmentOrientationDistribution
 *                                          OrientationDistribution.hpp:38-271
 *   AlignmentDistribution, SizeBasedAlignmentDistribution
 *                                          AlignmentDistribution.hpp,
 *                                          SizeBasedAlignmentDistribution.hpp
 *   DustProperties (abstract)              DustProperties.hpp:36-70
 *
 * generate_tasks() lays out the results exactly as the reference does
 * (sizes and wavelengths sorted ascending; per (composition, size,
 * wavelength) the extinction, absorption and scattering results in that
 * order; ResultKey::get_result_index gives the position) and registers one
 * task; QuickSched::execute_tasks() runs it: the grid's grains go to the
 * GPU in batches sized to the memory budget (min(maximum_memory_usage, 80 %
 * of the free device memory), per-grain estimate ctm_grain_memory_estimate)
 * and each batch runs the task-path convergence, the orientation average of
 * the alignment distribution, the observables and the shape average on the
 * device.  The per-thread auxiliary space manager has nothing to manage
 * (the library context holds the device scratch).  The caller deletes
 * tasks, resources, results, result_key and space_manager, as with the
 * reference.  Failures print a message and abort, as ctm_error does.
 */
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <fstream>
#include <ostream>
#include <string>
#include <vector>

#include "costuum_amd.hpp"

enum DustGrainType {
  DUSTGRAINTYPE_CARBON_PARALLEL = 0,
  DUSTGRAINTYPE_CARBON_PERPENDICULAR,
  DUSTGRAINTYPE_SILICON,
  NUMBER_OF_DUSTGRAINTYPES
};

// ---------------------------------------------------------------------------
// Scheduling vocabulary (QuickSchedWrapper.hpp:43-243).

class Task {
 public:
  virtual ~Task() {}
  virtual void execute(const int_fast32_t thread_id = 0) = 0;
  virtual int_fast32_t get_cost() const { return 1; }
};

class Resource {
 public:
  virtual ~Resource() {}
};

class Computable {
 public:
  virtual ~Computable() {}
  void make_available() { _was_computed = true; }
  bool was_computed() const { return _was_computed; }

 private:
  bool _was_computed = false;
};

// Runs the registered tasks in registration order: each task is one
// device-wide computation, so the thread count has no meaning here; the
// log, when asked for, records each task's wall time.
class QuickSched {
 public:
  QuickSched(const int_fast32_t number_of_threads,
             const bool write_log = false, const std::string log_name = "",
             const bool display_progress = true)
      : _number_of_threads(number_of_threads), _write_log(write_log),
        _log_name(log_name) {
    (void)display_progress;
  }
  void register_task(Task &t) { _tasks.push_back(&t); }
  void register_resource(Resource &) {}
  void execute_tasks(int_fast32_t number_of_threads = -1) {
    (void)number_of_threads;
    std::ofstream log;
    if (_write_log) log.open(_log_name);
    for (size_t i = 0; i < _tasks.size(); ++i) {
      const auto t0 = std::chrono::steady_clock::now();
      _tasks[i]->execute(0);
      const auto t1 = std::chrono::steady_clock::now();
      if (_write_log)
        log << "0\t0\t"
            << std::chrono::duration<double>(t1 - t0).count() << "\t0\t" << i
            << "\n";
    }
  }
  void print_type_dict(std::ostream &o) const { o << "0\tGPU table\n"; }
  int_fast32_t get_number_of_threads() const { return _number_of_threads; }

 private:
  int_fast32_t _number_of_threads;
  bool _write_log;
  std::string _log_name;
  std::vector<Task *> _tasks;
};

// The reference's per-thread T-matrix scratch; the library context holds
// the device scratch instead.
class TMatrixAuxiliarySpaceManager {
 public:
  virtual ~TMatrixAuxiliarySpaceManager() {}
};

// ---------------------------------------------------------------------------
// Results (Result.hpp, ResultKey.hpp and the three result classes).

enum ResultType {
  RESULTTYPE_EXTINCTIONCOEFFICIENTS = 0,
  RESULTTYPE_ABSORPTIONCOEFFICIENTS,
  RESULTTYPE_SCATTERINGMATRIX
};

class Result : public Resource {
 public:
  Result(int_fast32_t composition, float_type size, float_type wavelength,
         int_fast32_t type)
      : _composition(composition), _size(size), _wavelength(wavelength),
        _type(type) {}
  virtual ~Result() {}
  int_fast32_t get_composition() const { return _composition; }
  float_type get_size() const { return _size; }
  float_type get_wavelength() const { return _wavelength; }
  int_fast32_t get_type() const { return _type; }

 private:
  int_fast32_t _composition;
  float_type _size, _wavelength;
  int_fast32_t _type;
};

namespace costuum_amd_shim {
class TableTask;
[[noreturn]] inline void index_fail(const char *what) {
  fail(what, nullptr);
}
}  // namespace costuum_amd_shim

class ExtinctionCoefficientResult : public Result {
 public:
  ExtinctionCoefficientResult(int_fast32_t composition, float_type size,
                              float_type wavelength,
                              uint_fast32_t number_of_angles)
      : Result(composition, size, wavelength,
               RESULTTYPE_EXTINCTIONCOEFFICIENTS),
        _Qext(number_of_angles, 0.), _Qextpol(number_of_angles, 0.),
        _Qextcpol(number_of_angles, 0.) {}
  float_type get_Qext(uint_fast32_t i) const { return at(_Qext, i); }
  float_type get_Qextpol(uint_fast32_t i) const { return at(_Qextpol, i); }
  float_type get_Qextcpol(uint_fast32_t i) const { return at(_Qextcpol, i); }

 private:
  friend class costuum_amd_shim::TableTask;
  static float_type at(const std::vector<float_type> &v, uint_fast32_t i) {
    if (i >= v.size()) costuum_amd_shim::index_fail("result index");
    return v[i];
  }
  std::vector<float_type> _Qext, _Qextpol, _Qextcpol;
};

class AbsorptionCoefficientResult : public Result {
 public:
  AbsorptionCoefficientResult(int_fast32_t composition, float_type size,
                              float_type wavelength,
                              uint_fast32_t number_of_angles)
      : Result(composition, size, wavelength,
               RESULTTYPE_ABSORPTIONCOEFFICIENTS),
        _Qabs(number_of_angles, 0.), _Qabspol(number_of_angles, 0.) {}
  float_type get_Qabs(uint_fast32_t i) const { return at(_Qabs, i); }
  float_type get_Qabspol(uint_fast32_t i) const { return at(_Qabspol, i); }

 private:
  friend class costuum_amd_shim::TableTask;
  static float_type at(const std::vector<float_type> &v, uint_fast32_t i) {
    if (i >= v.size()) costuum_amd_shim::index_fail("result index");
    return v[i];
  }
  std::vector<float_type> _Qabs, _Qabspol;
};

// Z for every (theta, phi) of the scattering grid, index itheta * nphi +
// iphi (ScatteringMatrixTask.hpp:719-720).
class ScatteringMatrixResult : public Result {
 public:
  ScatteringMatrixResult(int_fast32_t composition, float_type size,
                         float_type wavelength,
                         uint_fast32_t number_of_angles)
      : Result(composition, size, wavelength, RESULTTYPE_SCATTERINGMATRIX),
        _Z(number_of_angles, Matrix<float_type>(4, 4)) {}
  const Matrix<float_type> &get_scattering_matrix(uint_fast32_t i) const {
    if (i >= _Z.size()) costuum_amd_shim::index_fail("result index");
    return _Z[i];
  }

 private:
  friend class costuum_amd_shim::TableTask;
  std::vector<Matrix<float_type>> _Z;
};

class ResultKey {
 public:
  ResultKey(const std::vector<int_fast32_t> &compositions,
            const std::vector<float_type> &sizes,
            const std::vector<float_type> &wavelengths,
            const uint_fast32_t number_of_results = 1)
      : _compositions(compositions), _sizes(sizes), _wavelengths(wavelengths),
        _number_of_results(number_of_results) {}
  size_t composition_size() const { return _compositions.size(); }
  size_t size_size() const { return _sizes.size(); }
  size_t wavelength_size() const { return _wavelengths.size(); }
  uint_fast32_t results_size() const { return _number_of_results; }
  int_fast32_t get_composition(size_t i) const { return _compositions.at(i); }
  float_type get_size(size_t i) const { return _sizes.at(i); }
  float_type get_wavelength(size_t i) const { return _wavelengths.at(i); }
  // composition-major, then size, wavelength, result (ResultKey.hpp:
  // 143-155)
  size_t get_result_index(const size_t composition_index,
                          const size_t size_index,
                          const size_t wavelength_index,
                          const uint_fast32_t result_index = 0) const {
    if (composition_index >= _compositions.size() ||
        size_index >= _sizes.size() ||
        wavelength_index >= _wavelengths.size() ||
        result_index >= _number_of_results)
      costuum_amd_shim::index_fail("ResultKey index out of range");
    return ((composition_index * _sizes.size() + size_index) *
                _wavelengths.size() +
            wavelength_index) *
               _number_of_results +
           result_index;
  }

 private:
  std::vector<int_fast32_t> _compositions;
  std::vector<float_type> _sizes, _wavelengths;
  uint_fast32_t _number_of_results;
};

// ---------------------------------------------------------------------------
// Shape distributions.

namespace costuum_amd_shim {
// Adapter: a member function of an object as a ctm_real_fn.
template <class F>
double call_fn(double x, void *arg) {
  return (*static_cast<const F *>(arg))(x);
}
template <class F>
double adaptive(const F &f, double a, double b, uint32_t start,
                uint32_t maxo, double abs_err, double rel_err) {
  double r = 0.;
  const int rc = ctm_adaptive_quadrature(&call_fn<F>, const_cast<F *>(&f), a,
                                         b, start, maxo, abs_err, rel_err, &r);
  if (rc < 0) fail("ctm_adaptive_quadrature", nullptr);
  return r;
}
}  // namespace costuum_amd_shim

// Gauss-Legendre points on [d_min, d_max] weighted by the density
// (ShapeDistribution.hpp:36-140); the default density is a Gaussian around
// d = 1 on [0.5, 1.5].
class ShapeDistribution {
 public:
  virtual ~ShapeDistribution() {}
  virtual float_type operator()(const float_type axis_ratio) const {
    if (axis_ratio >= 0.5 && axis_ratio <= 1.5) {
      const float_type x = axis_ratio - 1.;
      return std::exp(-0.5 * x * x / 0.027777778);
    }
    return 0.;
  }
  virtual float_type get_minimum_axis_ratio() const { return 0.5; }
  virtual float_type get_maximum_axis_ratio() const { return 1.5; }
  void evaluate(const uint_fast32_t npoints) {
    _shapes.assign(npoints, 0.);
    _weights.assign(npoints, 0.);
    costuum_amd_shim::check(
        ctm_gauss_legendre((uint32_t)npoints, get_minimum_axis_ratio(),
                           get_maximum_axis_ratio(), _shapes.data(),
                           _weights.data()),
        "ctm_gauss_legendre", nullptr);
    for (uint_fast32_t i = 0; i < npoints; ++i)
      _weights[i] *= (*this)(_shapes[i]);
  }
  uint_fast32_t get_number_of_points() const { return _shapes.size(); }
  float_type get_shape(uint_fast32_t i) const { return _shapes.at(i); }
  float_type get_weight(uint_fast32_t i) const { return _weights.at(i); }

 protected:
  std::vector<float_type> _shapes, _weights;
};

// One axis ratio with weight 1 (SingleShapeShapeDistribution.hpp:34-76).
class SingleShapeShapeDistribution : public ShapeDistribution {
 public:
  explicit SingleShapeShapeDistribution(const float_type axis_ratio) {
    _shapes.assign(1, axis_ratio);
    _weights.assign(1, 1.);
  }
  virtual float_type operator()(const float_type d) const {
    return d == _shapes[0] ? 1. : 0.;
  }
  virtual float_type get_minimum_axis_ratio() const { return _shapes[0]; }
  virtual float_type get_maximum_axis_ratio() const { return _shapes[0]; }
};

// Draine & Hensley (2017) CDE2 density 12 L (1 - L)^2 |dL/dd| in the axis
// ratio, limited where it drops below `cutoff` or to the central `fraction`
// of the prolate and oblate parts (DraineHensleyShapeDistribution.hpp:
// 67-370).
class DraineHensleyShapeDistribution : public ShapeDistribution {
 public:
  DraineHensleyShapeDistribution(const uint_fast32_t npoints,
                                 const float_type cutoff = 0.15,
                                 const float_type fraction = -1.) {
    if (fraction > 0.) {
      if (fraction >= 1.)
        costuum_amd_shim::fail("fraction must be below 1", nullptr);
      limits_from_fraction(fraction);
    } else {
      if (!(cutoff > 0. && cutoff < 0.5))
        costuum_amd_shim::fail("cutoff must lie in (0, 0.5)", nullptr);
      limits_from_cutoff(cutoff);
    }
    evaluate(npoints);
  }
  virtual float_type operator()(const float_type d) const {
    return (d >= _dmin && d <= _dmax) ? cde2(d) : 0.;
  }
  virtual float_type get_minimum_axis_ratio() const { return _dmin; }
  virtual float_type get_maximum_axis_ratio() const { return _dmax; }

  // depolarisation factor L(d) (DraineHensleyShapeDistribution.hpp:67-84)
  static double shape_factor(double d) {
    if (d < 1.) {
      const double e2 = 1. - d * d, e = std::sqrt(e2);
      return (1. - e2) * (0.5 * std::log((1. + e) / (1. - e)) / e - 1.) / e2;
    }
    if (d > 1.) {
      const double e2 = 1. - 1. / (d * d), e = std::sqrt(e2);
      return (1. - std::sqrt(1. - e2) * std::asin(e) / e) / e2;
    }
    return 1. / 3.;
  }
  // |dL/dd| (DraineHensleyShapeDistribution.hpp:93-121)
  static double jacobian(double d) {
    if (d < 1.) {
      const double e2 = 1. - d * d;
      if (e2 < 1.e-5) return 4. / 15. + 2. * e2 / 21. + 3. * e2 * e2 / 70.;
      const double e = std::sqrt(e2);
      return 0.5 * std::sqrt(1. - e2) *
             ((3. - e2) * std::log((1. + e) / (1. - e)) - 6. * e) /
             (e2 * e2 * e);
    }
    if (d > 1.) {
      const double e2 = 1. - 1. / (d * d);
      if (e2 < 1.e-5) return 4. / 15. - 2. * e2 / 21. - 11. * e2 * e2 / 210.;
      const double e = std::sqrt(e2);
      return (1. - e2) *
             ((3. - 2. * e2) * std::asin(e) - 3. * e * std::sqrt(1. - e2)) /
             (e2 * e2 * e);
    }
    return 4. / 15.;
  }
  static double cde2(double d) {
    const double L = shape_factor(d);
    return 12. * L * (L - 1.) * (L - 1.) * jacobian(d);
  }

 private:
  double _dmin = 1., _dmax = 1.;

  static double q(double a, double b) {
    auto f = [](double d) { return cde2(d); };
    return costuum_amd_shim::adaptive(f, a, b, 10, 1000, 1.e-10, 1.e-10);
  }
  void limits_from_fraction(double fraction) {
    const double prolate = q(0., 1.);
    double target = fraction * prolate;
    double lo = 0., hi = 1., half = 0.5, integral = q(half, 1.);
    while (std::fabs(integral - target) > 0.005 * std::fabs(integral + target)) {
      if (half <= 0.) costuum_amd_shim::fail("shape limit search", nullptr);
      if (integral > target) lo = half;
      else hi = half;
      half = 0.5 * (lo + hi);
      integral = q(half, 1.);
    }
    _dmin = half;
    target = fraction * (1. - prolate);
    lo = 1.;
    hi = 2.;
    integral = q(1., hi);
    while (integral < target) {
      hi *= 2.;
      integral = q(1., hi);
    }
    half = 0.5 * (lo + hi);
    integral = q(1., half);
    while (std::fabs(integral - target) > 1.e-5 * std::fabs(integral + target)) {
      if (half <= 0.) costuum_amd_shim::fail("shape limit search", nullptr);
      if (integral > target) hi = half;
      else lo = half;
      half = 0.5 * (lo + hi);
      integral = q(1., half);
    }
    _dmax = half;
  }
  void limits_from_cutoff(double cutoff) {
    double val = cde2(_dmin) - cutoff;
    while (val > 0.) {
      _dmin *= 0.5;
      val = cde2(_dmin) - cutoff;
    }
    double upper = 1.;
    if (!(cde2(upper) - cutoff > 0.))
      costuum_amd_shim::fail("cutoff above the density at d = 1", nullptr);
    while (std::fabs(_dmin - upper) > 0.005 * std::fabs(_dmin + upper)) {
      const double mid = 0.5 * (_dmin + upper);
      if (cde2(mid) - cutoff < 0.) _dmin = mid;
      else upper = mid;
    }
    val = cde2(_dmax) - cutoff;
    while (val > 0.) {
      _dmax += 1.;
      val = cde2(_dmax) - cutoff;
    }
    double lower = 1.;
    while (std::fabs(_dmax - lower) > 0.005 * std::fabs(_dmax + lower)) {
      const double mid = 0.5 * (lower + _dmax);
      if (cde2(mid) - cutoff > 0.) lower = mid;
      else _dmax = mid;
    }
  }
};

// ---------------------------------------------------------------------------
// Orientation and alignment distributions.

// p(beta) with Legendre coefficients p_n = int sin(b) p(b) d^n_00(b) db
// after normalisation (OrientationDistribution.hpp:38-271); the base class
// is the isotropic density 1/2 (exported to Python as
// RandomOrientationDistribution).
class OrientationDistribution {
 public:
  explicit OrientationDistribution(const uint_fast32_t nmax)
      : _coefficients(nmax + 1, 0.) {}
  virtual ~OrientationDistribution() {}
  virtual float_type operator()(const float_type beta,
                                const float_type cosbeta,
                                const float_type sinbeta) const {
    (void)beta;
    (void)cosbeta;
    (void)sinbeta;
    return 0.5;
  }
  // numerical coefficients, adaptive Gauss-Legendre quadrature
  // (OrientationDistribution.hpp:220-243)
  void initialise(const float_type absolute_error = 1.e-10,
                  const float_type relative_error = 1.e-5) {
    auto f = [this](double b) {
      return (*this)(b, std::cos(b), std::sin(b));
    };
    const int rc = ctm_orientation_coefficients(
        &costuum_amd_shim::call_fn<decltype(f)>, &f,
        (uint32_t)get_maximum_order(), absolute_error, relative_error,
        _coefficients.data());
    if (rc < 0) costuum_amd_shim::fail("ctm_orientation_coefficients", nullptr);
  }
  float_type get_coefficient(const uint_fast32_t n) const {
    return _coefficients.at(n);
  }
  uint_fast32_t get_maximum_order() const { return _coefficients.size() - 1; }
  virtual bool compute_alignment() const { return true; }
  const std::vector<float_type> &coefficients() const { return _coefficients; }

 protected:
  std::vector<float_type> _coefficients;
};

// p_0 = 1, p_2 = 3/2 <cos^2 beta> - 1/2 (MishchenkoOrientationDistribution
// .hpp:65-122).
class MishchenkoOrientationDistribution : public OrientationDistribution {
 public:
  MishchenkoOrientationDistribution(const uint_fast32_t nmax,
                                    const float_type cos2beta)
      : OrientationDistribution(nmax), _p2(1.5 * cos2beta - 0.5) {
    if (cos2beta < 0.2 || cos2beta > 0.6)
      costuum_amd_shim::fail("cos2beta must lie in [0.2, 0.6]", nullptr);
    _coefficients[0] = 1.;
    if (nmax >= 2) _coefficients[2] = _p2;
  }
  virtual float_type operator()(const float_type, const float_type cosbeta,
                                const float_type) const {
    return 0.5 + 1.25 * _p2 * (3. * cosbeta * cosbeta - 1.);
  }

 private:
  float_type _p2;
};

// Perfect alignment: oblate grains along the field (p_n = 1), prolate
// grains perpendicular (p_2i = (-1)^i (2i - 1)!! / (2i)!!)
// (DavisGreensteinOrientationDistribution.hpp:66-105).
class DavisGreensteinOrientationDistribution : public OrientationDistribution {
 public:
  DavisGreensteinOrientationDistribution(const uint_fast32_t nmax,
                                         const float_type axis_ratio)
      : OrientationDistribution(nmax) {
    if (axis_ratio < 1.) {
      double num = 1., den = 2.;
      _coefficients[0] = 1.;
      for (uint_fast32_t i = 1; 2 * i < _coefficients.size(); ++i) {
        _coefficients[2 * i] = ((i % 2 == 0) ? 1. : -1.) * num / den;
        num *= 2. * i + 1.;
        den *= 2. * (i + 1.);
      }
    } else {
      std::fill(_coefficients.begin(), _coefficients.end(), 1.);
    }
  }
};

// Alignment switched off: the ensemble T-matrix is the grain's own
// (DisabledAlignmentOrientationDistribution.hpp:36-81).
class DisabledAlignmentOrientationDistribution
    : public OrientationDistribution {
 public:
  DisabledAlignmentOrientationDistribution() : OrientationDistribution(0) {}
  virtual bool compute_alignment() const { return false; }
};

class AlignmentDistribution {
 public:
  virtual ~AlignmentDistribution() {}
  virtual const OrientationDistribution &get_distribution(
      const float_type equal_volume_radius,
      const float_type axis_ratio) const = 0;
};

// Grains larger than minimum_size aligned (oblate / prolate distribution of
// the chosen type: 0 Davis-Greenstein, 1 Mishchenko with <cos^2> 0.6 / 0.2,
// 2 disabled, 3 the given ones), smaller ones isotropic
// (SizeBasedAlignmentDistribution.hpp:39-147).
class SizeBasedAlignmentDistribution : public AlignmentDistribution {
 public:
  SizeBasedAlignmentDistribution(
      const float_type minimum_size,
      const int_fast32_t aligned_orientation_distribution_type,
      uint_fast32_t nmax,
      OrientationDistribution *oblate_aligned_orientation_distribution =
          nullptr,
      OrientationDistribution *prolate_aligned_orientation_distribution =
          nullptr)
      : _minimum_size(minimum_size) {
    _non_aligned = new OrientationDistribution(2 * nmax);
    _non_aligned->initialise();
    switch (aligned_orientation_distribution_type) {
      case 0:
        _oblate = new DavisGreensteinOrientationDistribution(2 * nmax, 2.);
        _prolate = new DavisGreensteinOrientationDistribution(2 * nmax, 0.5);
        break;
      case 1:
        _oblate = new MishchenkoOrientationDistribution(2 * nmax, 0.6);
        _prolate = new MishchenkoOrientationDistribution(2 * nmax, 0.2);
        break;
      case 2:
        _oblate = new DisabledAlignmentOrientationDistribution();
        _prolate = new DisabledAlignmentOrientationDistribution();
        break;
      case 3:
        if (!oblate_aligned_orientation_distribution ||
            !prolate_aligned_orientation_distribution)
          costuum_amd_shim::fail("custom alignment needs both distributions",
                                 nullptr);
        _oblate =
            new OrientationDistribution(*oblate_aligned_orientation_distribution);
        _prolate = new OrientationDistribution(
            *prolate_aligned_orientation_distribution);
        break;
      default:
        costuum_amd_shim::fail("Unknown orientation distribution!", nullptr);
    }
  }
  virtual ~SizeBasedAlignmentDistribution() {
    delete _non_aligned;
    delete _oblate;
    delete _prolate;
  }
  virtual const OrientationDistribution &get_distribution(
      const float_type equal_volume_radius,
      const float_type axis_ratio) const {
    if (equal_volume_radius > _minimum_size)
      return axis_ratio > 1. ? *_oblate : *_prolate;
    return *_non_aligned;
  }

 private:
  float_type _minimum_size;
  OrientationDistribution *_non_aligned = nullptr, *_oblate = nullptr,
                          *_prolate = nullptr;
};

// m(lambda, size, type) of a grain (DustProperties.hpp:36-70).
class DustProperties {
 public:
  virtual ~DustProperties() {}
  virtual std::complex<float_type> get_refractive_index(
      const float_type wavelength, const float_type grain_size,
      const int_fast32_t grain_type) const = 0;
};

// ---------------------------------------------------------------------------
// The table computation.

namespace costuum_amd_shim {

struct TableSetup {
  uint_fast32_t minimum_order, maximum_order, glfac;
  double tolerance;
  size_t maximum_memory_usage;
  const ShapeDistribution *shapes;
  const AlignmentDistribution *alignment;
  const DustProperties *dust;
  std::vector<int_fast32_t> compositions;
  std::vector<double> sizes, wavelengths, theta;
  uint_fast32_t ngauss;
  bool do_extinction, do_absorption, do_scattering, account_for_scattering;
  bool verbose;
};

// One task for the whole grid: convergence, orientation average,
// observables and shape average of every grain on the GPU, in batches
// that fit the memory budget.
class TableTask : public Task {
 public:
  TableTask(const TableSetup &setup, std::vector<Result *> results)
      : _s(setup), _results(std::move(results)) {}

  virtual void execute(const int_fast32_t = 0) {
    const TableSetup &s = _s;
    const size_t nshape = s.shapes->get_number_of_points();
    if (nshape == 0) fail("the shape distribution has no points", nullptr);
    std::vector<double> shapes(nshape), weights(nshape);
    for (size_t i = 0; i < nshape; ++i) {
      shapes[i] = s.shapes->get_shape(i);
      weights[i] = s.shapes->get_weight(i);
    }
    // interactions in the reference's loop order, one dust callback each
    struct Inter {
      size_t comp, size, lam;
      std::complex<double> mr;
    };
    std::vector<Inter> inter;
    for (size_t c = 0; c < s.compositions.size(); ++c)
      for (size_t a = 0; a < s.sizes.size(); ++a)
        for (size_t l = 0; l < s.wavelengths.size(); ++l)
          inter.push_back({c, a, l,
                           s.dust->get_refractive_index(
                               s.wavelengths[l], s.sizes[a],
                               s.compositions[c])});
    const int nth = (int)s.theta.size();
    const int nst = s.do_scattering ? nth : 0;
    const int ng = s.account_for_scattering ? (int)s.ngauss : 0;
    const int n_pairs = nth * (1 + ng) + nst;
    ctm_context *ctx = nullptr;
    check(ctm_create(&ctx, 0), "ctm_create", nullptr);
    size_t free_b = 0, total_b = 0;
    check_hip(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const double budget =
        std::min((double)s.maximum_memory_usage, 0.8 * (double)free_b);
    size_t i0 = 0;
    while (i0 < inter.size()) {
      // batch: interactions (each with all its shapes) within the budget,
      // at most 4096 grains
      size_t i1 = i0;
      double bytes = 0.;
      while (i1 < inter.size() && (i1 - i0 + 1) * nshape <= 4096) {
        double b = 0.;
        for (size_t k = 0; k < nshape; ++k)
          b += ctm_grain_memory_estimate(
              s.sizes[inter[i1].size], shapes[k],
              s.wavelengths[inter[i1].lam], (uint32_t)s.minimum_order,
              (uint32_t)s.maximum_order, n_pairs);
        if (i1 > i0 && bytes + b > budget) break;
        bytes += b;
        ++i1;
      }
      if (s.verbose)
        std::fprintf(stderr, "costuum_amd: interactions %zu..%zu of %zu\n",
                     i0, i1, inter.size());
      run_batch(ctx, inter, i0, i1, shapes, weights, nth, nst, ng);
      i0 = i1;
    }
    (void)ctm_destroy(ctx);
  }

 private:
  TableSetup _s;
  std::vector<Result *> _results;

  template <class InterVec>
  void run_batch(ctm_context *ctx, const InterVec &inter, size_t i0,
                 size_t i1, const std::vector<double> &shapes,
                 const std::vector<double> &weights, int nth, int nst,
                 int ng) {
    const TableSetup &s = _s;
    const size_t nshape = shapes.size(), ni = i1 - i0, n = ni * nshape;
    std::vector<double> h(5 * n);
    for (size_t i = 0; i < ni; ++i)
      for (size_t k = 0; k < nshape; ++k) {
        const size_t g = i * nshape + k;
        const auto &it = inter[i0 + i];
        h[g] = s.sizes[it.size];
        h[n + g] = shapes[k];
        h[2 * n + g] = s.wavelengths[it.lam];
        h[3 * n + g] = it.mr.real();
        h[4 * n + g] = it.mr.imag();
      }
    DeviceDoubles params(5 * n), sums(2 * n);
    check_hip(hipMemcpy(params.p, h.data(), h.size() * sizeof(double),
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    uint32_t *d_nmax = nullptr;
    int32_t *d_status = nullptr;
    check_hip(hipMalloc(reinterpret_cast<void **>(&d_nmax), 8 * n),
              "hipMalloc");
    d_status = reinterpret_cast<int32_t *>(d_nmax + n);
    ctm_convergence conv{(uint32_t)s.minimum_order, (uint32_t)s.maximum_order,
                         (uint32_t)s.glfac, CTM_PRECISION_FP64, s.tolerance};
    const int rc = ctm_tmatrix_batch(
        ctx, (int)n, params.p, params.p + n, params.p + 2 * n,
        params.p + 3 * n, params.p + 4 * n, &conv, d_nmax, d_status, sums.p,
        sums.p + n);
    (void)hipFree(d_nmax);
    check(rc, "ctm_tmatrix_batch", ctx);
    std::vector<uint32_t> nmax(n), ngs(n);
    std::vector<int32_t> status(n);
    check(ctm_batch_orders(ctx, nmax.data(), ngs.data(), status.data()),
          "ctm_batch_orders", ctx);
    // orientation average per grain (TaskManager.hpp:1112-1120)
    std::vector<const OrientationDistribution *> rows;
    std::vector<int32_t> dist(n, -1);
    uint32_t nmax_all = 1;
    for (size_t g = 0; g < n; ++g) nmax_all = std::max(nmax_all, nmax[g]);
    for (size_t g = 0; g < n; ++g) {
      const OrientationDistribution &od =
          s.alignment->get_distribution(h[g], h[n + g]);
      if (!od.compute_alignment()) continue;
      if (od.get_maximum_order() < 2 * nmax[g])
        fail("orientation distribution order below 2 nmax", nullptr);
      size_t r = 0;
      while (r < rows.size() && rows[r] != &od) ++r;
      if (r == rows.size()) rows.push_back(&od);
      dist[g] = (int32_t)r;
    }
    if (!rows.empty()) {
      size_t ncoef = 2 * (size_t)nmax_all + 1;
      for (const auto *r : rows)
        ncoef = std::max(ncoef, (size_t)r->get_maximum_order() + 1);
      std::vector<double> pN(rows.size() * ncoef, 0.);
      for (size_t r = 0; r < rows.size(); ++r)
        for (size_t k = 0; k <= rows[r]->get_maximum_order(); ++k)
          pN[r * ncoef + k] = rows[r]->get_coefficient(k);
      check(ctm_orientation_average(ctx, (int)rows.size(), (int)ncoef,
                                    pN.data(), dist.data()),
            "ctm_orientation_average", ctx);
    }
    // observables and the shape average (ShapeAveragingTask.hpp:127-166)
    const size_t zsz = (size_t)nst * 2 * nst * 16;
    DeviceDoubles q(5 * n * (size_t)std::max(nth, 1)),
        z(n * std::max<size_t>(zsz, 1));
    ctm_observable_spec spec{nth,
                             s.theta.data(),
                             s.do_extinction ? 1 : 0,
                             s.do_absorption ? 1 : 0,
                             s.account_for_scattering ? 1 : 0,
                             std::max(ng, 1),
                             nst,
                             1};
    double *qp[5];
    for (int k = 0; k < 5; ++k) qp[k] = q.p + (size_t)k * n * nth;
    check(ctm_observables(ctx, &spec, qp[0], qp[1], qp[2], qp[3], qp[4],
                          nst > 0 ? z.p : nullptr),
          "ctm_observables", ctx);
    DeviceDoubles w(n), qa(5 * ni * (size_t)std::max(nth, 1)),
        za(ni * std::max<size_t>(zsz, 1));
    {
      std::vector<double> hw(n);
      for (size_t g = 0; g < n; ++g) hw[g] = weights[g % nshape];
      check_hip(hipMemcpy(w.p, hw.data(), n * sizeof(double),
                          hipMemcpyHostToDevice),
                "hipMemcpy");
    }
    for (int k = 0; k < 5; ++k)
      check(ctm_weighted_reduce(ctx, (int)ni, (int)nshape, nth, qp[k], w.p, 1,
                                qa.p + (size_t)k * ni * nth),
            "ctm_weighted_reduce", ctx);
    if (nst > 0)
      check(ctm_weighted_reduce(ctx, (int)ni, (int)nshape, (int64_t)zsz, z.p,
                                w.p, 1, za.p),
            "ctm_weighted_reduce", ctx);
    check(ctm_synchronize(ctx), "ctm_synchronize", ctx);
    const std::vector<double> hq = qa.host(5 * ni * nth);
    const std::vector<double> hz =
        nst > 0 ? za.host(ni * zsz) : std::vector<double>();
    const uint_fast32_t nres = (s.do_extinction ? 1 : 0) +
                               (s.do_absorption ? 1 : 0) +
                               (s.do_scattering ? 1 : 0);
    for (size_t i = 0; i < ni; ++i) {
      const auto &it = inter[i0 + i];
      size_t r = ((it.comp * s.sizes.size() + it.size) * s.wavelengths.size() +
                  it.lam) *
                 nres;
      auto col = [&](int k, int j) { return hq[((size_t)k * ni + i) * nth + j]; };
      if (s.do_extinction) {
        auto *R = static_cast<ExtinctionCoefficientResult *>(_results[r++]);
        for (int j = 0; j < nth; ++j) {
          R->_Qext[j] = col(0, j);
          R->_Qextpol[j] = col(1, j);
          R->_Qextcpol[j] = col(2, j);
        }
      }
      if (s.do_absorption) {
        auto *R = static_cast<AbsorptionCoefficientResult *>(_results[r++]);
        for (int j = 0; j < nth; ++j) {
          R->_Qabs[j] = col(3, j);
          R->_Qabspol[j] = col(4, j);
        }
      }
      if (s.do_scattering) {
        auto *R = static_cast<ScatteringMatrixResult *>(_results[r++]);
        for (size_t a = 0; a < (size_t)nst * 2 * nst; ++a)
          for (int e = 0; e < 16; ++e)
            R->_Z[a](e / 4, e % 4) = hz[i * zsz + a * 16 + e];
      }
    }
  }
};

}  // namespace costuum_amd_shim

// The reference's TaskManager (TaskManager.hpp:101-1330) with the same
// constructor, parameter lists and generate_tasks signature.
class TaskManager {
 public:
  TaskManager(const uint_fast32_t minimum_order,
              const uint_fast32_t maximum_order,
              const uint_fast32_t gauss_legendre_factor,
              const float_type tolerance, const size_t maximum_memory_usage,
              const ShapeDistribution &shape_distribution,
              const AlignmentDistribution &alignment_distribution,
              const DustProperties &dust_properties)
      : _minimum_order(minimum_order), _maximum_order(maximum_order),
        _glfac(gauss_legendre_factor), _tolerance(tolerance),
        _maximum_memory_usage(maximum_memory_usage),
        _shape_distribution(shape_distribution),
        _alignment_distribution(alignment_distribution),
        _dust_properties(dust_properties) {}

  void add_composition(const int_fast32_t composition) {
    _compositions.push_back(composition);
  }
  void add_size(const float_type size) { _sizes.push_back(size); }
  void add_wavelength(const float_type wavelength) {
    _wavelengths.push_back(wavelength);
  }

  void generate_tasks(const std::vector<float_type> &theta,
                      const uint_fast32_t ngauss, QuickSched &quicksched,
                      std::vector<Task *> &tasks,
                      std::vector<Resource *> &resources,
                      ResultKey *&result_key, std::vector<Result *> &results,
                      TMatrixAuxiliarySpaceManager *&space_manager,
                      const bool do_extinction, const bool do_absorption,
                      const bool do_scattering,
                      const bool account_for_scattering,
                      const bool verbose = false,
                      const bool write_memory_log = false,
                      const std::string memory_log_file_name = "") {
    std::sort(_sizes.begin(), _sizes.end());
    std::sort(_wavelengths.begin(), _wavelengths.end());
    const uint_fast32_t nres = (do_extinction ? 1 : 0) +
                               (do_absorption ? 1 : 0) +
                               (do_scattering ? 1 : 0);
    result_key = new ResultKey(_compositions, _sizes, _wavelengths, nres);
    const uint_fast32_t nth = theta.size();
    results.clear();
    for (int_fast32_t c : _compositions)
      for (float_type a : _sizes)
        for (float_type l : _wavelengths) {
          if (do_extinction)
            results.push_back(new ExtinctionCoefficientResult(c, a, l, nth));
          if (do_absorption)
            results.push_back(new AbsorptionCoefficientResult(c, a, l, nth));
          if (do_scattering)
            results.push_back(
                new ScatteringMatrixResult(c, a, l, nth * 2 * nth));
        }
    costuum_amd_shim::TableSetup setup{
        _minimum_order, _maximum_order, _glfac, _tolerance,
        _maximum_memory_usage, &_shape_distribution, &_alignment_distribution,
        &_dust_properties, _compositions, _sizes, _wavelengths, theta, ngauss,
        do_extinction, do_absorption, do_scattering, account_for_scattering,
        verbose};
    Task *t = new costuum_amd_shim::TableTask(setup, results);
    quicksched.register_task(*t);
    tasks.clear();
    tasks.push_back(t);
    resources.clear();
    space_manager = new TMatrixAuxiliarySpaceManager();
    if (write_memory_log) {
      std::ofstream log(memory_log_file_name);
      log << "# device memory budget (bytes)\t" << _maximum_memory_usage
          << "\n";
    }
  }

 private:
  uint_fast32_t _minimum_order, _maximum_order, _glfac;
  float_type _tolerance;
  size_t _maximum_memory_usage;
  const ShapeDistribution &_shape_distribution;
  const AlignmentDistribution &_alignment_distribution;
  const DustProperties &_dust_properties;
  std::vector<int_fast32_t> _compositions;
  std::vector<float_type> _sizes, _wavelengths;
};
