// Exercises the C++ task-result surface of costuum_amd_tasks.hpp the way
// the reference's callers do (testTaskManager.cpp:127-235): TaskManager,
// add_*, QuickSched, generate_tasks, execute_tasks, then the results read
// through ResultKey::get_result_index and the result getters.  Prints one
// JSON line: the shape points and weights, the absorption and extinction
// coefficients [comp][size][lambda][theta][Qabs, Qabspol, Qext, Qextpol,
// Qextcpol], the scattering matrices [comp][size][lambda][angle][16] and
// the axes as sorted by generate_tasks.
#include <cstdio>
#include <vector>

#include "costuum_amd_tasks.hpp"

// m = 1.4 + 0.1 t + 0.05 lambda + i (0.02 + 0.01 t)
class TestDust : public DustProperties {
 public:
  std::complex<float_type> get_refractive_index(
      const float_type wavelength, const float_type,
      const int_fast32_t grain_type) const override {
    return {1.4 + 0.1 * grain_type + 0.05 * wavelength,
            0.02 + 0.01 * grain_type};
  }
};

int main() {
  DraineHensleyShapeDistribution shapes(3);
  SizeBasedAlignmentDistribution alignment(0.1, 1, 50);
  TestDust dust;
  TaskManager tm(4, 60, 2, 1.e-4, 10000000000ull, shapes, alignment, dust);
  tm.add_composition(DUSTGRAINTYPE_SILICON);
  tm.add_composition(DUSTGRAINTYPE_CARBON_PARALLEL);
  tm.add_size(0.25);  // unsorted on purpose
  tm.add_size(0.08);
  tm.add_wavelength(1.1);
  tm.add_wavelength(0.55);
  const std::vector<float_type> theta = {0.3, 1.2, 2.0};
  QuickSched quicksched(4, false, "");
  std::vector<Task *> tasks;
  std::vector<Resource *> resources;
  ResultKey *key = nullptr;
  std::vector<Result *> results;
  TMatrixAuxiliarySpaceManager *space = nullptr;
  tm.generate_tasks(theta, 12, quicksched, tasks, resources, key, results,
                    space, true, true, true, true);
  quicksched.execute_tasks();

  std::printf("{\"shapes\": [");
  for (uint_fast32_t i = 0; i < shapes.get_number_of_points(); ++i)
    std::printf("%s[%.17g, %.17g]", i ? ", " : "", shapes.get_shape(i),
                shapes.get_weight(i));
  std::printf("], \"sizes\": [%.17g, %.17g], \"wavelengths\": [%.17g, %.17g]",
              key->get_size(0), key->get_size(1), key->get_wavelength(0),
              key->get_wavelength(1));
  std::printf(", \"n_results\": %zu, \"q\": [", results.size());
  bool first = true;
  for (size_t c = 0; c < key->composition_size(); ++c)
    for (size_t a = 0; a < key->size_size(); ++a)
      for (size_t l = 0; l < key->wavelength_size(); ++l) {
        const auto &E = *static_cast<ExtinctionCoefficientResult *>(
            results[key->get_result_index(c, a, l, 0)]);
        const auto &A = *static_cast<AbsorptionCoefficientResult *>(
            results[key->get_result_index(c, a, l, 1)]);
        if (E.get_composition() != key->get_composition(c) ||
            A.get_size() != key->get_size(a) ||
            E.get_wavelength() != key->get_wavelength(l)) {
          std::fprintf(stderr, "result parameters out of order\n");
          return 1;
        }
        for (size_t j = 0; j < theta.size(); ++j) {
          std::printf("%s%.17g, %.17g, %.17g, %.17g, %.17g",
                      first ? "" : ", ", A.get_Qabs(j), A.get_Qabspol(j),
                      E.get_Qext(j), E.get_Qextpol(j), E.get_Qextcpol(j));
          first = false;
        }
      }
  std::printf("], \"z\": [");
  first = true;
  for (size_t c = 0; c < key->composition_size(); ++c)
    for (size_t a = 0; a < key->size_size(); ++a)
      for (size_t l = 0; l < key->wavelength_size(); ++l) {
        const auto &S = *static_cast<ScatteringMatrixResult *>(
            results[key->get_result_index(c, a, l, 2)]);
        for (size_t k = 0; k < theta.size() * 2 * theta.size(); ++k) {
          const Matrix<float_type> &Z = S.get_scattering_matrix(k);
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
              std::printf("%s%.17g", first ? "" : ", ", Z(i, j));
              first = false;
            }
        }
      }
  std::printf("]}\n");
  for (Task *t : tasks) delete t;
  for (Resource *r : resources) delete r;
  for (Result *r : results) delete r;
  delete key;
  delete space;
  return 0;
}
