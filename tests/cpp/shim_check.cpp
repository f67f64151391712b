// Exercises the C++ TMatrixCalculator / TMatrix surface of costuum_amd.hpp
// on the grains of the fixture rows given on the command line (one line
// "r d lambda mr_re mr_im" each on stdin) and prints one JSON line per
// grain: converged nmax and ngauss, the two T-matrix sums, three elements,
// Z at two angle sets, K, and the same for the isotropic ensemble.
#include <cstdio>
#include <vector>

#include "costuum_amd.hpp"

// Isotropic orientation distribution: p_0 = 1, p_n = 0 (its coefficients
// as OrientationDistribution::get_coefficient returns them analytically).
struct Isotropic {
  double get_coefficient(size_t n) const { return n == 0 ? 1. : 0.; }
};

static void print_matrix(const char *name, const costuum_amd_shim::Matrix4 &M) {
  std::printf("\"%s\": [", name);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      std::printf("%.17g%s", M(i, j), i == 3 && j == 3 ? "]" : ", ");
}

int main() {
  double r, d, lam, mre, mim;
  while (std::scanf("%lf %lf %lf %lf %lf", &r, &d, &lam, &mre, &mim) == 5) {
    TMatrix *T = TMatrixCalculator::calculate_TMatrix(
        1., d, r, lam, 200, 1.e-4, 2, std::complex<double>(mre, mim), 500);
    const uint_fast32_t N = T->get_nmax();
    std::printf("{\"nmax\": %u, \"ngauss\": %u, ", (unsigned)N,
                (unsigned)T->get_ngauss());
    std::printf("\"qsca\": %.17g, \"qext\": %.17g, ",
                T->get_scattering_coefficient(),
                T->get_extinction_coefficient());
    const std::complex<double> t1 = (*T)(0, 1, 0, 0, 1, 0);
    const std::complex<double> t2 = (*T)(1, 2, 1, 1, 2, 1);
    const std::complex<double> t3 = (*T)(0, N, N, 1, N, N);
    std::printf("\"t\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g], ",
                t1.real(), t1.imag(), t2.real(), t2.imag(), t3.real(),
                t3.imag());
    print_matrix("z1", T->get_scattering_matrix(0., 0., 0., 0., 1.2, 0.));
    std::printf(", ");
    print_matrix("z2", T->get_scattering_matrix(0.4, 0.9, 1.1, 2.0, 2.5, 0.3));
    std::printf(", ");
    print_matrix("k", T->get_extinction_matrix(0.3, 0.7, 0.9, 0.2));
    TMatrix *A = TMatrixCalculator::apply_orientation_distribution(
        *T, Isotropic());
    std::printf(", \"avg_qext\": %.17g, \"avg_qsca\": %.17g, ",
                A->get_extinction_coefficient(),
                A->get_scattering_coefficient());
    print_matrix("avg_z", A->get_scattering_matrix(0., 0., 0.5, 0., 1.7, 0.));
    std::printf("}\n");
    std::fflush(stdout);
    delete A;
    delete T;
  }
  return 0;
}
