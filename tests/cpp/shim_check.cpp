// Exercises the C++ TMatrixCalculator / TMatrix surface of costuum_amd.hpp
// on the grains of the fixture rows given on the command line (one line
// "r d lambda mr_re mr_im" each on stdin) and prints one JSON line per
// grain: converged nmax and ngauss, the two T-matrix sums, three elements,
// Z at two angle sets, K, and the same for the isotropic ensemble.
#include <cmath>
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
    // forward scattering matrix: particle orientation, particle frame, and
    // the sine/cosine form of the particle frame
    const Matrix<std::complex<double>> S1 =
        T->get_forward_scattering_matrix(0.4, 0.9, 1.1, 2.0, 2.5, 0.3);
    const Matrix<std::complex<double>> S2 =
        T->get_forward_scattering_matrix(1.1, 2.0, 2.5, 0.3);
    const Matrix<std::complex<double>> S3 = T->get_forward_scattering_matrix(
        std::cos(1.1), std::sin(1.1), 1. / std::sin(1.1), std::cos(2.0),
        std::sin(2.0), std::cos(2.5), std::sin(2.5), 1. / std::sin(2.5),
        std::cos(0.3), std::sin(0.3));
    std::printf(", \"s\": [");
    const Matrix<std::complex<double>> *Ss[3] = {&S1, &S2, &S3};
    for (int k = 0; k < 3; ++k)
      for (int e = 0; e < 4; ++e)
        std::printf("%.17g, %.17g%s", (*Ss[k])(e / 2, e % 2).real(),
                    (*Ss[k])(e / 2, e % 2).imag(),
                    k == 2 && e == 3 ? "]" : ", ");
    double red[3];
    T->get_reduced_extinction_matrix(0.3, 0.7, 0.9, 0.2, red);
    std::printf(", \"reduced_k\": [%.17g, %.17g, %.17g]", red[0], red[1],
                red[2]);
    // absorption cross sections: explicit quadrature, pairs, C_abs only,
    // and the average over incidence
    {
      const uint_fast32_t ng = 6;
      std::vector<double> x(ng), w(ng), th(ng), ph(ng), pw(ng);
      ctm_gauss_legendre(ng, -1., 1., x.data(), w.data());
      ctm_gauss_legendre(ng, 0., 2. * M_PI, ph.data(), pw.data());
      for (uint_fast32_t i = 0; i < ng; ++i) th[i] = std::acos(x[i]);
      double c2[2];
      T->get_absorption_cross_sections(0.7, ng, true, true, th, w, ph, pw,
                                       c2);
      const double thetas[2] = {0.7, 1.9};
      double pairs[4], only[2];
      T->get_absorption_cross_sections(thetas, 2, ng, pairs);
      T->get_absorption_cross_section(thetas, 2, ng, only);
      std::printf(", \"cabs_quad\": [%.17g, %.17g], \"cabs_pairs\": "
                  "[%.17g, %.17g, %.17g, %.17g], \"cabs\": [%.17g, %.17g], "
                  "\"cabs_avg\": %.17g",
                  c2[0], c2[1], pairs[0], pairs[1], pairs[2], pairs[3],
                  only[0], only[1], T->get_average_absorption_cross_section(ng));
    }
    TMatrix *A = TMatrixCalculator::apply_orientation_distribution(
        *T, Isotropic());
    std::printf(", \"avg_qext\": %.17g, \"avg_qsca\": %.17g, ",
                A->get_extinction_coefficient(),
                A->get_scattering_coefficient());
    print_matrix("avg_z", A->get_scattering_matrix(0., 0., 0.5, 0., 1.7, 0.));
    // set_element: doubling T^(11)_{1 1 0} changes the queries accordingly
    T->set_element(0, 1, 0, 0, 1, 0, 2. * (*T)(0, 1, 0, 0, 1, 0));
    std::printf(", \"set_t\": [%.17g, %.17g], ",
                (*T)(0, 1, 0, 0, 1, 0).real(), (*T)(0, 1, 0, 0, 1, 0).imag());
    print_matrix("set_z", T->get_scattering_matrix(0.4, 0.9, 1.1, 2.0, 2.5,
                                                   0.3));
    std::printf("}\n");
    std::fflush(stdout);
    delete A;
    delete T;
  }
  return 0;
}
