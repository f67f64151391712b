/*
 * ctm_oracle.h -- CPU FP64 restatement of CosTuuM's T-matrix hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the CPU baseline timed by bench.py.  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it; the product library
 * (costuum_amd) never links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Complex values
 * cross the ABI as interleaved (re, im) doubles.
 *
 * T-matrix layout ("reference block layout", TMatrix.hpp:1028-1046 and
 * TMatrixResource.hpp:195-247): for m = 0..nmax one row-major block of
 * (2 nm) x (2 nm) complex values, nm = nmax (m = 0) or nmax + 1 - m (m > 0),
 * blocks concatenated in m order.  Row/column index i*nm + (n - nmin) with
 * i in {0, 1} and nmin = max(m, 1).
 */
#ifndef CTM_ORACLE_H
#define CTM_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* a1: SpecialFunctions.hpp:647-715 */
void ctmo_gauss_legendre(unsigned order, double *x, double *w);
/* a1: SpecialFunctions.hpp:738-751 */
void ctmo_gauss_legendre_ab(unsigned order, double a, double b, double *x,
                            double *w);
/* a4: SpecialFunctions.hpp:558-583 */
double ctmo_rv_over_rs(double axis_ratio);
/* a3: SpecialFunctions.hpp:873-893 */
void ctmo_r_dr_spheroid(const double *costheta, unsigned size,
                        double axis_ratio, double *r2, double *dr_over_r);
/* a6: SpecialFunctions.hpp:202-240 (real and complex argument) */
void ctmo_sph_j_real(unsigned nmax, double z, double *j, double *dj);
void ctmo_sph_j_cplx(unsigned nmax, const double *z, double *j, double *dj);
/* a7: SpecialFunctions.hpp:107-140 */
void ctmo_sph_y_real(unsigned nmax, double z, double *y, double *dy);
/* a9: SpecialFunctions.hpp:303-371 */
void ctmo_wigner_dn_0m(double cosx, unsigned nmax, unsigned m, double *y,
                       double *dy);
/* a22: SpecialFunctions.hpp:398-495 */
void ctmo_wigner_dn_0m_sinx(double cosx, double sinx, double sinx_inv,
                            unsigned nmax, unsigned m, double *y, double *dy);
/* a13: Matrix.hpp:261-384, in place on a row-major n x n complex matrix */
void ctmo_plu_inverse(unsigned n, double *A);
/* a20: SpecialFunctions.hpp:973-1108 */
double ctmo_clebsch_gordan(int n1, int m1, int n2, int m2, int N, int M);

/* Number of complex entries of a full T-matrix of order nmax. */
size_t ctmo_tmatrix_size(unsigned nmax);
/* Offset (in complex entries) of the m block. */
size_t ctmo_tmatrix_offset(unsigned nmax, unsigned m);

/*
 * a8 + a11/a12: one m block of the T-matrix of a spheroid (task-path
 * formulas, TMatrixResource.hpp:579-776 and :893-1120) at the given order and
 * number of Gauss-Legendre nodes per hemisphere.  R_V and wavelength in the
 * same length unit.  T receives (2 nm)^2 complex values.
 */
void ctmo_tmatrix_m(double rv, double axis_ratio, double wavelength,
                    double mr_re, double mr_im, unsigned nmax, unsigned ngauss,
                    unsigned m, double *T);
/* All m blocks at fixed (nmax, ngauss); T has ctmo_tmatrix_size(nmax)
 * complex entries. */
void ctmo_tmatrix_all(double rv, double axis_ratio, double wavelength,
                      double mr_re, double mr_im, unsigned nmax,
                      unsigned ngauss, double *T);
/* m = 0 convergence sums (TMatrixResource.hpp:743-756). */
void ctmo_m0_qsums(unsigned nmax, const double *T0, double *qsca,
                   double *qext);
/* a14: TMatrix.hpp:2055-2110 over a full T-matrix. */
void ctmo_qcoeffs(unsigned nmax, const double *T, double *qsca, double *qext);

/*
 * a18: task-path convergence (TaskManager.hpp:494-499,1146-1197,
 * TMatrixResource.hpp:743-775): orders min_order..max_order-1 with
 * ngauss = glfac * order.  Returns 0 when converged, 1 when the last order
 * was reached (the reference then silently keeps the last order).
 */
int ctmo_converge_task(double rv, double axis_ratio, double wavelength,
                       double mr_re, double mr_im, unsigned min_order,
                       unsigned max_order, unsigned glfac, double tolerance,
                       unsigned *nmax_out, unsigned *ngauss_out,
                       double *dqs_out, double *dqe_out);
/*
 * a17: TMatrixCalculator::calculate_TMatrix two-phase convergence
 * (TMatrixCalculator.hpp:64-213).  Returns 0 on success, 1 when a limit is hit
 * (the reference aborts).
 */
int ctmo_converge_calculator(double ratio_of_radii, double axis_ratio,
                             double axi, double wavelength,
                             unsigned maximum_order, double tolerance,
                             unsigned ndgs, double mr_re, double mr_im,
                             unsigned maximum_ngauss, unsigned *nmax_out,
                             unsigned *ngauss_out, double *rv_out);

/*
 * Batched task-path convergence plus full T-matrix, OpenMP over grains
 * (the CPU baseline).  Outputs per grain: converged order, status, full-T
 * Q_sca/Q_ext sums.  If T_out is not NULL, T-matrices are written
 * contiguously at T_offsets[g] (complex entries).
 */
void ctmo_batch_task(int n, const double *rv, const double *axis_ratio,
                     const double *wavelength, const double *mr_re,
                     const double *mr_im, unsigned min_order,
                     unsigned max_order, unsigned glfac, double tolerance,
                     int nthreads, unsigned *nmax_out, int *status_out,
                     double *qsca_out, double *qext_out);

/* CPU baseline of the batched workload: T-matrices plus Q_ext and Q_abs
 * (scattered-light quadrature ngauss x ngauss) at ntheta incidence angles,
 * averaged with weights wtheta. */
void ctmo_batch_task_obs(int n, const double *rv, const double *axis_ratio,
                         const double *wavelength, const double *mr_re,
                         const double *mr_im, unsigned min_order,
                         unsigned max_order, unsigned glfac, double tolerance,
                         int nthreads, unsigned ntheta, const double *theta,
                         const double *wtheta, unsigned ngauss,
                         unsigned *nmax_out, int *status_out,
                         double *qext_avg, double *qabs_avg);

/* ctmo_batch_task_obs with optional full-T sums and per-angle values
 * ([n][ntheta]); with nmax_forced, no convergence loop (orders forced). */
void ctmo_batch_task_obs_ext(int n, const double *rv, const double *axis_ratio,
                             const double *wavelength, const double *mr_re,
                             const double *mr_im, unsigned min_order,
                             unsigned max_order, unsigned glfac,
                             double tolerance, int nthreads, unsigned ntheta,
                             const double *theta, const double *wtheta,
                             unsigned ngauss, const unsigned *nmax_forced,
                             unsigned *nmax_out, int *status_out,
                             double *qext_avg, double *qabs_avg,
                             double *qsca_sum, double *qext_sum,
                             double *qext_theta, double *qabs_theta);

/* a23: ExtinctionCoefficientTask.hpp:626-653 (per incidence angle). */
void ctmo_extinction(unsigned nmax, const double *T, double wavelength,
                     double rv, unsigned ntheta, const double *theta,
                     double *qext, double *qextpol, double *qextcpol);
/* a24: AbsorptionCoefficientTask.hpp:722-781. */
void ctmo_absorption(unsigned nmax, const double *T, double wavelength,
                     double rv, unsigned ntheta, const double *theta,
                     unsigned ngauss, int account_for_scattering,
                     double *qabs, double *qabspol);
/* a25: ScatteringMatrixTask.hpp:698-751 on its (theta, phi) grid. */
void ctmo_scattering_grid(unsigned nmax, const double *T, double wavelength,
                          unsigned ntheta, int use_gl, double *Z);
/* a26: S, Z, K at particle orientation (alpha, beta) for laboratory
 * directions (TMatrix.hpp:1246-1482,1891-2010); out gets 40 doubles. */
void ctmo_scattering_query(unsigned nmax, const double *T, double wavelength,
                           double alpha, double beta, double th_in,
                           double ph_in, double th_out, double ph_out,
                           double *out);
/* a19: p_N of an orientation distribution (kind 0 isotropic, 1 Mishchenko
 * with <cos^2 beta> = param, 2 Davis-Greenstein with axis ratio param). */
void ctmo_orientation_coefficients(int kind, double param, unsigned nmax,
                                   double *pN);
/* a21: AlignmentAverageTask.hpp:115-215 / TMatrixCalculator.hpp:229-307. */
void ctmo_orientation_average(unsigned nmax, const double *Tin,
                              const double *pN, double *Tout);

#ifdef __cplusplus
}
#endif

#endif /* CTM_ORACLE_H */
