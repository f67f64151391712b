/*
 * ctm_oracle.c -- CPU FP64 restatement of CosTuuM's T-matrix hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the CPU
 * baseline of bench.py.  The product library never links or calls this code.
 *
 * Each function follows the reference algorithm step for step (same
 * recurrences, same pivoting rule, same convergence flags) so that its
 * results agree with the reference's to rounding; citations are
 * file:line into SKIRT/CosTuuM's src/.
 */
#include "ctm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* Starting order of the downward ratio recurrence (SpecialFunctions.hpp:40). */
#define CTMO_BESSEL_NSTART 800u

/* ------------------------------------------------------------------------ */
/* a1: Gauss-Legendre nodes and weights on [-1, 1], Newton iteration on the
 * Legendre three-term recurrence with the Lether-Wenston initial guess and
 * a 1e-16 tolerance that relaxes x10 per iteration past 100
 * (SpecialFunctions.hpp:647-715).  Nodes ascend; x[i] = -x[order-1-i]. */
void ctmo_gauss_legendre(unsigned order, double *x, double *w) {
  const unsigned odd = order % 2u;
  const unsigned half = order / 2u + odd;
  const double inv = 1.0 / order;
  const double inv2 = inv * inv;
  for (unsigned i = 0; i < half; ++i) {
    const double t = M_PI * (i + 0.75) / (order + 0.5);
    double xi = (1. - 0.125 * inv2 + (5. / 38.) * inv * inv2 -
                 (2. / 25.) * inv2 * inv2 * (1. - (14. / 39.) / (t * t))) *
                cos(t);
    double tol = 1.e-16;
    unsigned iter = 0;
    double pinv, step;
    do {
      ++iter;
      if (iter > 100) {
        tol *= 10.;
      }
      double p_prev = 1.;
      double p_cur = xi;
      for (unsigned j = 1; j < order; ++j) {
        const double p_old = p_prev;
        p_prev = p_cur;
        p_cur = xi * p_prev + (xi * p_prev - p_old) * j / (j + 1.);
      }
      /* 1 / [(1 - x^2) P_n'(x)] */
      pinv = 1. / ((p_prev - xi * p_cur) * order);
      step = pinv * p_cur * (1. - xi * xi);
      xi = xi - step;
    } while (fabs(step) > tol * fabs(xi));
    const unsigned k = order - i - 1;
    x[k] = xi;
    w[k] = 2. * pinv * pinv * (1. - xi * xi);
    if (i != half - 1 || odd != 1) {
      x[i] = -x[k];
      w[i] = w[k];
    }
  }
}

/* a1: rescaled rule on [a, b] (SpecialFunctions.hpp:738-751). */
void ctmo_gauss_legendre_ab(unsigned order, double a, double b, double *x,
                            double *w) {
  ctmo_gauss_legendre(order, x, w);
  const double half_width = 0.5 * (b - a);
  const double mid = 0.5 * (a + b);
  for (unsigned i = 0; i < order; ++i) {
    x[i] = half_width * x[i] + mid;
    w[i] *= half_width;
  }
}

/* a4: R_V / R_S for a spheroid with axis ratio d = a/b
 * (SpecialFunctions.hpp:558-583). */
double ctmo_rv_over_rs(double d) {
  if (d > 1.) {
    const double d2 = d * d;
    const double e = sqrt(1. - 1. / d2);
    const double c = cbrt(d2);
    return 1. / sqrt(0.25 * (2. * c + log((1. + e) / (1. - e)) / (e * c * c)));
  } else if (d < 1.) {
    const double c = cbrt(d);
    const double e = sqrt(1. - d * d);
    return 1. / sqrt(0.5 * (c * c + asin(e) / (e * c)));
  }
  return 1.;
}

/* a3: r^2(theta) / R_V^2 and (1/r) dr/dtheta for a spheroid, mirror-filled
 * from the first half of the nodes (SpecialFunctions.hpp:873-893). */
void ctmo_r_dr_spheroid(const double *costheta, unsigned size, double d,
                        double *r2, double *dr_over_r) {
  const double a = cbrt(d);
  const double a2 = a * a;
  const double d2 = d * d;
  const double d2m1 = d2 - 1.;
  for (unsigned i = 0; i < size / 2; ++i) {
    const double c2 = costheta[i] * costheta[i];
    const double s2 = (1. - costheta[i]) * (1. + costheta[i]);
    const double s = sqrt(s2);
    const double r2a2 = 1. / (s2 + d2 * c2);
    r2[i] = a2 * r2a2;
    r2[size - i - 1] = r2[i];
    dr_over_r[i] = r2a2 * costheta[i] * s * d2m1;
    dr_over_r[size - i - 1] = -dr_over_r[i];
  }
}

/* a6: j_n(z) and [z j_n(z)]'/z for n = 1..nmax, real z.  Downward ratio
 * recurrence from n = 800, then forward products
 * (SpecialFunctions.hpp:202-240). */
void ctmo_sph_j_real(unsigned nmax, double z, double *j, double *dj) {
  double rho[CTMO_BESSEL_NSTART];
  const double zinv = 1. / z;
  rho[CTMO_BESSEL_NSTART - 1] = z / (2. * CTMO_BESSEL_NSTART + 1.);
  for (unsigned idx = CTMO_BESSEL_NSTART - 1; idx >= 1; --idx) {
    rho[idx - 1] = 1. / ((2. * idx + 1.) * zinv - rho[idx]);
  }
  const double j0 = sin(z) * zinv;
  j[0] = rho[0] * j0;
  dj[0] = j0 - j[0] * zinv;
  for (unsigned i = 1; i < nmax; ++i) {
    j[i] = rho[i] * j[i - 1];
    dj[i] = j[i - 1] - (i + 1.) * zinv * j[i];
  }
}

static void sph_j_cplx(unsigned nmax, cplx z, cplx *j, cplx *dj) {
  cplx rho[CTMO_BESSEL_NSTART];
  const cplx zinv = 1. / z;
  rho[CTMO_BESSEL_NSTART - 1] = z / (2. * CTMO_BESSEL_NSTART + 1.);
  for (unsigned idx = CTMO_BESSEL_NSTART - 1; idx >= 1; --idx) {
    rho[idx - 1] = 1. / ((2. * idx + 1.) * zinv - rho[idx]);
  }
  const cplx j0 = csin(z) * zinv;
  j[0] = rho[0] * j0;
  dj[0] = j0 - j[0] * zinv;
  for (unsigned i = 1; i < nmax; ++i) {
    j[i] = rho[i] * j[i - 1];
    dj[i] = j[i - 1] - (i + 1.) * zinv * j[i];
  }
}

void ctmo_sph_j_cplx(unsigned nmax, const double *z, double *j, double *dj) {
  sph_j_cplx(nmax, z[0] + I * z[1], (cplx *)j, (cplx *)dj);
}

/* a7: y_n(x) and [x y_n(x)]'/x for n = 1..nmax by forward recurrence from
 * the closed forms of y_1, y_2 (SpecialFunctions.hpp:107-140). */
void ctmo_sph_y_real(unsigned nmax, double z, double *y, double *dy) {
  const double c = cos(z);
  const double s = sin(z);
  const double zi = 1. / z;
  const double zi2 = zi * zi;
  const double zi3 = zi2 * zi;
  double y0 = -c * zi2 - s * zi;
  double y1 = (-3. * zi3 + zi) * c - 3. * zi2 * s;
  y[0] = y0;
  dy[0] = -zi * (c + y0);
  if (nmax > 1) {
    y[1] = y1;
    dy[1] = y0 - 2. * zi * y1;
  }
  for (unsigned i = 2; i < nmax; ++i) {
    y[i] = (2. * i + 1.) * zi * y[i - 1] - y[i - 2];
    dy[i] = y[i - 1] - (i + 1.) * zi * y[i];
  }
}

/* a9: Wigner d^n_{0m}(x) and d/dx for n = 1..nmax by the three-term
 * recurrence started at A_m sin^m x; zeros for n < m
 * (SpecialFunctions.hpp:303-371). */
void ctmo_wigner_dn_0m(double cosx, unsigned nmax, unsigned m, double *y,
                       double *dy) {
  const double sinx = sqrt((1. - cosx) * (1. + cosx));
  const double sinxinv = 1. / sinx;
  if (m == 0) {
    double dprev = 1.;
    double dcur = cosx;
    for (unsigned n = 0; n < nmax; ++n) {
      const double np1 = n + 1.;
      const double np2 = n + 2.;
      const double tnp3 = 2. * np1 + 1.;
      const double dnext = (tnp3 * cosx * dcur - np1 * dprev) / np2;
      const double deriv = sinxinv * np1 * np2 * (dnext - dprev) / tnp3;
      y[n] = dcur;
      dy[n] = deriv;
      dprev = dcur;
      dcur = dnext;
    }
  } else {
    const double m2 = (double)m * (double)m;
    double dprev = 0.;
    double dcur = 1.;
    for (unsigned i = 0; i < m; ++i) {
      const double i2 = 2. * (i + 1.);
      dcur *= sqrt((i2 - 1.) / i2) * sinx;
    }
    for (unsigned n = 0; n + 1 < m; ++n) {
      y[n] = 0.;
      dy[n] = 0.;
    }
    for (unsigned n = m - 1; n < nmax; ++n) {
      const double np1 = n + 1.;
      const double np2 = n + 2.;
      const double tnp3 = 2. * np1 + 1.;
      const double s1 = sqrt(np1 * np1 - m2);
      const double s2 = sqrt(np2 * np2 - m2);
      const double dnext = (tnp3 * cosx * dcur - s1 * dprev) / s2;
      const double deriv = sinxinv * (np1 * s2 * dnext - np2 * s1 * dprev) / tnp3;
      y[n] = dcur;
      dy[n] = deriv;
      dprev = dcur;
      dcur = dnext;
    }
  }
}

/* a22: d^n_{0m}(x)/sin(x) and d/dx with the |cos x| ~ 1 special case
 * (SpecialFunctions.hpp:398-495). */
void ctmo_wigner_dn_0m_sinx(double cosx, double sinx, double sinx_inv,
                            unsigned nmax, unsigned m, double *y, double *dy) {
  (void)sinx;
  if (fabs(1. - fabs(cosx)) <= 1.e-10) {
    if (m == 1) {
      int sign = 1;
      for (unsigned n = 1; n < nmax + 1; ++n) {
        double dn = 0.5 * sqrt(n * (n + 1.));
        if (cosx < 0.) {
          dn *= sign;
        }
        y[n - 1] = dn;
        if (cosx < 0.) {
          dn = -dn;
        }
        dy[n - 1] = dn;
        sign = -sign;
      }
    } else {
      for (unsigned i = 0; i < nmax; ++i) {
        y[i] = 0.;
        dy[i] = 0.;
      }
    }
    return;
  }
  if (m == 0) {
    double dprev = 1.;
    double dcur = cosx;
    for (unsigned n = 0; n < nmax; ++n) {
      const double np1 = n + 1.;
      const double np2 = n + 2.;
      const double tnp3 = 2. * np1 + 1.;
      const double dnext = (tnp3 * cosx * dcur - np1 * dprev) / np2;
      const double deriv = sinx_inv * np1 * np2 * (dnext - dprev) / tnp3;
      y[n] = dcur * sinx_inv;
      dy[n] = deriv;
      dprev = dcur;
      dcur = dnext;
    }
  } else {
    const double m2 = (double)m * (double)m;
    double dprev = 0.;
    double dcur = 1.;
    for (unsigned i = 0; i < m; ++i) {
      const double i2 = 2. * (i + 1.);
      dcur *= sqrt((i2 - 1.) / i2) * sinx;
    }
    for (unsigned n = 0; n + 1 < m; ++n) {
      y[n] = 0.;
      dy[n] = 0.;
    }
    for (unsigned n = m - 1; n < nmax; ++n) {
      const double np1 = n + 1.;
      const double np2 = n + 2.;
      const double tnp3 = 2. * np1 + 1.;
      const double s1 = sqrt(np1 * np1 - m2);
      const double s2 = sqrt(np2 * np2 - m2);
      const double dnext = (tnp3 * cosx * dcur - s1 * dprev) / s2;
      const double deriv =
          sinx_inv * (np1 * s2 * dnext - np2 * s1 * dprev) / tnp3;
      y[n] = dcur * sinx_inv;
      dy[n] = deriv;
      dprev = dcur;
      dcur = dnext;
    }
  }
}

/* a13: explicit inverse by Doolittle PLU with partial pivoting on |A_ji|
 * (first maximum wins), inversion of U, solve inv(A) L = inv(U), column
 * exchanges in reverse pivot order (Matrix.hpp:261-384). */
static void plu_inverse(unsigned n, cplx *A, unsigned *piv, cplx *work) {
#define AE(r, c) A[(size_t)(r) * n + (c)]
  for (unsigned i = 0; i < n; ++i) {
    unsigned imax = i;
    double smax = cabs(AE(i, i));
    for (unsigned j = i + 1; j < n; ++j) {
      const double s = cabs(AE(j, i));
      if (s > smax) {
        smax = s;
        imax = j;
      }
    }
    piv[i] = imax;
    if (imax != i) {
      for (unsigned j = 0; j < n; ++j) {
        const cplx t = AE(i, j);
        AE(i, j) = AE(imax, j);
        AE(imax, j) = t;
      }
    }
    const cplx inv = 1. / AE(i, i);
    for (unsigned j = i + 1; j < n; ++j) {
      AE(j, i) *= inv;
    }
    for (unsigned j = i + 1; j < n; ++j) {
      for (unsigned k = i + 1; k < n; ++k) {
        AE(j, k) -= AE(j, i) * AE(i, k);
      }
    }
  }
  for (unsigned i = 0; i < n; ++i) {
    AE(i, i) = 1. / AE(i, i);
    for (unsigned j = i + 1; j < n; ++j) {
      cplx acc = 0.;
      for (unsigned k = i; k < j; ++k) {
        acc -= AE(i, k) * AE(k, j);
      }
      AE(i, j) = acc / AE(j, j);
    }
  }
  for (unsigned jp1 = n; jp1 > 0; --jp1) {
    const unsigned j = jp1 - 1;
    for (unsigned i = 0; i < jp1; ++i) {
      work[i] = AE(i, j);
    }
    for (unsigned i = jp1; i < n; ++i) {
      work[i] = 0.;
    }
    for (unsigned k = jp1; k < n; ++k) {
      for (unsigned i = 0; i < n; ++i) {
        work[i] -= AE(i, k) * AE(k, j);
      }
    }
    for (unsigned i = 0; i < n; ++i) {
      AE(i, j) = work[i];
    }
  }
  for (unsigned jp1 = n; jp1 > 0; --jp1) {
    const unsigned j = jp1 - 1;
    const unsigned jp = piv[j];
    if (jp != j) {
      for (unsigned i = 0; i < n; ++i) {
        const cplx t = AE(i, j);
        AE(i, j) = AE(i, jp);
        AE(i, jp) = t;
      }
    }
  }
#undef AE
}

void ctmo_plu_inverse(unsigned n, double *A) {
  unsigned *piv = (unsigned *)malloc(sizeof(unsigned) * (n ? n : 1));
  cplx *work = (cplx *)malloc(sizeof(cplx) * (n ? n : 1));
  plu_inverse(n, (cplx *)A, piv, work);
  free(piv);
  free(work);
}

/* ------------------------------------------------------------------------ */
/* T-matrix core. */

size_t ctmo_tmatrix_offset(unsigned nmax, unsigned m) {
  size_t off = 0;
  for (unsigned mm = 0; mm < m; ++mm) {
    const size_t nm = (mm == 0) ? nmax : nmax + 1 - mm;
    off += 4 * nm * nm;
  }
  return off;
}

size_t ctmo_tmatrix_size(unsigned nmax) {
  return ctmo_tmatrix_offset(nmax, nmax + 1);
}

/* Per-(grain, order, ngauss) tables: the quadrature (GaussBasedResources.hpp
 * :132-152), geometry (ParticleGeometryResource.hpp:135-149), interaction
 * variables and Bessel tables (InteractionResource.hpp:75-78,478-527) for the
 * first ngauss nodes of the 2 ngauss-point rule (the hemisphere the assembly
 * sums over). */
typedef struct {
  unsigned nmax, ngauss;
  double k2;
  cplx k2mr;
  double *costh, *weight, *sinv, *sin2inv, *r2, *dr;
  double *krinv;
  cplx *krmrinv;
  double *jkr, *djkr, *ykr, *dykr; /* [ngauss][nmax] */
  cplx *jkrmr, *djkrmr;            /* [ngauss][nmax] */
} tables_t;

static void tables_build(tables_t *t, double rv, double d, double lambda,
                         cplx mr, unsigned nmax, unsigned ngauss) {
  const unsigned nn = 2 * ngauss;
  t->nmax = nmax;
  t->ngauss = ngauss;
  const double k = 2. * M_PI / lambda;
  t->k2 = k * k;
  const cplx kmr = mr * k;
  t->k2mr = mr * t->k2;
  double *x = (double *)malloc(sizeof(double) * nn);
  double *w = (double *)malloc(sizeof(double) * nn);
  double *r2 = (double *)malloc(sizeof(double) * nn);
  double *dr = (double *)malloc(sizeof(double) * nn);
  ctmo_gauss_legendre(nn, x, w);
  ctmo_r_dr_spheroid(x, nn, d, r2, dr);
  t->costh = (double *)malloc(sizeof(double) * ngauss);
  t->weight = (double *)malloc(sizeof(double) * ngauss);
  t->sinv = (double *)malloc(sizeof(double) * ngauss);
  t->sin2inv = (double *)malloc(sizeof(double) * ngauss);
  t->r2 = (double *)malloc(sizeof(double) * ngauss);
  t->dr = (double *)malloc(sizeof(double) * ngauss);
  t->krinv = (double *)malloc(sizeof(double) * ngauss);
  t->krmrinv = (cplx *)malloc(sizeof(cplx) * ngauss);
  const size_t tsz = (size_t)ngauss * nmax;
  t->jkr = (double *)malloc(sizeof(double) * tsz);
  t->djkr = (double *)malloc(sizeof(double) * tsz);
  t->ykr = (double *)malloc(sizeof(double) * tsz);
  t->dykr = (double *)malloc(sizeof(double) * tsz);
  t->jkrmr = (cplx *)malloc(sizeof(cplx) * tsz);
  t->djkrmr = (cplx *)malloc(sizeof(cplx) * tsz);
  for (unsigned ig = 0; ig < ngauss; ++ig) {
    t->costh[ig] = x[ig];
    t->weight[ig] = w[ig];
    const double s2i = 1. / ((1. - x[ig]) * (1. + x[ig]));
    t->sin2inv[ig] = s2i;
    t->sinv[ig] = sqrt(s2i);
    t->r2[ig] = r2[ig];
    t->dr[ig] = dr[ig];
    const double r = sqrt(r2[ig]);
    const double kr = k * r * rv;
    const cplx krmr = kmr * r * rv;
    t->krinv[ig] = 1. / kr;
    t->krmrinv[ig] = 1. / krmr;
    ctmo_sph_j_real(nmax, kr, t->jkr + (size_t)ig * nmax,
                    t->djkr + (size_t)ig * nmax);
    ctmo_sph_y_real(nmax, kr, t->ykr + (size_t)ig * nmax,
                    t->dykr + (size_t)ig * nmax);
    sph_j_cplx(nmax, krmr, t->jkrmr + (size_t)ig * nmax,
               t->djkrmr + (size_t)ig * nmax);
  }
  free(x);
  free(w);
  free(r2);
  free(dr);
}

static void tables_free(tables_t *t) {
  free(t->costh);
  free(t->weight);
  free(t->sinv);
  free(t->sin2inv);
  free(t->r2);
  free(t->dr);
  free(t->krinv);
  free(t->krmrinv);
  free(t->jkr);
  free(t->djkr);
  free(t->ykr);
  free(t->dykr);
  free(t->jkrmr);
  free(t->djkrmr);
}

/* a5: a_{n1 n2} = sqrt((2n1+1)(2n2+1) / (n1(n1+1) n2(n2+1)))
 * (NBasedResources.hpp:115-137). */
static double ann(unsigned n1, unsigned n2) {
  const double s1 = sqrt((2. * n1 + 1.) / (n1 * (n1 + 1.)));
  const double s2 = sqrt((2. * n2 + 1.) / (n2 * (n2 + 1.)));
  return s1 * s2;
}

/* a11/a12: Q, RgQ assembly for one m (TMatrixResource.hpp:601-709 for m = 0,
 * :927-1100 for m > 0), explicit inverse and T_m = -RgQ Q^-1
 * (:712-730, :1102-1119). */
static void tmatrix_m_from_tables(const tables_t *t, unsigned m, cplx *T) {
  const unsigned nmax = t->nmax;
  const unsigned nm = (m == 0) ? nmax : nmax + 1 - m;
  const unsigned nmin = (m == 0) ? 1 : m;
  const unsigned n2m = 2 * nm;
  cplx *Q = (cplx *)calloc((size_t)n2m * n2m, sizeof(cplx));
  cplx *RgQ = (cplx *)calloc((size_t)n2m * n2m, sizeof(cplx));
  double *wd = (double *)malloc(sizeof(double) * nmax);
  double *dwd = (double *)malloc(sizeof(double) * nmax);
  const double k2 = t->k2;
  const cplx k2mr = t->k2mr;
  const double mm = (double)m;
  const double m2 = mm * mm;
#define QE(M, r, c) M[(size_t)(r) * n2m + (c)]
  for (unsigned ig = 0; ig < t->ngauss; ++ig) {
    ctmo_wigner_dn_0m(t->costh[ig], nmax, m, wd, dwd);
    const double *jr = t->jkr + (size_t)ig * nmax;
    const double *djr = t->djkr + (size_t)ig * nmax;
    const double *yr = t->ykr + (size_t)ig * nmax;
    const double *dyr = t->dykr + (size_t)ig * nmax;
    const cplx *jm = t->jkrmr + (size_t)ig * nmax;
    const cplx *djm = t->djkrmr + (size_t)ig * nmax;
    const double krinv = t->krinv[ig];
    const cplx krmrinv = t->krmrinv[ig];
    const double dr = t->dr[ig];
    const double wr2 = t->weight[ig] * t->r2[ig];
    for (unsigned n1 = nmin; n1 <= nmax; ++n1) {
      const double n1n1p1 = n1 * (n1 + 1.);
      const unsigned i1 = n1 - nmin;
      const unsigned ii1 = i1 + nm;
      for (unsigned n2 = nmin; n2 <= nmax; ++n2) {
        const int even = ((n1 + n2) % 2u) == 0;
        if (m == 0 && !even) {
          continue;
        }
        const double n2n2p1 = n2 * (n2 + 1.);
        const unsigned i2 = n2 - nmin;
        const unsigned ii2 = i2 + nm;
        const double w1 = wd[n1 - 1], dw1 = dwd[n1 - 1];
        const double w2 = wd[n2 - 1], dw2 = dwd[n2 - 1];
        const double jk = jr[n1 - 1];
        const cplx h = jk + I * yr[n1 - 1];
        const double djk = djr[n1 - 1];
        const cplx dh = djk + I * dyr[n1 - 1];
        const cplx jmn2 = jm[n2 - 1];
        const cplx djmn2 = djm[n2 - 1];
        const cplx c1 = jmn2 * jk;
        const cplx b1 = jmn2 * h;
        const cplx c2 = jmn2 * djk;
        const cplx b2 = jmn2 * dh;
        const cplx c4 = djmn2 * jk;
        const cplx b4 = djmn2 * h;
        const double a12 = ann(n1, n2);
        if (!even) {
          /* off-diagonal blocks J11/J22 (TMatrixResource.hpp:993-1044) */
          const double dsi = t->sinv[ig] * mm * wr2;
          const cplx c6 = djmn2 * djk;
          const cplx b6 = djmn2 * dh;
          const cplx c7 = c4 * krinv;
          const cplx b7 = b4 * krinv;
          const cplx c8 = c2 * krmrinv;
          const cplx b8 = b2 * krmrinv;
          const double e1 = dsi * (w1 * dw2 + dw1 * w2);
          const double f = dsi * dr * (w1 * w2);
          const double e2 = f * n1n1p1;
          const double e3 = f * n2n2p1;
          const cplx J11 = I * a12 * (e1 * b1);
          const cplx RgJ11 = I * a12 * (e1 * c1);
          const cplx J22 = I * a12 * (e1 * b6 + e2 * b7 + e3 * b8);
          const cplx RgJ22 = I * a12 * (e1 * c6 + e2 * c7 + e3 * c8);
          QE(Q, i1, ii2) += k2mr * J11 + k2 * J22;
          QE(RgQ, i1, ii2) += k2mr * RgJ11 + k2 * RgJ22;
          QE(Q, ii1, i2) += k2mr * J22 + k2 * J11;
          QE(RgQ, ii1, i2) += k2mr * RgJ22 + k2 * RgJ11;
        } else {
          /* diagonal blocks J12/J21 (TMatrixResource.hpp:623-705,
           * :1045-1097) */
          const cplx c3 = krinv * c1;
          const cplx b3 = krinv * b1;
          const cplx c5 = c1 * krmrinv;
          const cplx b5 = b1 * krmrinv;
          const double dssi = t->sin2inv[ig] * m2;
          const double f1 = wr2 * ((w1 * w2) * dssi + dw1 * dw2);
          const double f2 = wr2 * dr * n1n1p1 * (w1 * dw2);
          const double f3 = wr2 * dr * n2n2p1 * (dw1 * w2);
          const cplx J12 = -a12 * (f1 * b2 + f2 * b3);
          const cplx RgJ12 = -a12 * (f1 * c2 + f2 * c3);
          const cplx J21 = a12 * (f1 * b4 + f3 * b5);
          const cplx RgJ21 = a12 * (f1 * c4 + f3 * c5);
          QE(Q, i1, i2) += k2mr * J21 + k2 * J12;
          QE(RgQ, i1, i2) += k2mr * RgJ21 + k2 * RgJ12;
          QE(Q, ii1, ii2) += k2mr * J12 + k2 * J21;
          QE(RgQ, ii1, ii2) += k2mr * RgJ12 + k2 * RgJ21;
        }
      }
    }
  }
  unsigned *piv = (unsigned *)malloc(sizeof(unsigned) * n2m);
  cplx *work = (cplx *)malloc(sizeof(cplx) * n2m);
  plu_inverse(n2m, Q, piv, work);
  for (unsigned i = 0; i < n2m; ++i) {
    for (unsigned j = 0; j < n2m; ++j) {
      cplx acc = 0.;
      for (unsigned k = 0; k < n2m; ++k) {
        acc -= QE(RgQ, i, k) * QE(Q, k, j);
      }
      T[(size_t)i * n2m + j] = acc;
    }
  }
#undef QE
  free(piv);
  free(work);
  free(Q);
  free(RgQ);
  free(wd);
  free(dwd);
}

void ctmo_tmatrix_m(double rv, double d, double lambda, double mr_re,
                    double mr_im, unsigned nmax, unsigned ngauss, unsigned m,
                    double *T) {
  tables_t t;
  tables_build(&t, rv, d, lambda, mr_re + I * mr_im, nmax, ngauss);
  tmatrix_m_from_tables(&t, m, (cplx *)T);
  tables_free(&t);
}

void ctmo_tmatrix_all(double rv, double d, double lambda, double mr_re,
                      double mr_im, unsigned nmax, unsigned ngauss,
                      double *T) {
  tables_t t;
  tables_build(&t, rv, d, lambda, mr_re + I * mr_im, nmax, ngauss);
  for (unsigned m = 0; m <= nmax; ++m) {
    tmatrix_m_from_tables(&t, m, (cplx *)T + ctmo_tmatrix_offset(nmax, m));
  }
  tables_free(&t);
}

/* m = 0 convergence sums (TMatrixResource.hpp:743-756). */
void ctmo_m0_qsums(unsigned nmax, const double *T0d, double *qsca,
                   double *qext) {
  const cplx *T0 = (const cplx *)T0d;
  const unsigned n2 = 2 * nmax;
  double s = 0., e = 0.;
  for (unsigned n = 1; n <= nmax; ++n) {
    const double dn1 = 2. * n + 1.;
    const cplx a = T0[(size_t)(n - 1) * n2 + (n - 1)];
    const cplx b = T0[(size_t)(nmax + n - 1) * n2 + (nmax + n - 1)];
    s += dn1 * ((creal(a) * creal(a) + cimag(a) * cimag(a)) +
                (creal(b) * creal(b) + cimag(b) * cimag(b)));
    e += dn1 * (creal(a) + creal(b));
  }
  *qsca = s;
  *qext = e;
}

/* a14: Q_sca = sum_m f_m sum |T^{ij}_{n1 n2}|^2, Q_ext = sum_m f_m sum_n
 * Re T^{ij}_{nn} with f_0 = 1, f_{m>0} = 2 (TMatrix.hpp:2055-2110). */
void ctmo_qcoeffs(unsigned nmax, const double *Td, double *qsca,
                  double *qext) {
  const cplx *T = (const cplx *)Td;
  double s = 0., e = 0.;
  for (unsigned m = 0; m <= nmax; ++m) {
    const unsigned nm = (m == 0) ? nmax : nmax + 1 - m;
    const unsigned n2m = 2 * nm;
    const cplx *B = T + ctmo_tmatrix_offset(nmax, m);
    const double f = (m == 0) ? 1. : 2.;
    double bs = 0., be = 0.;
    for (unsigned r = 0; r < n2m; ++r) {
      for (unsigned c = 0; c < n2m; ++c) {
        const cplx v = B[(size_t)r * n2m + c];
        bs += creal(v) * creal(v) + cimag(v) * cimag(v);
      }
    }
    for (unsigned i = 0; i < nm; ++i) {
      be += creal(B[(size_t)i * n2m + i]) +
            creal(B[(size_t)i * n2m + nm + i]) +
            creal(B[(size_t)(nm + i) * n2m + i]) +
            creal(B[(size_t)(nm + i) * n2m + nm + i]);
    }
    s += f * bs;
    e += f * be;
  }
  *qsca = s;
  *qext = e;
}

static void m0_sums_at(double rv, double d, double lambda, cplx mr,
                       unsigned nmax, unsigned ngauss, double *qs,
                       double *qe) {
  tables_t t;
  tables_build(&t, rv, d, lambda, mr, nmax, ngauss);
  cplx *T0 = (cplx *)malloc(sizeof(cplx) * 4 * (size_t)nmax * nmax);
  tmatrix_m_from_tables(&t, 0, T0);
  ctmo_m0_qsums(nmax, (const double *)T0, qs, qe);
  free(T0);
  tables_free(&t);
}

/* a18: task-path convergence.  The first order sets the flags to -1 and can
 * never converge; later orders compare |(old - new) / new| of both sums with
 * the tolerance (TMatrixResource.hpp:743-775). */
int ctmo_converge_task(double rv, double d, double lambda, double mr_re,
                       double mr_im, unsigned min_order, unsigned max_order,
                       unsigned glfac, double tol, unsigned *nmax_out,
                       unsigned *ngauss_out, double *dqs_out,
                       double *dqe_out) {
  const cplx mr = mr_re + I * mr_im;
  double old_s = 0., old_e = 0.;
  double dqs = -2., dqe = -2.;
  unsigned last = min_order;
  for (unsigned n = min_order; n < max_order; ++n) {
    const unsigned ng = glfac * n;
    double qs, qe;
    m0_sums_at(rv, d, lambda, mr, n, ng, &qs, &qe);
    last = n;
    if (dqs == -2.) {
      dqs = -1.;
      dqe = -1.;
    } else {
      dqs = fabs((old_s - qs) / qs);
      dqe = fabs((old_e - qe) / qe);
    }
    old_s = qs;
    old_e = qe;
    if (dqs > 0. && dqs <= tol && dqe <= tol) {
      *nmax_out = n;
      *ngauss_out = ng;
      if (dqs_out) *dqs_out = dqs;
      if (dqe_out) *dqe_out = dqe;
      return 0;
    }
  }
  *nmax_out = last;
  *ngauss_out = glfac * last;
  if (dqs_out) *dqs_out = dqs;
  if (dqe_out) *dqe_out = dqe;
  return 1;
}

/* a17: two-phase convergence of TMatrixCalculator::calculate_TMatrix
 * (TMatrixCalculator.hpp:64-213): nmax from max(4, x + 4.05 x^(1/3)) at
 * ngauss = ndgs nmax, then ngauss from ndgs nmax + 1; old sums carry over. */
int ctmo_converge_calculator(double ratio_of_radii, double d, double axi,
                             double lambda, unsigned maximum_order,
                             double tol, unsigned ndgs, double mr_re,
                             double mr_im, unsigned maximum_ngauss,
                             unsigned *nmax_out, unsigned *ngauss_out,
                             double *rv_out) {
  const cplx mr = mr_re + I * mr_im;
  if (fabs(ratio_of_radii - 1.) > 1.e-8) {
    ratio_of_radii = ctmo_rv_over_rs(d);
  }
  const double rv = ratio_of_radii * axi;
  const double xev = 2. * M_PI * rv / lambda;
  double guess = xev + 4.05 * cbrt(xev);
  if (guess < 4.) {
    guess = 4.;
  }
  unsigned nmax = (unsigned)guess;
  double old_e = 0., old_s = 0., dext = 1., dsca = 1.;
  *rv_out = rv;
  while (nmax < maximum_order && (dext > tol || dsca > tol)) {
    double qs, qe;
    m0_sums_at(rv, d, lambda, mr, nmax, ndgs * nmax, &qs, &qe);
    dsca = fabs((old_s - qs) / qs);
    dext = fabs((old_e - qe) / qe);
    old_e = qe;
    old_s = qs;
    ++nmax;
  }
  if (nmax == maximum_order) {
    *nmax_out = nmax;
    *ngauss_out = 0;
    return 1;
  }
  --nmax;
  unsigned ngauss = nmax * ndgs + 1;
  dext = tol + 1.;
  dsca = tol + 1.;
  while (ngauss < maximum_ngauss && (dext > tol || dsca > tol)) {
    double qs, qe;
    m0_sums_at(rv, d, lambda, mr, nmax, ngauss, &qs, &qe);
    dsca = fabs((old_s - qs) / qs);
    dext = fabs((old_e - qe) / qe);
    old_e = qe;
    old_s = qs;
    ++ngauss;
  }
  if (ngauss == maximum_ngauss) {
    *nmax_out = nmax;
    *ngauss_out = ngauss;
    return 1;
  }
  --ngauss;
  *nmax_out = nmax;
  *ngauss_out = ngauss;
  return 0;
}

void ctmo_batch_task(int n, const double *rv, const double *d,
                     const double *lambda, const double *mr_re,
                     const double *mr_im, unsigned min_order,
                     unsigned max_order, unsigned glfac, double tol,
                     int nthreads, unsigned *nmax_out, int *status_out,
                     double *qsca_out, double *qext_out) {
  /* Same task decomposition as the reference's graph
   * (TaskManager.hpp:1146-1197,1314-1327): one order chain per grain, then
   * independent (grain, m) blocks, each level spread over the threads. */
#ifdef _OPENMP
  if (nthreads > 0) {
    omp_set_num_threads(nthreads);
  }
#endif
  unsigned *ngauss = (unsigned *)malloc(sizeof(unsigned) * (n ? n : 1));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int g = 0; g < n; ++g) {
    double dqs, dqe;
    status_out[g] = ctmo_converge_task(rv[g], d[g], lambda[g], mr_re[g],
                                       mr_im[g], min_order, max_order, glfac,
                                       tol, &nmax_out[g], &ngauss[g], &dqs,
                                       &dqe);
  }
  tables_t *tabs = (tables_t *)malloc(sizeof(tables_t) * (n ? n : 1));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int g = 0; g < n; ++g) {
    tables_build(&tabs[g], rv[g], d[g], lambda[g], mr_re[g] + I * mr_im[g],
                 nmax_out[g], ngauss[g]);
  }
  /* (grain, m) work list, largest blocks first */
  size_t *first = (size_t *)malloc(sizeof(size_t) * (n + 1));
  first[0] = 0;
  for (int g = 0; g < n; ++g) first[g + 1] = first[g] + nmax_out[g] + 1;
  const size_t nwork = first[n];
  int *wg = (int *)malloc(sizeof(int) * (nwork ? nwork : 1));
  unsigned *wm = (unsigned *)malloc(sizeof(unsigned) * (nwork ? nwork : 1));
  double *bs = (double *)calloc(nwork ? nwork : 1, sizeof(double));
  double *be = (double *)calloc(nwork ? nwork : 1, sizeof(double));
  for (int g = 0; g < n; ++g) {
    for (unsigned m = 0; m <= nmax_out[g]; ++m) {
      wg[first[g] + m] = g;
      wm[first[g] + m] = m;
    }
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (long q = 0; q < (long)nwork; ++q) {
    const int g = wg[q];
    const unsigned m = wm[q];
    const unsigned nmax = nmax_out[g];
    const unsigned nm = (m == 0) ? nmax : nmax + 1 - m;
    const unsigned n2m = 2 * nm;
    cplx *T = (cplx *)malloc(sizeof(cplx) * (size_t)n2m * n2m);
    tmatrix_m_from_tables(&tabs[g], m, T);
    double s = 0., e = 0.;
    for (size_t k = 0; k < (size_t)n2m * n2m; ++k) {
      s += creal(T[k]) * creal(T[k]) + cimag(T[k]) * cimag(T[k]);
    }
    for (unsigned i = 0; i < nm; ++i) {
      e += creal(T[(size_t)i * n2m + i]) + creal(T[(size_t)i * n2m + nm + i]) +
           creal(T[(size_t)(nm + i) * n2m + i]) +
           creal(T[(size_t)(nm + i) * n2m + nm + i]);
    }
    const double f = (m == 0) ? 1. : 2.;
    bs[q] = f * s;
    be[q] = f * e;
    free(T);
  }
  for (int g = 0; g < n; ++g) {
    double s = 0., e = 0.;
    for (size_t q = first[g]; q < first[g + 1]; ++q) {
      s += bs[q];
      e += be[q];
    }
    qsca_out[g] = s;
    qext_out[g] = e;
    tables_free(&tabs[g]);
  }
  free(tabs);
  free(first);
  free(wg);
  free(wm);
  free(bs);
  free(be);
  free(ngauss);
}

/* ------------------------------------------------------------------------ */
/* a20: Clebsch-Gordan coefficients <n1 m1 n2 m2 | N M> by the ladder
 * recursion: top row M = N from C(m1 = n1) = 1 downward, normalised, then
 * lower M by the lowering recursion; negative M by the symmetry
 * (-1)^(N-n1-n2) (SpecialFunctions.hpp:973-1108).  The table C has
 * (N + 1) rows of 2 n1 + 1 entries indexed [M][m1 + n1]. */
static void cg_table(int n1, int n2, int N, double *C) {
  const int ns = 2 * n1 + 1;
  memset(C, 0, sizeof(double) * (size_t)(N + 1) * ns);
  C[N * ns + n1 + n1] = 1.;
  double norm = 1.;
  const int lo = (-n1 > N - n2 - 1) ? -n1 : N - n2 - 1;
  for (int m1 = n1 - 1; m1 >= lo; --m1) {
    const double cp2 = sqrt((n2 - N + m1 + 1.) * (n2 + N - m1));
    const double cp1 = sqrt((n1 - m1) * (n1 + m1 + 1.));
    const double next = -cp2 * C[N * ns + m1 + 1 + n1] / cp1;
    C[N * ns + m1 + n1] = next;
    norm += next * next;
  }
  const double inorm = 1. / sqrt(norm);
  for (int m1 = n1; m1 >= lo; --m1) {
    C[N * ns + m1 + n1] *= inorm;
  }
  for (int M = N - 1; M >= 0; --M) {
    const double cmi = 1. / sqrt((N + M + 1.) * (N - M));
    const int a = (-n1 > M - n2) ? -n1 : M - n2;
    const int b = (n1 < M + n2) ? n1 : M + n2;
    for (int m1 = a; m1 <= b; ++m1) {
      const int m2 = M - m1;
      double t1 = 0., t2 = 0.;
      if (m1 + 1 <= n1) {
        t1 = sqrt((n1 + m1 + 1.) * (n1 - m1)) * C[(M + 1) * ns + m1 + 1 + n1];
      }
      if (m2 + 1 <= n2) {
        t2 = sqrt((n2 + m2 + 1.) * (n2 - m2)) * C[(M + 1) * ns + m1 + n1];
      }
      C[M * ns + m1 + n1] = cmi * (t1 + t2);
    }
  }
}

double ctmo_clebsch_gordan(int n1, int m1, int n2, int m2, int N, int M) {
  if (m1 + m2 != M) {
    return 0.;
  }
  double *C = (double *)malloc(sizeof(double) * (size_t)(N + 1) * (2 * n1 + 1));
  cg_table(n1, n2, N, C);
  double v;
  if (M >= 0) {
    v = C[M * (2 * n1 + 1) + m1 + n1];
  } else {
    const int sgn = ((N - n1 - n2) % 2 == 0) ? 1 : -1;
    v = sgn * C[-M * (2 * n1 + 1) + n1 - m1];
  }
  free(C);
  return v;
}

/* Packed M = 0 family <n1 m1 n2 -m1 | N 0> for m1 = -nmin..nmin,
 * nmin = min(n1, n2) (SpecialFunctions.hpp:1131-1244); out has 2 nmin + 1
 * entries; scratch has (N + 1)(2 n1 + 1) entries. */
static void cg_m0_family(int n1, int n2, int N, double *out, double *scratch) {
  cg_table(n1, n2, N, scratch);
  const int nmin = (n1 < n2) ? n1 : n2;
  for (int i = 0; i < 2 * nmin + 1; ++i) {
    out[i] = scratch[i - nmin + n1];
  }
}

/* ------------------------------------------------------------------------ */
/* Observables and averaging. */

/* T(i1, n1, m, i2, n2, m) of a full T-matrix in the reference block layout
 * (TMatrix.hpp:1028-1046). */
static cplx tm_el(const cplx *T, unsigned nmax, int i1, unsigned n1,
                  unsigned m, int i2, unsigned n2) {
  const unsigned nm = (m == 0) ? nmax : nmax + 1 - m;
  const unsigned nmin = (m == 0) ? 1 : m;
  const cplx *B = T + ctmo_tmatrix_offset(nmax, m);
  return B[(size_t)(i1 * nm + n1 - nmin) * (2 * nm) + i2 * nm + n2 - nmin];
}

/* c_{n1 n2} = i^(n1 - n2 - 1) a_{n1 n2} (NBasedResources.hpp:115-137). */
static cplx cnn(unsigned n1, unsigned n2) {
  static const cplx ipow[4] = {1., I, -1., -I};
  const int e = (((int)n1 - (int)n2 - 1) % 4 + 4) % 4;
  return ipow[e] * ann(n1, n2);
}

/* pi/tau tables at one angle for all m <= nmax (the special Wigner
 * resources, ExtinctionCoefficientTask.hpp:242-415): pi[m][n-1] =
 * m d^n_{0m}/sin, tau[m][n-1] = d/dtheta d^n_{0m}. */
static void pitau(double costh, double sinth, double sininv, unsigned nmax,
                  double *pi, double *tau) {
  for (unsigned m = 0; m <= nmax; ++m) {
    double *p = pi + (size_t)m * nmax;
    double *t = tau + (size_t)m * nmax;
    ctmo_wigner_dn_0m_sinx(costh, sinth, sininv, nmax, m, p, t);
    for (unsigned n = 0; n < nmax; ++n) p[n] *= m;
  }
}

static void angle_trig(double theta, double *c, double *s, double *si) {
  *c = cos(theta);
  *s = sqrt((1. - *c) * (1. + *c));
  *si = (*s != 0.) ? 1. / *s : 9000.;
}

/* a23-a25: forward-scattering-type S matrix from T for an incoming
 * direction with pi/tau tables (pin, tin) and an outgoing one (pout, tout),
 * e^{i m (phi_out - phi_in)} given by (cphi, sphi); divided by k
 * (ExtinctionCoefficientTask.hpp:517-619). */
static void smatrix(const cplx *T, unsigned nmax, double k, const double *pin,
                    const double *tin, const double *pout, const double *tout,
                    double cphi, double sphi, cplx *S) {
  S[0] = S[1] = S[2] = S[3] = 0.;
  const cplx step = cphi + I * sphi;
  cplx e = 1.;
  for (unsigned m = 0; m <= nmax; ++m) {
    const unsigned nmin = (m == 0) ? 1 : m;
    const double fcos = 2. * creal(e);
    const double fsin = 2. * cimag(e);
    e *= step;
    const double *pi_i = pin + (size_t)m * nmax, *ta_i = tin + (size_t)m * nmax;
    const double *pi_o = pout + (size_t)m * nmax,
                 *ta_o = tout + (size_t)m * nmax;
    for (unsigned nn = nmin; nn <= nmax; ++nn) {
      const double pnn = pi_i[nn - 1], tnn = ta_i[nn - 1];
      for (unsigned n = nmin; n <= nmax; ++n) {
        const double pn = pi_o[n - 1], tn = ta_o[n - 1];
        const cplx c = cnn(nn, n);
        const cplx T11 = tm_el(T, nmax, 0, n, m, 0, nn);
        const cplx T22 = tm_el(T, nmax, 1, n, m, 1, nn);
        if (m == 0) {
          const cplx f = c * tn * tnn;
          S[0] += f * T22;
          S[3] += f * T11;
        } else {
          const cplx T12 = tm_el(T, nmax, 0, n, m, 1, nn);
          const cplx T21 = tm_el(T, nmax, 1, n, m, 0, nn);
          const cplx rf = c * fcos, jf = c * fsin;
          const double pp = pn * pnn, pt = pn * tnn, tp = tn * pnn,
                       tt = tn * tnn;
          S[0] += rf * (T11 * pp + T21 * tp + T12 * pt + T22 * tt);
          S[1] += jf * (T11 * pt + T21 * tt + T12 * pp + T22 * tp);
          S[2] -= jf * (T11 * tp + T21 * pp + T12 * tt + T22 * pt);
          S[3] += rf * (T11 * tt + T21 * pt + T12 * tp + T22 * pp);
        }
      }
    }
  }
  const double kinv = 1. / k;
  for (int q = 0; q < 4; ++q) S[q] *= kinv;
}

/* a23: Q_ext, Q_extpol, Q_extcpol at the given incidence angles
 * (ExtinctionCoefficientTask.hpp:626-653), normalised by pi R_V^2. */
void ctmo_extinction(unsigned nmax, const double *Td, double wavelength,
                     double rv, unsigned ntheta, const double *theta,
                     double *qext, double *qextpol, double *qextcpol) {
  const cplx *T = (const cplx *)Td;
  const double k = 2. * M_PI / wavelength;
  const size_t tsz = (size_t)(nmax + 1) * nmax;
  double *pi = (double *)malloc(sizeof(double) * tsz);
  double *ta = (double *)malloc(sizeof(double) * tsz);
  const double norm = M_PI * rv * rv;
  for (unsigned i = 0; i < ntheta; ++i) {
    double c, s, si;
    angle_trig(theta[i], &c, &s, &si);
    pitau(c, s, si, nmax, pi, ta);
    cplx S[4];
    smatrix(T, nmax, k, pi, ta, pi, ta, 1., 0., S);
    const double pre = 2. * M_PI / k;
    qext[i] = pre * cimag(S[0] + S[3]) / norm;
    qextpol[i] = pre * cimag(S[0] - S[3]) / norm;
    qextcpol[i] = pre * creal(S[3] - S[0]) / norm;
  }
  free(pi);
  free(ta);
}

/* a24: Q_abs, Q_abspol at the given outgoing angles; with
 * account_for_scattering the GL(ngauss) x GL(ngauss) quadrature over
 * incoming directions of Z00 / Z10 is subtracted
 * (AbsorptionCoefficientTask.hpp:722-781, grid :125-295). */
void ctmo_absorption(unsigned nmax, const double *Td, double wavelength,
                     double rv, unsigned ntheta, const double *theta,
                     unsigned ngauss, int account_for_scattering,
                     double *qabs, double *qabspol) {
  const cplx *T = (const cplx *)Td;
  const double k = 2. * M_PI / wavelength;
  const size_t tsz = (size_t)(nmax + 1) * nmax;
  double *pio = (double *)malloc(sizeof(double) * tsz);
  double *tao = (double *)malloc(sizeof(double) * tsz);
  double *pii = (double *)malloc(sizeof(double) * tsz * (ngauss ? ngauss : 1));
  double *tai = (double *)malloc(sizeof(double) * tsz * (ngauss ? ngauss : 1));
  double *ct = (double *)malloc(sizeof(double) * (ngauss ? ngauss : 1));
  double *wt = (double *)malloc(sizeof(double) * (ngauss ? ngauss : 1));
  double *ph = (double *)malloc(sizeof(double) * (ngauss ? ngauss : 1));
  double *wp = (double *)malloc(sizeof(double) * (ngauss ? ngauss : 1));
  if (account_for_scattering && ngauss > 0) {
    ctmo_gauss_legendre(ngauss, ct, wt);
    ctmo_gauss_legendre_ab(ngauss, 0., 2. * M_PI, ph, wp);
    for (unsigned g = 0; g < ngauss; ++g) {
      const double s = sqrt((1. - ct[g]) * (1. + ct[g]));
      pitau(ct[g], s, (s != 0.) ? 1. / s : 9000., nmax, pii + g * tsz,
            tai + g * tsz);
    }
  }
  const double norm = M_PI * rv * rv;
  for (unsigned i = 0; i < ntheta; ++i) {
    double c, s, si;
    angle_trig(theta[i], &c, &s, &si);
    pitau(c, s, si, nmax, pio, tao);
    cplx S[4];
    smatrix(T, nmax, k, pio, tao, pio, tao, 1., 0., S);
    const double pre = 2. * M_PI / k;
    double a = pre * cimag(S[0] + S[3]);
    double ap = pre * cimag(S[0] - S[3]);
    if (account_for_scattering) {
      for (unsigned ti = 0; ti < ngauss; ++ti) {
        for (unsigned pj = 0; pj < ngauss; ++pj) {
          cplx St[4];
          smatrix(T, nmax, k, pii + ti * tsz, tai + ti * tsz, pio, tao,
                  cos(ph[pj]), -sin(ph[pj]), St);
          const double w = wt[ti] * wp[pj];
          double n2[4];
          for (int q = 0; q < 4; ++q)
            n2[q] = creal(St[q]) * creal(St[q]) + cimag(St[q]) * cimag(St[q]);
          a -= 0.5 * (n2[0] + n2[1] + n2[2] + n2[3]) * w;
          ap -= 0.5 * (n2[0] + n2[1] - n2[2] - n2[3]) * w;
        }
      }
    }
    qabs[i] = a / norm;
    qabspol[i] = ap / norm;
  }
  free(pio);
  free(tao);
  free(pii);
  free(tai);
  free(ct);
  free(wt);
  free(ph);
  free(wp);
}

/* Z (4x4, row-major) from S (ScatteringMatrixTask.hpp:722-748). */
static void z_from_s(const cplx *S, double *Z) {
  const cplx S0 = S[0], S1 = S[1], S2 = S[2], S3 = S[3];
  Z[0] = creal(0.5 * (S0 * conj(S0) + S1 * conj(S1) + S2 * conj(S2) + S3 * conj(S3)));
  Z[1] = creal(0.5 * (S0 * conj(S0) - S1 * conj(S1) + S2 * conj(S2) - S3 * conj(S3)));
  Z[2] = creal(-S0 * conj(S1) - S3 * conj(S2));
  Z[3] = creal(I * (S0 * conj(S1) - S3 * conj(S2)));
  Z[4] = creal(0.5 * (S0 * conj(S0) + S1 * conj(S1) - S2 * conj(S2) - S3 * conj(S3)));
  Z[5] = creal(0.5 * (S0 * conj(S0) - S1 * conj(S1) - S2 * conj(S2) + S3 * conj(S3)));
  Z[6] = creal(-S0 * conj(S1) + S3 * conj(S2));
  Z[7] = creal(I * (S0 * conj(S1) + S3 * conj(S2)));
  Z[8] = creal(-S0 * conj(S2) - S3 * conj(S1));
  Z[9] = creal(-S0 * conj(S2) + S3 * conj(S1));
  Z[10] = creal(S0 * conj(S3) + S1 * conj(S2));
  Z[11] = creal(-I * (S0 * conj(S3) + S2 * conj(S1)));
  Z[12] = creal(I * (S2 * conj(S0) + S3 * conj(S1)));
  Z[13] = creal(I * (S2 * conj(S0) - S3 * conj(S1)));
  Z[14] = creal(-I * (S3 * conj(S0) - S1 * conj(S2)));
  Z[15] = creal(S3 * conj(S0) - S1 * conj(S2));
}

/* a25: get_table's scattering matrix: incidence theta_in = pi/2, phi_in = 0;
 * outgoing theta on GL(ntheta) in cos (or uniform i pi/ntheta), phi on
 * GL(2 ntheta) in [0, 2pi] (or uniform); Z[(itheta * nphi + iphi) * 16]
 * (ScatteringMatrixTask.hpp:107-277,698-751, TaskManager.hpp:832-833). */
void ctmo_scattering_grid(unsigned nmax, const double *Td, double wavelength,
                          unsigned ntheta, int use_gl, double *Z) {
  const cplx *T = (const cplx *)Td;
  const double k = 2. * M_PI / wavelength;
  const unsigned nphi = 2 * ntheta;
  const size_t tsz = (size_t)(nmax + 1) * nmax;
  double *pin = (double *)malloc(sizeof(double) * tsz);
  double *tin = (double *)malloc(sizeof(double) * tsz);
  double *po = (double *)malloc(sizeof(double) * tsz);
  double *to = (double *)malloc(sizeof(double) * tsz);
  double *cth = (double *)malloc(sizeof(double) * ntheta);
  double *phi = (double *)malloc(sizeof(double) * nphi);
  double *ign = (double *)malloc(sizeof(double) * nphi);
  double c, s, si;
  angle_trig(M_PI_2, &c, &s, &si);
  pitau(c, s, si, nmax, pin, tin);
  if (use_gl) {
    ctmo_gauss_legendre(ntheta, cth, ign);
    ctmo_gauss_legendre_ab(nphi, 0., 2. * M_PI, phi, ign);
  } else {
    for (unsigned i = 0; i < ntheta; ++i) cth[i] = cos(i * (M_PI / ntheta));
    for (unsigned i = 0; i < nphi; ++i) phi[i] = i * (2. * M_PI / nphi);
  }
  for (unsigned it = 0; it < ntheta; ++it) {
    const double co = cth[it];
    const double so = sqrt((1. - co) * (1. + co));
    pitau(co, so, (so != 0.) ? 1. / so : 9000., nmax, po, to);
    for (unsigned ip = 0; ip < nphi; ++ip) {
      cplx S[4];
      smatrix(T, nmax, k, pin, tin, po, to, cos(phi[ip]), sin(phi[ip]), S);
      z_from_s(S, Z + ((size_t)it * nphi + ip) * 16);
    }
  }
  free(pin);
  free(tin);
  free(po);
  free(to);
  free(cth);
  free(phi);
  free(ign);
}

/* a19: Legendre coefficients p_N of an orientation distribution.
 * kind 0: isotropic p(beta) = 1/2, integrated numerically like
 * OrientationDistribution::initialise (OrientationDistribution.hpp:220-243):
 * normalisation, then p_N = int_0^pi sin b p(b) d^N_00(b) db by adaptive GL
 * (SpecialFunctions.hpp:772-810) with tolerances 1e-10 / 1e-5.
 * kind 1: Mishchenko p0 = 1, p2 = 1.5 <cos^2> - 0.5
 * (MishchenkoOrientationDistribution.hpp:105-119).
 * kind 2: Davis-Greenstein for axis ratio `param`
 * (DavisGreensteinOrientationDistribution.hpp:76-102). */
typedef struct {
  unsigned order;
  double norm;
} pn_args;

static double iso_integrand(double beta, const pn_args *a) {
  const double cb = cos(beta), sb = sin(beta);
  double wd = 1.;
  if (a->order > 0) {
    double *y = (double *)malloc(sizeof(double) * a->order);
    double *dy = (double *)malloc(sizeof(double) * a->order);
    ctmo_wigner_dn_0m(cb, a->order, 0, y, dy);
    wd = y[a->order - 1];
    free(y);
    free(dy);
  }
  return sb * (a->norm * 0.5) * wd;
}

static double gl_quad_adaptive(const pn_args *a, int normalisation,
                               unsigned start, unsigned maxo, double abs_err,
                               double rel_err) {
  unsigned order = start;
  double old = 0., ad = abs_err + 1., rd = rel_err + 1.;
  while (order < maxo && ad > abs_err && rd > rel_err) {
    double *x = (double *)malloc(sizeof(double) * order);
    double *w = (double *)malloc(sizeof(double) * order);
    ctmo_gauss_legendre_ab(order, 0., M_PI, x, w);
    double q = 0.;
    for (unsigned i = 0; i < order; ++i) {
      q += w[i] * (normalisation ? sin(x[i]) * 0.5 : iso_integrand(x[i], a));
    }
    free(x);
    free(w);
    ad = fabs(q - old);
    rd = ad / fabs(q + old);
    old = q;
    ++order;
  }
  return old;
}

void ctmo_orientation_coefficients(int kind, double param, unsigned nmax,
                                   double *pN) {
  for (unsigned i = 0; i <= nmax; ++i) pN[i] = 0.;
  if (kind == 0) {
    pn_args a = {0, 1.};
    const double norm = gl_quad_adaptive(&a, 1, 100, 1000, 1e-10, 1e-5);
    a.norm = 1. / norm;
    pN[0] = 1.;
    for (unsigned i = 1; i <= nmax; ++i) {
      a.order = i;
      pN[i] = gl_quad_adaptive(&a, 0, 2 * (i + 1), 100 * (i + 1), 1e-10, 1e-5);
    }
  } else if (kind == 1) {
    pN[0] = 1.;
    if (nmax >= 2) pN[2] = 1.5 * param - 0.5;
  } else if (kind == 2) {
    if (param < 1.) {
      double num = 1., den = 2.;
      pN[0] = 1.;
      for (unsigned i = 1; 2 * i <= nmax; ++i) {
        const int sgn = (i % 2 == 0) ? 1 : -1;
        pN[2 * i] = sgn * num / den;
        num *= 2. * i + 1.;
        den *= 2. * (i + 1.);
      }
    } else {
      for (unsigned i = 0; i <= nmax; ++i) pN[i] = 1.;
    }
  }
}

/* a21: ensemble T-matrix <T>_{m n1 n2} = sum_N p_N sum_m1 (+-) f C C T_m1
 * (Mishchenko 1991 eq. 3.27; AlignmentAverageTask.hpp:115-215,
 * TMatrixCalculator.hpp:229-307).  pN has at least 2 nmax + 1 entries. */
void ctmo_orientation_average(unsigned nmax, const double *Tind,
                              const double *pN, double *Toutd) {
  const cplx *Tin = (const cplx *)Tind;
  cplx *Tout = (cplx *)Toutd;
  memset(Tout, 0, sizeof(cplx) * ctmo_tmatrix_size(nmax));
  double *C = (double *)malloc(sizeof(double) * (2 * nmax + 1));
  double *scratch =
      (double *)malloc(sizeof(double) * (2 * nmax + 1) * (2 * nmax + 1));
  for (unsigned m = 0; m <= nmax; ++m) {
    const int msign = (m % 2 == 0) ? 1 : -1;
    const unsigned nmin = (m == 0) ? 1 : m;
    const unsigned nm = (m == 0) ? nmax : nmax + 1 - m;
    cplx *B = Tout + ctmo_tmatrix_offset(nmax, m);
    for (unsigned n1 = nmin; n1 <= nmax; ++n1) {
      for (unsigned n2 = nmin; n2 <= nmax; ++n2) {
        cplx t[2][2] = {{0., 0.}, {0., 0.}};
        const unsigned lo = (n1 > n2) ? n1 - n2 : n2 - n1;
        const unsigned hi = n1 + n2;
        const unsigned M = (n1 < n2) ? n1 : n2;
        for (unsigned N = lo; N <= hi; ++N) {
          const double p = pN[N];
          /* a zero coefficient adds exact zeros: skip its CG family */
          if (p == 0.) {
            continue;
          }
          cg_m0_family((int)n1, (int)n2, (int)N, C, scratch);
          for (unsigned m1 = 0; m1 <= M; ++m1) {
            const double f = (m1 == 0) ? 1. : 2.;
            const int sg = ((m1 % 2 == 0) ? 1 : -1) * msign;
            const double cg = sg * f * p * C[M + m] * C[M + m1];
            if ((N + hi) % 2 == 0) {
              t[0][0] += cg * tm_el(Tin, nmax, 0, n1, m1, 0, n2);
              t[1][1] += cg * tm_el(Tin, nmax, 1, n1, m1, 1, n2);
            } else {
              t[0][1] += cg * tm_el(Tin, nmax, 0, n1, m1, 1, n2);
              t[1][0] += cg * tm_el(Tin, nmax, 1, n1, m1, 0, n2);
            }
          }
        }
        for (int i1 = 0; i1 < 2; ++i1)
          for (int i2 = 0; i2 < 2; ++i2)
            B[(size_t)(i1 * nm + n1 - nmin) * (2 * nm) + i2 * nm + n2 - nmin] =
                t[i1][i2];
      }
    }
  }
  free(C);
  free(scratch);
}

/* The CPU baseline of the batched workload: task-path T-matrices as
 * ctmo_batch_task, then Q_ext and Q_abs (with the scattered-light
 * quadrature of ngauss x ngauss incoming directions) at ntheta incidence
 * angles, averaged with weights wtheta; parallel over (grain, m) blocks and
 * over (grain, angle) pairs.  Optional outputs (NULL to skip): the full-T
 * sums Q_sca, Q_ext (TMatrix.hpp:2055-2110) and the per-angle Q_ext, Q_abs
 * ([n][ntheta]).  With nmax_forced (not NULL) no convergence loop runs:
 * every grain is computed at nmax_forced[g], ngauss = glfac nmax_forced[g]
 * (status 0), as for the sensitivity of a converged grain to its inputs. */
void ctmo_batch_task_obs_ext(int n, const double *rv, const double *d,
                             const double *lambda, const double *mr_re,
                             const double *mr_im, unsigned min_order,
                             unsigned max_order, unsigned glfac, double tol,
                             int nthreads, unsigned ntheta,
                             const double *theta, const double *wtheta,
                             unsigned ngauss, const unsigned *nmax_forced,
                             unsigned *nmax_out, int *status_out,
                             double *qext_avg, double *qabs_avg,
                             double *qsca_sum, double *qext_sum,
                             double *qext_theta, double *qabs_theta) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  unsigned *ng = (unsigned *)malloc(sizeof(unsigned) * (n ? n : 1));
  if (nmax_forced) {
    for (int g = 0; g < n; ++g) {
      nmax_out[g] = nmax_forced[g];
      ng[g] = glfac * nmax_forced[g];
      status_out[g] = 0;
    }
  } else {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int g = 0; g < n; ++g) {
      double dqs, dqe;
      status_out[g] = ctmo_converge_task(rv[g], d[g], lambda[g], mr_re[g],
                                         mr_im[g], min_order, max_order,
                                         glfac, tol, &nmax_out[g], &ng[g],
                                         &dqs, &dqe);
    }
  }
  tables_t *tabs = (tables_t *)malloc(sizeof(tables_t) * (n ? n : 1));
  cplx **Ts = (cplx **)malloc(sizeof(cplx *) * (n ? n : 1));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int g = 0; g < n; ++g) {
    tables_build(&tabs[g], rv[g], d[g], lambda[g], mr_re[g] + I * mr_im[g],
                 nmax_out[g], ng[g]);
    Ts[g] = (cplx *)malloc(sizeof(cplx) * ctmo_tmatrix_size(nmax_out[g]));
  }
  size_t *first = (size_t *)malloc(sizeof(size_t) * (n + 1));
  first[0] = 0;
  for (int g = 0; g < n; ++g) first[g + 1] = first[g] + nmax_out[g] + 1;
  const long nwork = (long)first[n];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (long q = 0; q < nwork; ++q) {
    int g = 0;
    while ((size_t)q >= first[g + 1]) ++g;
    const unsigned m = (unsigned)(q - (long)first[g]);
    tmatrix_m_from_tables(&tabs[g], m,
                          Ts[g] + ctmo_tmatrix_offset(nmax_out[g], m));
  }
  double *qe = (double *)calloc((size_t)n * ntheta + 1, sizeof(double));
  double *qa = (double *)calloc((size_t)n * ntheta + 1, sizeof(double));
  const long npairs = (long)n * ntheta;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (long q = 0; q < npairs; ++q) {
    const int g = (int)(q / ntheta);
    const unsigned j = (unsigned)(q % ntheta);
    double e, ep, ec, a, ap;
    ctmo_extinction(nmax_out[g], (const double *)Ts[g], lambda[g], rv[g], 1,
                    theta + j, &e, &ep, &ec);
    ctmo_absorption(nmax_out[g], (const double *)Ts[g], lambda[g], rv[g], 1,
                    theta + j, ngauss, 1, &a, &ap);
    qe[q] = e;
    qa[q] = a;
  }
  for (int g = 0; g < n; ++g) {
    double se = 0., sa = 0.;
    for (unsigned j = 0; j < ntheta; ++j) {
      se += wtheta[j] * qe[(size_t)g * ntheta + j];
      sa += wtheta[j] * qa[(size_t)g * ntheta + j];
    }
    qext_avg[g] = se;
    qabs_avg[g] = sa;
    if (qsca_sum || qext_sum) {
      double s, e;
      ctmo_qcoeffs(nmax_out[g], (const double *)Ts[g], &s, &e);
      if (qsca_sum) qsca_sum[g] = s;
      if (qext_sum) qext_sum[g] = e;
    }
    if (qext_theta)
      memcpy(qext_theta + (size_t)g * ntheta, qe + (size_t)g * ntheta,
             sizeof(double) * ntheta);
    if (qabs_theta)
      memcpy(qabs_theta + (size_t)g * ntheta, qa + (size_t)g * ntheta,
             sizeof(double) * ntheta);
    tables_free(&tabs[g]);
    free(Ts[g]);
  }
  free(qe);
  free(qa);
  free(first);
  free(tabs);
  free(Ts);
  free(ng);
}

void ctmo_batch_task_obs(int n, const double *rv, const double *d,
                         const double *lambda, const double *mr_re,
                         const double *mr_im, unsigned min_order,
                         unsigned max_order, unsigned glfac, double tol,
                         int nthreads, unsigned ntheta, const double *theta,
                         const double *wtheta, unsigned ngauss,
                         unsigned *nmax_out, int *status_out,
                         double *qext_avg, double *qabs_avg) {
  ctmo_batch_task_obs_ext(n, rv, d, lambda, mr_re, mr_im, min_order,
                          max_order, glfac, tol, nthreads, ntheta, theta,
                          wtheta, ngauss, NULL, nmax_out, status_out,
                          qext_avg, qabs_avg, NULL, NULL, NULL, NULL);
}

/* a26: S, Z and K for particle orientation (alpha, beta) and laboratory
 * directions (TMatrix.hpp:1246-1482 rotations and particle-frame S,
 * :1891-1952 Z, :1972-2010 K).  out: S (4 re/im pairs), Z[16], K[16]. */
void ctmo_scattering_query(unsigned nmax, const double *Td, double wavelength,
                           double alpha, double beta, double th_in,
                           double ph_in, double th_out, double ph_out,
                           double *out) {
  const cplx *T = (const cplx *)Td;
  const double k = 2. * M_PI / wavelength;
  const double ca = cos(alpha), sa = sin(alpha), cb = cos(beta), sb = sin(beta);
  double ctl[2] = {cos(th_in), cos(th_out)}, stl[2] = {sin(th_in), sin(th_out)};
  double cpl[2] = {cos(ph_in), cos(ph_out)}, spl[2] = {sin(ph_in), sin(ph_out)};
  double ctp[2], stp[2], stpi[2], cpp[2], spp[2];
  double R[2][2][2];
  const double B[3][3] = {{ca * cb, sa * cb, -sb}, {-sa, ca, 0.},
                          {ca * sb, sa * sb, cb}};
  for (int d = 0; d < 2; ++d) {
    const double cpr = cpl[d] * ca + spl[d] * sa;
    const double spr = spl[d] * ca - cpl[d] * sa;
    ctp[d] = ctl[d] * cb + stl[d] * sb * cpr;
    stp[d] = sin(acos(ctp[d]));
    if (stp[d] != 0.) {
      stpi[d] = 1. / stp[d];
      cpp[d] = (cb * stl[d] * cpr - sb * ctl[d]) * stpi[d];
      spp[d] = stl[d] * spr * stpi[d];
    } else {
      cpp[d] = 1.;
      spp[d] = 0.;
      stpi[d] = 9000.;
    }
    const double AL[3][2] = {{ctl[d] * cpl[d], -spl[d]},
                             {ctl[d] * spl[d], cpl[d]},
                             {-stl[d], 0.}};
    const double AP[2][3] = {{ctp[d] * cpp[d], ctp[d] * spp[d], -stp[d]},
                             {-spp[d], cpp[d], 0.}};
    double C[3][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 2; ++j) {
        C[i][j] = 0.;
        for (int q = 0; q < 3; ++q) C[i][j] += B[i][q] * AL[q][j];
      }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        R[d][i][j] = 0.;
        for (int q = 0; q < 3; ++q) R[d][i][j] += AP[i][q] * C[q][j];
      }
  }
  {
    const double det = 1. / (R[1][0][0] * R[1][1][1] - R[1][0][1] * R[1][1][0]);
    const double tmp = R[1][0][0];
    R[1][0][0] = R[1][1][1] * det;
    R[1][0][1] = -R[1][0][1] * det;
    R[1][1][0] = -R[1][1][0] * det;
    R[1][1][1] = tmp * det;
  }
  /* particle-frame S (TMatrix.hpp:1109-1229) */
  const size_t tsz = (size_t)(nmax + 1) * nmax;
  double *pin = (double *)malloc(sizeof(double) * tsz);
  double *tin = (double *)malloc(sizeof(double) * tsz);
  double *pout = (double *)malloc(sizeof(double) * tsz);
  double *tout = (double *)malloc(sizeof(double) * tsz);
  pitau(ctp[0], stp[0], stpi[0], nmax, pin, tin);
  pitau(ctp[1], stp[1], stpi[1], nmax, pout, tout);
  cplx S[4];
  smatrix(T, nmax, k, pin, tin, pout, tout, cpp[1] * cpp[0] + spp[1] * spp[0],
          spp[1] * cpp[0] - cpp[1] * spp[0], S);
  free(pin);
  free(tin);
  free(pout);
  free(tout);
  const cplx c11 = S[0] * R[0][0][0] + S[1] * R[0][1][0];
  const cplx c12 = S[0] * R[0][0][1] + S[1] * R[0][1][1];
  const cplx c21 = S[2] * R[0][0][0] + S[3] * R[0][1][0];
  const cplx c22 = S[2] * R[0][0][1] + S[3] * R[0][1][1];
  cplx F[4];
  F[0] = R[1][0][0] * c11 + R[1][0][1] * c21;
  F[1] = R[1][0][0] * c12 + R[1][0][1] * c22;
  F[2] = R[1][1][0] * c11 + R[1][1][1] * c21;
  F[3] = R[1][1][0] * c12 + R[1][1][1] * c22;
  for (int q = 0; q < 4; ++q) {
    out[2 * q] = creal(F[q]);
    out[2 * q + 1] = cimag(F[q]);
  }
  z_from_s(F, out + 8);
  double *K = out + 24;
  const double pre = 2. * M_PI / k;
  K[0] = pre * cimag(F[0] + F[3]);
  K[5] = K[10] = K[15] = K[0];
  K[1] = pre * cimag(F[0] - F[3]);
  K[4] = K[1];
  K[2] = -pre * cimag(F[1] + F[2]);
  K[8] = K[2];
  K[3] = pre * creal(F[2] - F[1]);
  K[12] = K[3];
  K[6] = pre * cimag(F[2] - F[1]);
  K[9] = -K[6];
  K[7] = -pre * creal(F[1] + F[2]);
  K[13] = -K[7];
  K[11] = pre * creal(F[3] - F[0]);
  K[14] = -K[11];
}
