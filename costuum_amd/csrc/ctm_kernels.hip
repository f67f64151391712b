// ctm_kernels.hip -- gfx950 kernels of the batched T-matrix path (FP64).
//
// K0 k_gauss_legendre  Gauss-Legendre nodes per distinct G
//                      (SpecialFunctions.hpp:647-715)
// K2 k_wigner          d^n_{0m}, d/dtheta per distinct (N, m)
//                      (SpecialFunctions.hpp:303-371, WignerDResources.hpp)
// K1 k_tables          per (grain, N): node scalars and Bessel tables
//                      (InteractionResource.hpp:478-527,
//                      SpecialFunctions.hpp:107-240,873-893)
// K3 k_assemble        Q/RgQ class blocks, one 16x16 tile of (n1, n2)
//                      pairs per workgroup (TMatrixResource.hpp:601-709,
//                      927-1100)
// K4 k_invert_solve_*  Gauss-Jordan inverse of Q_c (registers, LDS or
//                      global scratch) and T_c = -RgQ_c Q_c^-1
//                      (Matrix.hpp:261-384, TMatrixResource.hpp:712-730,
//                      1102-1119)
// K5 k_converge        per-grain order test (TMatrixResource.hpp:743-775)
// K5 k_full_sums       per-grain Q_sca/Q_ext over all m (TMatrix.hpp:2055-2110)
#include "ctm_device.hpp"
#include "ctm_kernels.hpp"
#include "ctm_gj.hpp"

namespace ctm {

// Starting order of the downward ratio recurrence (SpecialFunctions.hpp:40).
constexpr int kBesselStart = 800;

// The reference starts the ratio recurrence at n = 800 for every argument.
// For n well above max(N, |z|) the recurrence contracts by ~(|z|/2n)^2 per
// step, so a start at max(N, |z|) + 10 |z|^(1/3) + 50 already reaches the
// same fixed point: the ratios come out bit-identical to the 800 start (checked
// over |z| in [1e-4, 500], arg z in [0, 1.2], N up to 200, at twice this
// margin's precision need).  Never above 800.
__device__ inline int bessel_start(int N, double absz) {
  const int a = (int)absz;
  const int ns = (N > a ? N : a) + (int)(10. * cbrt(absz)) + 50;
  return ns < kBesselStart ? ns : kBesselStart;
}

// ---------------------------------------------------------------------------
// K0: Gauss-Legendre nodes.  Thread i runs the Newton iteration of root i of
// the 2G-point rule (Lether-Wenston guess, tolerance 1e-16 relaxed x10 per
// iteration past 100) and stores node i = -root, the first half in ascending
// order.  The Legendre recurrence takes its ratios j / (j + 1) from a table
// in LDS (one FP64 division per step otherwise).
constexpr int kGlRatios = 2048;

__global__ __launch_bounds__(256) void k_gauss_legendre(const GlKey *keys,
                                                        double *gl) {
  __shared__ double ratio[kGlRatios];
  const GlKey key = keys[blockIdx.x];
  const int G = key.G;
  double *x = gl + key.off;
  double *w = x + G;
  const unsigned order = 2u * (unsigned)G;
  const unsigned ntab = order < (unsigned)kGlRatios ? order : kGlRatios;
  for (unsigned j = threadIdx.x; j < ntab; j += blockDim.x)
    ratio[j] = j / (j + 1.);
  __syncthreads();
  const double inv = 1.0 / order;
  const double inv2 = inv * inv;
  for (int i = threadIdx.x; i < G; i += blockDim.x) {
    const double t = M_PI * (i + 0.75) / (order + 0.5);
    double xi = (1. - 0.125 * inv2 + (5. / 38.) * inv * inv2 -
                 (2. / 25.) * inv2 * inv2 * (1. - (14. / 39.) / (t * t))) *
                cos(t);
    double tol = 1.e-16;
    int iter = 0;
    double pinv, step, last = 1.;
    do {
      ++iter;
      if (iter > 100) tol *= 10.;
      double p_prev = 1.;
      double p_cur = xi;
#pragma unroll 8
      for (unsigned j = 1; j < ntab; ++j) {
        const double p_old = p_prev;
        p_prev = p_cur;
        p_cur = xi * p_prev + (xi * p_prev - p_old) * ratio[j];
      }
      for (unsigned j = ntab; j < order; ++j) {
        const double p_old = p_prev;
        p_prev = p_cur;
        p_cur = xi * p_prev + (xi * p_prev - p_old) * (j / (j + 1.));
      }
      pinv = 1. / ((p_prev - xi * p_cur) * order);
      step = pinv * p_cur * (1. - xi * xi);
      xi = xi - step;
      // Newton converges quadratically down to the rounding level of the
      // recurrence (a few 1e-16 relative, more near the ends of large
      // rules), where the steps stop shrinking; the 1e-16 test alone would
      // keep iterating there until the tolerance has relaxed past
      // iteration 100.  Such a node is converged: stop.
      if (iter > 3 && fabs(step) >= 0.25 * last) break;
      if (iter > 4 && fabs(step) <= 4.e-16) break;
      last = fabs(step);
    } while (fabs(step) > tol * fabs(xi));
    x[i] = -xi;
    w[i] = 2. * pinv * pinv * (1. - xi * xi);
  }
}

// ---------------------------------------------------------------------------
// K2: Wigner d^n_{0m}(theta) and its derivative at the cached nodes.
__global__ void k_wigner(const WigKey *keys, const double *gl, double *wig) {
  const WigKey key = keys[blockIdx.x];
  const int N = key.N, m = key.m, G = key.G;
  const int nmin = m > 0 ? m : 1;
  const int nm = N + 1 - nmin;
  double *wd = wig + key.off;
  double *wdd = wd + (int64_t)nm * G;
  const double *cth = gl + key.gl_off;
  for (int ig = threadIdx.x; ig < G; ig += blockDim.x) {
    const double cosx = cth[ig];
    const double sinx = sqrt((1. - cosx) * (1. + cosx));
    const double sinxinv = 1. / sinx;
    if (m == 0) {
      double dprev = 1., dcur = cosx;
      for (int n = 0; n < N; ++n) {
        const double np1 = n + 1., np2 = n + 2.;
        const double tnp3 = 2. * np1 + 1.;
        const double dnext = (tnp3 * cosx * dcur - np1 * dprev) / np2;
        const double deriv = sinxinv * np1 * np2 * (dnext - dprev) / tnp3;
        wd[(int64_t)n * G + ig] = dcur;
        wdd[(int64_t)n * G + ig] = deriv;
        dprev = dcur;
        dcur = dnext;
      }
    } else {
      const double m2 = (double)m * (double)m;
      double dprev = 0., dcur = 1.;
      for (int i = 0; i < m; ++i) {
        const double i2 = 2. * (i + 1.);
        dcur *= sqrt((i2 - 1.) / i2) * sinx;
      }
      for (int n = m - 1; n < N; ++n) {
        const double np1 = n + 1., np2 = n + 2.;
        const double tnp3 = 2. * np1 + 1.;
        const double s1 = sqrt(np1 * np1 - m2);
        const double s2 = sqrt(np2 * np2 - m2);
        const double dnext = (tnp3 * cosx * dcur - s1 * dprev) / s2;
        const double deriv =
            sinxinv * (np1 * s2 * dnext - np2 * s1 * dprev) / tnp3;
        const int nl = n + 1 - nmin;
        wd[(int64_t)nl * G + ig] = dcur;
        wdd[(int64_t)nl * G + ig] = deriv;
        dprev = dcur;
        dcur = dnext;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K1: per-(grain, order) node scalars and Bessel tables.  One thread per
// node over all sets of the list (a thread finds its set by binary search
// in the sets' node offsets, which start at node_base for the first set of
// the launch).  j_n (at kr and at m kr): the ratio recurrence
// r_n = j_n / j_{n-1} runs down from the start order, and the product of the
// ratios below N gives j_{N-1} = j_0 r_1 ... r_{N-1}; from j_N and j_{N-1}
// the three-term recurrence runs down (the direction in which j_n, the
// minimal solution, is stable) and writes every j_n once.  Only when
// j_{N-1} lies below the normal range (tiny arguments at high order) are
// the ratios parked in the output rows and multiplied up instead.  Only the
// functions are stored (the readers form the derivatives from neighbouring
// orders): j_0, y_0 and j_0(m kr) go with the node scalars.
constexpr double kMillerFloor = 1e-280;
// dn (device list counts, may be NULL): the set and node counts are read
// from dn[0], dn[1] (the grid covers at least the nodes).
__global__ __launch_bounds__(128) void k_tables(const TableSet *sets,
                                                int n_sets, int node_base,
                                                int n_nodes,
                                                const GrainParams gp,
                                                const double *gl,
                                                double *tables,
                                                const int32_t *dn) {
  if (dn) {
    n_sets = dn[0];
    n_nodes = dn[1];
  }
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_nodes) return;
  const int tid = node_base + t;
  int lo = 0, hi = n_sets;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (sets[mid].node0 <= tid) lo = mid; else hi = mid;
  }
  const TableSet s = sets[lo];
  const int N = s.N, G = s.G;
  const double rv = gp.radius[s.grain];
  const double d = gp.axis_ratio[s.grain];
  const double lam = gp.wavelength[s.grain];
  const cdbl mr = cmk(gp.mr_re[s.grain], gp.mr_im[s.grain]);
  const double *cth = gl + s.gl_off;
  const double *wgt = cth + G;
  double *S = tables + s.off;
  double *B = S + (int64_t)kNodeScalars * G;
  const int64_t NG = (int64_t)N * G;
  double *bj = B, *by = B + NG, *bjmr = B + 2 * NG, *bjmi = B + 3 * NG;

  const double k = 2. * M_PI / lam;
  const cdbl kmr = cscale(mr, k);
  const double a = cbrt(d);
  const double a2 = a * a;
  const double d2 = d * d;
  const double d2m1 = d2 - 1.;
  {
    const int ig = tid - s.node0;
    const double c = cth[ig];
    const double w = wgt[ig];
    const double s2 = (1. - c) * (1. + c);
    const double s2i = 1. / s2;
    const double sn = sqrt(s2);
    const double r2a2 = 1. / (s2 + d2 * (c * c));
    const double r2 = a2 * r2a2;
    const double dr = r2a2 * c * sn * d2m1;
    const double r = sqrt(r2);
    const double kr = k * r * rv;
    const cdbl krmr = cmk(kmr.x * r * rv, kmr.y * r * rv);
    const double krinv = 1. / kr;
    const cdbl krmrinv = crcp(krmr);
    S[0 * G + ig] = w * r2;
    S[1 * G + ig] = dr;
    S[2 * G + ig] = sqrt(s2i);
    S[3 * G + ig] = s2i;
    S[4 * G + ig] = krinv;
    S[5 * G + ig] = krmrinv.x;
    S[6 * G + ig] = krmrinv.y;

    // j_n(kr) and j_n(m kr): both ratio recurrences in one loop from the
    // larger start order (a start above the needed one reaches the same
    // ratios), so the two dependency chains overlap
    const int ns = bessel_start(N, kr);
    const int nsm = bessel_start(N, sqrt(krmr.x * krmr.x + krmr.y * krmr.y));
    const int ntop = ns > nsm ? ns : nsm;
    const double j0 = sin(kr) * krinv;
    S[7 * G + ig] = j0;
    // sin(z) = sin x cosh y + i cos x sinh y
    const cdbl szm = cmk(sin(krmr.x) * cosh(krmr.y), cos(krmr.x) * sinh(krmr.y));
    const cdbl jm0 = cmul(szm, krmrinv);
    S[9 * G + ig] = jm0.x;
    S[10 * G + ig] = jm0.y;
    double rho_top = 0., prod = 1.;
    cdbl rhom_top = cmk(0., 0.), prodm = cmk(1., 0.);
    {
      const double zinv = krinv;
      const cdbl zinvm = krmrinv;
      double rho = kr / (2. * ntop + 1.);
      cdbl rhom = cscale(krmr, 1. / (2. * ntop + 1.));
      for (int idx = ntop - 1; idx >= 1; --idx) {
        const double f = 2. * idx + 1.;
        rho = 1. / (f * zinv - rho);
        const cdbl den = csub(cscale(zinvm, f), rhom);
        const double dinv = 1. / (den.x * den.x + den.y * den.y);
        rhom = cmk(den.x * dinv, -den.y * dinv);
        if (idx == N) {
          rho_top = rho;
          rhom_top = rhom;
        }
        if (idx < N) {
          prod *= rho;
          prodm = cmul(prodm, rhom);
        }
      }
    }
    bool parked;
    {
      const double zinv = krinv;
      double a0 = j0 * prod;  // j_{N-1}
      parked = !(fabs(a0) >= kMillerFloor);
      if (!parked) {
        double a1 = a0 * rho_top;  // j_N
        bj[(int64_t)(N - 1) * G + ig] = a1;
        for (int n = N - 1; n >= 1; --n) {
          bj[(int64_t)(n - 1) * G + ig] = a0;
          const double am = (2. * n + 1.) * zinv * a0 - a1;
          a1 = a0;
          a0 = am;
        }
      }
    }
    if (parked) {
      const double zinv = krinv;
      double rho = kr / (2. * ns + 1.);
      for (int idx = ns - 1; idx >= 1; --idx) {
        rho = 1. / ((2. * idx + 1.) * zinv - rho);
        if (idx - 1 < N) bj[(int64_t)(idx - 1) * G + ig] = rho;
      }
      double jprev = bj[ig] * j0;
      bj[ig] = jprev;
      // the parked ratios are read in chunks of 8 ahead of the chunk's
      // stores, so the loads overlap instead of waiting one by one
      for (int i0 = 1; i0 < N; i0 += 8) {
        double rho8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          rho8[u] = i0 + u < N ? bj[(int64_t)(i0 + u) * G + ig] : 0.;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int i = i0 + u;
          if (i >= N) break;
          const double ji = rho8[u] * jprev;
          bj[(int64_t)i * G + ig] = ji;
          jprev = ji;
        }
      }
    }
    // y_n(kr): forward recurrence from the closed forms of y_1, y_2.
    {
      const double cz = cos(kr), sz = sin(kr);
      const double zi = krinv, zi2 = zi * zi, zi3 = zi2 * zi;
      double ym2 = -cz * zi2 - sz * zi;
      S[8 * G + ig] = -cz * zi;
      by[ig] = ym2;
      if (N > 1) {
        double ym1 = (-3. * zi3 + zi) * cz - 3. * zi2 * sz;
        by[G + ig] = ym1;
        for (int i = 2; i < N; ++i) {
          const double yi = (2. * i + 1.) * zi * ym1 - ym2;
          by[(int64_t)i * G + ig] = yi;
          ym2 = ym1;
          ym1 = yi;
        }
      }
    }
    // j_n(m kr), as j_n(kr) in complex arithmetic
    {
      const cdbl zinv = krmrinv;
      cdbl a0 = cmul(jm0, prodm);  // j_{N-1}(m kr)
      parked = !(fmax(fabs(a0.x), fabs(a0.y)) >= kMillerFloor);
      if (!parked) {
        cdbl a1 = cmul(a0, rhom_top);  // j_N(m kr)
        bjmr[(int64_t)(N - 1) * G + ig] = a1.x;
        bjmi[(int64_t)(N - 1) * G + ig] = a1.y;
        for (int n = N - 1; n >= 1; --n) {
          const int64_t o = (int64_t)(n - 1) * G + ig;
          bjmr[o] = a0.x;
          bjmi[o] = a0.y;
          const cdbl am = csub(cmul(cscale(zinv, 2. * n + 1.), a0), a1);
          a1 = a0;
          a0 = am;
        }
      }
    }
    if (parked) {
      const cdbl zinv = krmrinv;
      cdbl rho = cscale(krmr, 1. / (2. * nsm + 1.));
      for (int idx = nsm - 1; idx >= 1; --idx) {
        const double f = 2. * idx + 1.;
        rho = crcp(csub(cscale(zinv, f), rho));
        if (idx - 1 < N) {
          bjmr[(int64_t)(idx - 1) * G + ig] = rho.x;
          bjmi[(int64_t)(idx - 1) * G + ig] = rho.y;
        }
      }
      cdbl jprev = cmul(cmk(bjmr[ig], bjmi[ig]), jm0);
      bjmr[ig] = jprev.x;
      bjmi[ig] = jprev.y;
      for (int i0 = 1; i0 < N; i0 += 8) {
        cdbl rho8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int64_t o = (int64_t)(i0 + u) * G + ig;
          rho8[u] = i0 + u < N ? cmk(bjmr[o], bjmi[o]) : cmk(0., 0.);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int i = i0 + u;
          if (i >= N) break;
          const int64_t o = (int64_t)i * G + ig;
          const cdbl ji = cmul(rho8[u], jprev);
          bjmr[o] = ji.x;
          bjmi[o] = ji.y;
          jprev = ji;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K3: Q/RgQ assembly.  Every (n1, n2) pair of an m block is a sum over the
// quadrature nodes of products of a row factor (depends on n1 and the node)
// and a column factor (n2 and the node).  The five column factors
//   R1 = a2 d2 [z j]'_2   R2 = a2 d2' [z j]'_2   R3 = a2 n2(n2+1) d2 j_2
//   R4 = a2 d2 j_2        R5 = a2 d2' j_2        (j_2 = j_n2(m kr))
// are shared by all pairs; the row factors differ for pairs with n1 + n2
// even (J21 = A1R1 + A2R2 + A3R3, J12 = A4R4 + A5R5) and odd
// (J11 = i(B5R5 + B4R4), J22 = i(B2R2 + B1R1 + B3R3)), and the RgJ sums use
// the same factors with h_n1 -> j_n1.  Then
//   Qx = k^2 m J_a + k^2 J_b  -> row (half 0, n1), class n1 mod 2
//   Qy = k^2 m J_b + k^2 J_a  -> row (half 1, n1), class (n1 + 1) mod 2
// with (J_a, J_b) = (J21, J12) for even and (J11, J22) for odd pairs
// (TMatrixResource.hpp:623-705, 993-1097).
//
// Workgroup: 256 threads = 4 waves on a tile of 16 MT rows of pairs, each
// thread accumulating an MT x MT micro-tile (MT = 1 for blocks of order up
// to 16, MT = 2 above: twice the multiply-adds per LDS operand).  For m > 0
// the tile has 16 MT columns; wave w takes rows of parity w & 1 and columns
// of parity w >> 1, so every wave sees a single pair parity; lane (lr, lc)
// takes rows 2 (lr + 8 a) + (w & 1) and columns 2 (lc + 8 b) + (w >> 1),
// a, b < MT.  For m = 0 only pairs with n1 + n2 even contribute
// (TMatrixResource.hpp:622), so the tile is 32 MT columns wide and all four
// waves take even pairs: wave w takes rows of parity w & 1 and the columns
// of the same parity in half w >> 1 of the tile.
// Workgroups find their (item, tile) by binary search in the prefix sum of
// tile counts over the cost-sorted item list.
constexpr int kLeftStride = 34;   // 32 factors + 2 pad doubles per row
constexpr int kRightStride = 10;  // 5 complex factors per column

// Row factors (32 doubles) of block row p at node ig, zero beyond the block
// or the node range.
__device__ __forceinline__ void row_factor_entry(
    double *L, int p, int ig, int G, int nm, int nmin, double mm,
    const double *S, const double *B, int64_t NG, const double *Wd,
    const double *Wdd) {
  if (ig < G && p < nm) {
    const int n1 = nmin + p;
    const double d1 = Wd[(int64_t)p * G + ig];
    const double dd1 = Wdd[(int64_t)p * G + ig];
    double j1, dj1, y1, dy1;
    bessel_kr(S, B, NG, G, n1, ig, j1, dj1, y1, dy1);
    const double wr2 = S[ig], dr = S[G + ig], s1 = S[2 * G + ig];
    const double s2 = S[3 * G + ig], krinv = S[4 * G + ig];
    const cdbl imkr = cmk(S[5 * G + ig], S[6 * G + ig]);
    const double nn1 = n1 * (n1 + 1.);
    const double a1 = sqrt((2. * n1 + 1.) / nn1);
    const cdbl h1 = cmk(j1, y1), dh1 = cmk(dj1, dy1);
    const double m2s2 = mm * mm * s2;
    const double q = dr * krinv * nn1 * d1;
    // even-pair row factors
    const double fA1 = wr2 * m2s2 * d1 * a1;
    const double fA2 = wr2 * dd1 * a1;
    const double fA3 = wr2 * dr * dd1 * a1;
    const double fA5 = -wr2 * a1;
    const cdbl A1 = cscale(h1, fA1);
    const cdbl A2 = cscale(h1, fA2);
    const cdbl A3 = cmul(imkr, cscale(h1, fA3));
    const cdbl A4 = cscale(dh1, -fA1);
    const cdbl A5 = cscale(cadd(cscale(dh1, dd1), cscale(h1, q)), fA5);
    L[0] = A1.x; L[1] = A1.y; L[2] = A2.x; L[3] = A2.y;
    L[4] = A3.x; L[5] = A3.y; L[6] = A4.x; L[7] = A4.y;
    L[8] = A5.x; L[9] = A5.y;
    L[10] = fA1 * j1;
    L[11] = fA2 * j1;
    const cdbl A3r = cscale(imkr, fA3 * j1);
    L[12] = A3r.x; L[13] = A3r.y;
    L[14] = -fA1 * dj1;
    L[15] = fA5 * (dd1 * dj1 + q * j1);
    // odd-pair row factors (without the common factor i)
    const double dsi = s1 * mm * wr2;
    const double fB5 = dsi * d1 * a1;
    const double fB4 = dsi * dd1 * a1;
    const double fB1 = dsi * a1;
    const double fB3 = dsi * dr * d1 * a1;
    const cdbl B5 = cscale(h1, fB5);
    const cdbl B4 = cscale(h1, fB4);
    const cdbl B2 = cscale(dh1, fB5);
    const cdbl B1 = cscale(cadd(cscale(dh1, dd1), cscale(h1, q)), fB1);
    const cdbl B3 = cmul(imkr, cscale(dh1, fB3));
    L[16] = B5.x; L[17] = B5.y; L[18] = B4.x; L[19] = B4.y;
    L[20] = B2.x; L[21] = B2.y; L[22] = B1.x; L[23] = B1.y;
    L[24] = B3.x; L[25] = B3.y;
    L[26] = fB5 * j1;
    L[27] = fB4 * j1;
    L[28] = fB5 * dj1;
    L[29] = fB1 * (dd1 * dj1 + q * j1);
    const cdbl B3r = cscale(imkr, fB3 * dj1);
    L[30] = B3r.x; L[31] = B3r.y;
  } else {
#pragma unroll
    for (int i = 0; i < 32; ++i) L[i] = 0.;
  }
}

// Column factors (R1..R5, 10 doubles) of block column p at node ig, zero
// beyond the block or the node range.
__device__ __forceinline__ void col_factor_entry(
    double *R, int p, int ig, int G, int nm, int nmin, const double *S,
    const double *B, int64_t NG, const double *Wd, const double *Wdd) {
  if (ig < G && p < nm) {
    const int n2 = nmin + p;
    const double d2 = Wd[(int64_t)p * G + ig];
    const double dd2 = Wdd[(int64_t)p * G + ig];
    cdbl jm, djm;
    bessel_mkr(S, B, NG, G, n2, ig, jm, djm);
    const double nn2 = n2 * (n2 + 1.);
    const double a2 = sqrt((2. * n2 + 1.) / nn2);
    const cdbl R1 = cscale(djm, a2 * d2);
    const cdbl R2 = cscale(djm, a2 * dd2);
    const cdbl R3 = cscale(jm, a2 * nn2 * d2);
    const cdbl R4 = cscale(jm, a2 * d2);
    const cdbl R5 = cscale(jm, a2 * dd2);
    R[0] = R1.x; R[1] = R1.y; R[2] = R2.x; R[3] = R2.y;
    R[4] = R3.x; R[5] = R3.y; R[6] = R4.x; R[7] = R4.y;
    R[8] = R5.x; R[9] = R5.y;
  } else {
#pragma unroll
    for (int i = 0; i < 10; ++i) R[i] = 0.;
  }
}

// Stages the row factors (32 per row and node: even-pair A1..A5, RgJ
// factors, odd-pair B5..B3, RgJ factors) of the TR rows of tile row tr and
// the column factors (R1..R5) of the cw columns of tile column tc for the
// nodes ig0 .. ig0 + CH - 1 into LDS; entries beyond the block or the node
// range are zero.
template <int TR, int CH>
__device__ __forceinline__ void stage_factors(
    double (*Lsh)[TR][kLeftStride], double (*Rsh)[2 * TR][kRightStride],
    int t, int nthreads, int ig0, int G, int nm, int nmin, int tr, int tc,
    int cw, double mm, const double *S, const double *B, int64_t NG,
    const double *Wd, const double *Wdd) {
  const int nstage = TR * CH + cw * CH;
  for (int e = t; e < nstage; e += nthreads) {
    if (e < TR * CH) {
      const int r = e / CH, kk = e % CH;
      row_factor_entry(Lsh[kk][r], tr * TR + r, ig0 + kk, G, nm, nmin, mm, S,
                       B, NG, Wd, Wdd);
    } else {
      const int tt = e - TR * CH;
      const int c = tt / CH, kk = tt % CH;
      col_factor_entry(Rsh[kk][c], tc * cw + c, ig0 + kk, G, nm, nmin, S, B,
                       NG, Wd, Wdd);
    }
  }
}

// Adds the staged nodes kbeg, kbeg + kstep, ... < kmax of one LDS stage to
// the MT x MT micro-tile of pairs (rows rr, columns cc: indices into the
// staged row and column factors) of a thread; odd selects the odd-pair
// factors.  Lb/Rb: staged row/column factors, lnode/rnode doubles per node.
template <int MT>
__device__ __forceinline__ void accumulate_nodes(
    const double *Lb, int lnode, const double *Rb, int rnode, int kbeg,
    int kmax, int kstep, bool odd, const int (&rr)[MT], const int (&cc)[MT],
    cdbl (&Ja)[MT][MT], cdbl (&Jb)[MT][MT], cdbl (&RJa)[MT][MT],
    cdbl (&RJb)[MT][MT]) {
  if (!odd) {
    for (int kk = kbeg; kk < kmax; kk += kstep) {
      cdbl R1[MT], R2[MT], R3[MT], R4[MT], R5[MT];
#pragma unroll
      for (int b = 0; b < MT; ++b) {
        const double *R = Rb + kk * rnode + cc[b] * kRightStride;
        R1[b] = cmk(R[0], R[1]); R2[b] = cmk(R[2], R[3]);
        R3[b] = cmk(R[4], R[5]); R4[b] = cmk(R[6], R[7]);
        R5[b] = cmk(R[8], R[9]);
      }
#pragma unroll
      for (int a = 0; a < MT; ++a) {
        const double *L = Lb + kk * lnode + rr[a] * kLeftStride;
        const cdbl L1 = cmk(L[0], L[1]), L2 = cmk(L[2], L[3]);
        const cdbl L3 = cmk(L[4], L[5]), L4 = cmk(L[6], L[7]);
        const cdbl L5 = cmk(L[8], L[9]), L8 = cmk(L[12], L[13]);
        const double l6 = L[10], l7 = L[11], l9 = L[14], l10 = L[15];
#pragma unroll
        for (int b = 0; b < MT; ++b) {
          cfma(Ja[a][b], L1, R1[b]);
          cfma(Ja[a][b], L2, R2[b]);
          cfma(Ja[a][b], L3, R3[b]);
          cfma(Jb[a][b], L4, R4[b]);
          cfma(Jb[a][b], L5, R5[b]);
          rfma(RJa[a][b], l6, R1[b]);
          rfma(RJa[a][b], l7, R2[b]);
          cfma(RJa[a][b], L8, R3[b]);
          rfma(RJb[a][b], l9, R4[b]);
          rfma(RJb[a][b], l10, R5[b]);
        }
      }
    }
  } else {
    for (int kk = kbeg; kk < kmax; kk += kstep) {
      cdbl R1[MT], R2[MT], R3[MT], R4[MT], R5[MT];
#pragma unroll
      for (int b = 0; b < MT; ++b) {
        const double *R = Rb + kk * rnode + cc[b] * kRightStride;
        R1[b] = cmk(R[0], R[1]); R2[b] = cmk(R[2], R[3]);
        R3[b] = cmk(R[4], R[5]); R4[b] = cmk(R[6], R[7]);
        R5[b] = cmk(R[8], R[9]);
      }
#pragma unroll
      for (int a = 0; a < MT; ++a) {
        const double *L = Lb + kk * lnode + rr[a] * kLeftStride + 16;
        const cdbl L1 = cmk(L[0], L[1]), L2 = cmk(L[2], L[3]);
        const cdbl L3 = cmk(L[4], L[5]), L4 = cmk(L[6], L[7]);
        const cdbl L5 = cmk(L[8], L[9]), L8 = cmk(L[14], L[15]);
        const double l6 = L[10], l7 = L[11], l9 = L[12], l10 = L[13];
#pragma unroll
        for (int b = 0; b < MT; ++b) {
          cfma(Ja[a][b], L1, R5[b]);
          cfma(Ja[a][b], L2, R4[b]);
          cfma(Jb[a][b], L3, R2[b]);
          cfma(Jb[a][b], L4, R1[b]);
          cfma(Jb[a][b], L5, R3[b]);
          rfma(RJa[a][b], l6, R5[b]);
          rfma(RJa[a][b], l7, R4[b]);
          rfma(RJb[a][b], l9, R2[b]);
          rfma(RJb[a][b], l10, R1[b]);
          cfma(RJb[a][b], L8, R3[b]);
        }
      }
    }
  }
}

// Writes the micro-tile's pairs (block rows pr, columns pc; those beyond the
// block skipped) of Q and RgQ into the block's class layout: Qx = k^2 m J_a +
// k^2 J_b to row class n1 mod 2, Qy = k^2 m J_b + k^2 J_a to the other
// (odd pairs carry a factor i).
template <int MT>
__device__ __forceinline__ void store_pairs(
    const GrainParams &gp, int grain, int nm, int nmin, cdbl *Q,
    const int (&pr)[MT], const int (&pc)[MT], bool odd,
    const cdbl (&Ja)[MT][MT], const cdbl (&Jb)[MT][MT],
    const cdbl (&RJa)[MT][MT], const cdbl (&RJb)[MT][MT]) {
  const double lam = gp.wavelength[grain];
  const double k = 2. * M_PI / lam;
  const double k2 = k * k;
  const cdbl k2mr = cscale(cmk(gp.mr_re[grain], gp.mr_im[grain]), k2);
  const int64_t nm2 = (int64_t)nm * nm;
#pragma unroll
  for (int a = 0; a < MT; ++a) {
#pragma unroll
    for (int b = 0; b < MT; ++b) {
      const int p1 = pr[a];
      const int p2 = pc[b];
      if (p1 >= nm || p2 >= nm) continue;
      cdbl ja = Ja[a][b], jb = Jb[a][b], rja = RJa[a][b], rjb = RJb[a][b];
      if (odd) {
        // multiply by i: (x, y) -> (-y, x)
        ja = cmk(-ja.y, ja.x);
        jb = cmk(-jb.y, jb.x);
        rja = cmk(-rja.y, rja.x);
        rjb = cmk(-rjb.y, rjb.x);
      }
      const cdbl Qx = cadd(cmul(k2mr, ja), cscale(jb, k2));
      const cdbl Qy = cadd(cmul(k2mr, jb), cscale(ja, k2));
      const cdbl RQx = cadd(cmul(k2mr, rja), cscale(rjb, k2));
      const cdbl RQy = cadd(cmul(k2mr, rjb), cscale(rja, k2));
      const int n1 = nmin + p1;
      const int cx = n1 & 1;
      const int64_t e = (int64_t)p1 * nm + p2;
      Q[cx * nm2 + e] = Qx;
      Q[(1 - cx) * nm2 + e] = Qy;
      Q[(2 + cx) * nm2 + e] = RQx;
      Q[(2 + 1 - cx) * nm2 + e] = RQy;
    }
  }
}

// Sum of a partial sum over the lanes that differ in bit 2 (columns split)
// and bit 5 (rows split) of the lane index; every lane of a group gets it.
__device__ __forceinline__ void add_lane_halves(cdbl &v, bool split_c,
                                                bool split_r) {
  if (split_c) {
    v.x += __shfl_xor(v.x, 4);
    v.y += __shfl_xor(v.y, 4);
  }
  if (split_r) {
    v.x += __shfl_xor(v.x, 32);
    v.y += __shfl_xor(v.y, 32);
  }
}

// One workgroup per tile.  dn (device list counts, may be NULL): the item
// and tile counts are read from dn[0], dn[1]; workgroups beyond the tiles
// return at once.
template <int MT>
__global__ __launch_bounds__(256, 2) void k_assemble(
    const int32_t *order, const int32_t *tile_prefix, int n_order,
    const Item *items, const TableSet *sets, const WigKey *wkeys,
    const GrainParams gp, const double *tables, const double *wig,
    cdbl *qbuf, const int32_t *dn) {
  constexpr int TR = 16 * MT;            // rows of pairs per tile
  // quadrature nodes per LDS stage: two blocks per CU for MT = 2, 3 (their
  // registers allow no more); MT = 1 (120 registers) runs four blocks per
  // CU on 28 KiB of LDS each (measured faster than 8, 6 or 2 nodes)
  constexpr int CH = MT == 1 ? 4 : MT == 2 ? 4 : 2;
  __shared__ double Lsh[CH][TR][kLeftStride];
  __shared__ double Rsh[CH][2 * TR][kRightStride];

  const int bid = blockIdx.x;
  if (dn) {
    if (bid >= dn[1]) return;
    n_order = dn[0];
  }
  int lo = 0, hi = n_order;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_prefix[mid] <= bid) lo = mid; else hi = mid;
  }
  const Item it = items[order[lo]];
  const TableSet ts = sets[it.set];
  const WigKey wk = wkeys[it.wig];
  const int N = it.N, m = it.m, G = it.G;
  const int nmin = m > 0 ? m : 1;
  const int nm = N + 1 - nmin;
  const bool m0 = (m == 0);
  const int cw = m0 ? 2 * TR : TR;
  const int ntc = (nm + cw - 1) / cw;
  const int lt = bid - tile_prefix[lo];
  const int tr = lt / ntc, tc = lt - (lt / ntc) * ntc;
  const double *S = tables + ts.off;
  const double *B = S + (int64_t)kNodeScalars * G;
  const int64_t NG = (int64_t)N * G;
  const double *Wd = wig + wk.off;
  const double *Wdd = Wd + (int64_t)nm * G;

  const int t = threadIdx.x;
  const int wv = t >> 6, lane = t & 63;
  const int prow = wv & 1;
  const int pcol = m0 ? (wv & 1) : (wv >> 1);
  int cbase = m0 ? TR * (wv >> 1) : 0;
  int lr = lane >> 3, lc = lane & 7;
  // MT = 1 (orders <= 16): threads beyond the block sum a share of the
  // nodes instead.  For m = 0 the second half of the tile's columns is
  // always beyond the block, so waves 2 and 3 take the odd nodes of each
  // stage; for orders <= 8 a lane's row and column bits 2 pick one of four
  // node groups as well.  Node group sg of ns sums the stage nodes
  // kk = sg, sg + ns, ...; the partial sums are added at the end.
  int sg = 0, ns = 1;
  // MT = 2, 3: a tile whose rows (or columns) within the block all lie in
  // its first half maps its lanes onto that half, and the lanes freed take
  // every other node of a stage (lane bit 5 for rows, lane bit 2 for
  // columns; for m = 0 waves 2 and 3 when the block ends within the tile's
  // first TR columns), at most CH node groups
  bool split_w = false, split_r = false, split_c = false;
  if (MT == 1) {
    if (m0) {
      cbase = 0;
      sg = wv >> 1;
      ns = 2;
    }
    if (nm <= 8) {
      sg = sg * 4 + ((lr >> 2) << 1) + (lc >> 2);
      ns *= 4;
      lr &= 3;
      lc &= 3;
    }
  } else {
    const int vr = nm - tr * TR, vc = nm - tc * cw;
    if (m0 && vc <= TR) {
      cbase = 0;
      sg = wv >> 1;
      ns = 2;
      split_w = true;
    }
    if (!m0 && vc <= TR / 2 && 2 * ns <= CH) {
      sg = 2 * sg + (lc >> 2);
      ns *= 2;
      lc &= 3;
      split_c = true;
    }
    if (vr <= TR / 2 && 2 * ns <= CH) {
      sg = 2 * sg + (lr >> 2);
      ns *= 2;
      lr &= 3;
      split_r = true;
    }
  }
  // pair row of micro-tile row a: (lr mod 4) + 4 a in the tile's first half
  // for lr < 4, the same in its second half for lr >= 4 (MT = 1: lr + 8 a);
  // a split keeps every lane in the first half.  Columns alike.
  int r0, c0;
  if (MT == 1) {
    r0 = lr;
    c0 = lc;
  } else {
    r0 = (lr & 3) + (split_r ? 0 : 4 * MT * (lr >> 2));
    c0 = (lc & 3) + (split_c ? 0 : 4 * MT * (lc >> 2));
  }
  constexpr int kStep = MT == 1 ? 8 : 4;
  int rr[MT], cc[MT];
#pragma unroll
  for (int a = 0; a < MT; ++a) {
    rr[a] = 2 * (r0 + kStep * a) + prow;
    cc[a] = cbase + 2 * (c0 + kStep * a) + pcol;
  }
  const bool odd = !m0 && ((prow ^ pcol) != 0);

  cdbl Ja[MT][MT], Jb[MT][MT], RJa[MT][MT], RJb[MT][MT];
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int b = 0; b < MT; ++b) {
      Ja[a][b] = cmk(0., 0.);
      Jb[a][b] = cmk(0., 0.);
      RJa[a][b] = cmk(0., 0.);
      RJb[a][b] = cmk(0., 0.);
    }
  const double mm = (double)m;

  for (int ig0 = 0; ig0 < G; ig0 += CH) {
    __syncthreads();
    stage_factors<TR, CH>(Lsh, Rsh, t, 256, ig0, G, nm, nmin, tr, tc, cw, mm,
                          S, B, NG, Wd, Wdd);
    __syncthreads();
    const int kmax = (G - ig0) < CH ? (G - ig0) : CH;
    accumulate_nodes<MT>(&Lsh[0][0][0], TR * kLeftStride, &Rsh[0][0][0],
                         2 * TR * kRightStride, sg, kmax, ns, odd, rr, cc, Ja,
                         Jb, RJa, RJb);
  }

  if (MT == 1 && ns > 1) {
    // add the node groups: lane groups by shuffles, then waves 2 and 3
    // into waves 0 and 1 through LDS (the stage buffers are free now)
    double v[8] = {Ja[0][0].x,  Ja[0][0].y,  Jb[0][0].x,  Jb[0][0].y,
                   RJa[0][0].x, RJa[0][0].y, RJb[0][0].x, RJb[0][0].y};
    if (nm <= 8) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        v[q] += __shfl_xor(v[q], 4);
        v[q] += __shfl_xor(v[q], 32);
      }
    }
    if (m0) {
      double *red = &Lsh[0][0][0];
      __syncthreads();
      if (wv >= 2) {
#pragma unroll
        for (int q = 0; q < 8; ++q) red[q * 128 + (t - 128)] = v[q];
      }
      __syncthreads();
      if (wv < 2) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] += red[q * 128 + t];
      }
    }
    Ja[0][0] = cmk(v[0], v[1]);
    Jb[0][0] = cmk(v[2], v[3]);
    RJa[0][0] = cmk(v[4], v[5]);
    RJb[0][0] = cmk(v[6], v[7]);
    if (sg != 0) return;
  }
  if (MT > 1 && ns > 1) {
    // add the node groups: lane halves by shuffles (both lanes of a pair
    // then hold the same sums), then waves 2 and 3 into waves 0 and 1
    // through the row-factor stage buffer, EPR pairs of the micro-tile a
    // round
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
      for (int b = 0; b < MT; ++b) {
        add_lane_halves(Ja[a][b], split_c, split_r);
        add_lane_halves(Jb[a][b], split_c, split_r);
        add_lane_halves(RJa[a][b], split_c, split_r);
        add_lane_halves(RJb[a][b], split_c, split_r);
      }
    if (split_w) {
      constexpr int EPR = CH * TR * kLeftStride / 128 / 8;
      static_assert(EPR >= 1, "stage buffer too small for the reduction");
      double *red = &Lsh[0][0][0];
      const int tl = t & 127;
#pragma unroll
      for (int e0 = 0; e0 < MT * MT; e0 += EPR) {
        __syncthreads();
#pragma unroll
        for (int e = e0; e < e0 + EPR && e < MT * MT; ++e) {
          const int a = e / MT, b = e % MT;
          double *r = red + (int64_t)(e - e0) * 8 * 128 + tl;
          if (wv >= 2) {
            r[0] = Ja[a][b].x;    r[128] = Ja[a][b].y;
            r[256] = Jb[a][b].x;  r[384] = Jb[a][b].y;
            r[512] = RJa[a][b].x; r[640] = RJa[a][b].y;
            r[768] = RJb[a][b].x; r[896] = RJb[a][b].y;
          }
        }
        __syncthreads();
#pragma unroll
        for (int e = e0; e < e0 + EPR && e < MT * MT; ++e) {
          const int a = e / MT, b = e % MT;
          const double *r = red + (int64_t)(e - e0) * 8 * 128 + tl;
          if (wv < 2) {
            Ja[a][b].x += r[0];    Ja[a][b].y += r[128];
            Jb[a][b].x += r[256];  Jb[a][b].y += r[384];
            RJa[a][b].x += r[512]; RJa[a][b].y += r[640];
            RJb[a][b].x += r[768]; RJb[a][b].y += r[896];
          }
        }
      }
    }
    if (sg != 0) return;
  }
  int pr[MT], pc[MT];
#pragma unroll
  for (int a = 0; a < MT; ++a) {
    pr[a] = tr * TR + rr[a];
    pc[a] = tc * cw + cc[a];
  }
  store_pairs<MT>(gp, it.grain, nm, nmin, qbuf + it.q_off, pr, pc, odd, Ja,
                  Jb, RJa, RJb);
}

// K3 for blocks of order up to kPackRows: four blocks per workgroup, each on
// a kPackRows x kPackRows tile of 3 x 3 pairs per thread (the cheaper
// micro-tile per pair), so such a block fills a larger share of its tile than
// of a 32 x 32 or 16 x 16 one.  Lane (lr, lc) of a wave serves block
// 2 (lr / 4) + lc / 4 at rows 2 (lr mod 4 + 4 a) + prow and columns
// 2 (lc mod 4 + 4 b) + pcol; wave w takes the parity class (w & 1, w >> 1)
// for m > 0 as in k_assemble.  For m = 0 only the classes (0, 0) and (1, 1)
// exist: waves 2 and 3 repeat waves 0 and 1 on the odd nodes of each stage
// and their sums are added at the end.  The four blocks (consecutive in the
// list, which is sorted by G) are staged together, two nodes per stage;
// nodes beyond a block's G stage as zero.
struct PackBlock {
  const double *S, *B, *Wd, *Wdd;
  int64_t NG;
  int nm, nmin, G, m, item, pad;
};

// dn (may be NULL): the block count is read from dn[0]; workgroups beyond
// its groups of four return at once.
__global__ __launch_bounds__(256, 2) void k_assemble_pack4(
    const int32_t *order, int n_order, int m0, const Item *items,
    const TableSet *sets, const WigKey *wkeys, const GrainParams gp,
    const double *tables, const double *wig, cdbl *qbuf, const int32_t *dn) {
  constexpr int MT = 3, CH = 2, NB = 4, PR = kPackRows;
  __shared__ double Lsh[CH][NB * PR][kLeftStride];
  __shared__ double Rsh[CH][NB * PR][kRightStride];
  __shared__ PackBlock blk[NB];
  if (dn) {
    n_order = dn[0];
    if ((int)blockIdx.x * NB >= n_order) return;
  }
  const int t = threadIdx.x;
  if (t < NB) {
    const int q = blockIdx.x * NB + t;
    PackBlock pb{};
    pb.item = -1;
    if (q < n_order) {
      const Item it = items[order[q]];
      const TableSet ts = sets[it.set];
      const WigKey wk = wkeys[it.wig];
      pb.m = it.m;
      pb.nmin = it.m > 0 ? it.m : 1;
      pb.nm = it.N + 1 - pb.nmin;
      pb.G = it.G;
      pb.S = tables + ts.off;
      pb.B = pb.S + (int64_t)kNodeScalars * it.G;
      pb.NG = (int64_t)it.N * it.G;
      pb.Wd = wig + wk.off;
      pb.Wdd = pb.Wd + (int64_t)pb.nm * it.G;
      pb.item = order[q];
    }
    blk[t] = pb;
  }
  __syncthreads();
  const int gmax = max(max(blk[0].G, blk[1].G), max(blk[2].G, blk[3].G));
  const int wv = t >> 6, lane = t & 63;
  const int lr = lane >> 3, lc = lane & 7;
  const int b = ((lr >> 2) << 1) | (lc >> 2);
  const int prow = wv & 1;
  const int pcol = m0 ? prow : (wv >> 1);
  const int sg = m0 ? (wv >> 1) : 0, ns = m0 ? 2 : 1;
  int rr[MT], cc[MT];
#pragma unroll
  for (int a = 0; a < MT; ++a) {
    rr[a] = b * PR + 2 * ((lr & 3) + 4 * a) + prow;
    cc[a] = b * PR + 2 * ((lc & 3) + 4 * a) + pcol;
  }
  const bool odd = !m0 && ((prow ^ pcol) != 0);
  cdbl Ja[MT][MT], Jb[MT][MT], RJa[MT][MT], RJb[MT][MT];
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int c = 0; c < MT; ++c) {
      Ja[a][c] = cmk(0., 0.);
      Jb[a][c] = cmk(0., 0.);
      RJa[a][c] = cmk(0., 0.);
      RJb[a][c] = cmk(0., 0.);
    }
  for (int ig0 = 0; ig0 < gmax; ig0 += CH) {
    __syncthreads();
    for (int e = t; e < 2 * NB * PR * CH; e += 256) {
      if (e < NB * PR * CH) {
        const int r = e / CH, kk = e - (e / CH) * CH;
        const int bb = r / PR, p = r - bb * PR;
        const PackBlock &pb = blk[bb];
        row_factor_entry(Lsh[kk][r], p, ig0 + kk, pb.G, pb.nm, pb.nmin,
                         (double)pb.m, pb.S, pb.B, pb.NG, pb.Wd, pb.Wdd);
      } else {
        const int tt = e - NB * PR * CH;
        const int c = tt / CH, kk = tt - (tt / CH) * CH;
        const int bb = c / PR, p = c - bb * PR;
        const PackBlock &pb = blk[bb];
        col_factor_entry(Rsh[kk][c], p, ig0 + kk, pb.G, pb.nm, pb.nmin, pb.S,
                         pb.B, pb.NG, pb.Wd, pb.Wdd);
      }
    }
    __syncthreads();
    const int kmax = (gmax - ig0) < CH ? (gmax - ig0) : CH;
    accumulate_nodes<MT>(&Lsh[0][0][0], NB * PR * kLeftStride, &Rsh[0][0][0],
                         NB * PR * kRightStride, sg, kmax, ns, odd, rr, cc, Ja,
                         Jb, RJa, RJb);
  }
  if (m0) {
    // add the odd-node sums of waves 2 and 3 to waves 0 and 1 (J, then RgJ)
    double *red = &Lsh[0][0][0];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      __syncthreads();
      if (wv >= 2) {
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
          for (int c = 0; c < MT; ++c) {
            const cdbl u = half ? RJa[a][c] : Ja[a][c];
            const cdbl v = half ? RJb[a][c] : Jb[a][c];
            const int q = (a * MT + c) * 4;
            red[(q + 0) * 128 + t - 128] = u.x;
            red[(q + 1) * 128 + t - 128] = u.y;
            red[(q + 2) * 128 + t - 128] = v.x;
            red[(q + 3) * 128 + t - 128] = v.y;
          }
      }
      __syncthreads();
      if (wv < 2) {
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
          for (int c = 0; c < MT; ++c) {
            const int q = (a * MT + c) * 4;
            cdbl &u = half ? RJa[a][c] : Ja[a][c];
            cdbl &v = half ? RJb[a][c] : Jb[a][c];
            u.x += red[(q + 0) * 128 + t];
            u.y += red[(q + 1) * 128 + t];
            v.x += red[(q + 2) * 128 + t];
            v.y += red[(q + 3) * 128 + t];
          }
      }
    }
    if (wv >= 2) return;
  }
  const PackBlock &pb = blk[b];
  if (pb.item < 0) return;
  const Item it = items[pb.item];
  int pr[MT], pc[MT];
#pragma unroll
  for (int a = 0; a < MT; ++a) {
    pr[a] = rr[a] - b * PR;
    pc[a] = cc[a] - b * PR;
  }
  store_pairs<MT>(gp, it.grain, pb.nm, pb.nmin, qbuf + it.q_off, pr, pc, odd,
                  Ja, Jb, RJa, RJb);
}

// ---------------------------------------------------------------------------
// K4: T_c = -RgQ_c Q_c^-1 for one (item, class).  Q_c is inverted in place
// by Gauss-Jordan elimination with partial pivoting on |Q_ij| (largest
// modulus, first index on ties: the rule of Matrix.hpp:280-293); the column
// exchanges of the final unpermutation (Matrix.hpp:367-383) are folded into
// the product's column index.  For m = 0 the class block splits further into
// the rows of even and odd p (only n1 + n2 even couples,
// TMatrixResource.hpp:622): the two sub-blocks are stored apart (permuted
// index) and eliminated concurrently.  The product streams RgQ_c from HBM
// in groups of four rows per wave.  With diag_only only the diagonal of T is
// formed (all the convergence test needs).
// Shared end of the solve kernels: the column unpermutation of the
// Gauss-Jordan inverse Ainv (held in A, LDS or global), the product
// T_c = -RgQ_c Ainv (or only its diagonal) and the per-block sums.  RgQ is
// read from HBM ahead of its use: a wave loads its four rows of the product
// at once, coalesced (lane l holds column lo + l + 64 h of each row, KH
// slots h), and broadcasts element k by readlane, so the memory latency is
// paid once per row group instead of once per k; the diagonal loads its
// row in batches of 16.
template <int NT, int KH>
__device__ __forceinline__ void solve_tail(
    const Item &it, int task, int cls, int nm, int nb, int sz0,
    const int *off, const int *sz, const cdbl *A0, int ld0, const cdbl *A1,
    int ld1, const int *piv, int *src,
    const cdbl *RQg, cdbl *tbuf, double *qparts, int diag_only,
    double (*red)[4], cdbl *ainv_out) {
  constexpr int NW = NT / 64;
  const int m = it.m;
  const int nmin = m > 0 ? m : 1;
  const int64_t nm2 = (int64_t)nm * nm;
  auto pidx = [&](int i) {
    return nb == 2 ? (i < sz0 ? 2 * i : 2 * (i - sz0) + 1) : i;
  };
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  // element (k, c) of sub-block b: sub-block b starts at Ab[b] (its element
  // (off[b], off[b])) with row stride ld[b]
  const cdbl *Ab[2] = {A0, A1};
  const int ld[2] = {ld0, ld1};
  auto aget = [&](int b, int k, int c) -> cdbl {
    return Ab[b][(int64_t)(k - off[b]) * ld[b] + (c - off[b])];
  };
  if (t == 0) {
    for (int j = 0; j < nm; ++j) src[j] = j;
    for (int k = nm - 1; k >= 0; --k) {
      const int p = piv[k];
      if (p != k) {
        const int tmp = src[k];
        src[k] = src[p];
        src[p] = tmp;
      }
    }
  }
  __syncthreads();
  if (ainv_out) {
    // Q_c^-1 in the class layout: Q^-1(pidx(k), pidx(c)) = A(k, src(c))
    cdbl *Ai = ainv_out + it.t_off + cls * nm2;
    for (int k = t >> 5; k < nm; k += NT >> 5) {
      const bool bk = k >= sz0;
      for (int c = t & 31; c < nm; c += 32) {
        const bool same = nb == 1 || (c >= sz0) == bk;
        Ai[(int64_t)pidx(k) * nm + pidx(c)] =
            same ? aget(bk ? 1 : 0, k, src[c]) : cmk(0., 0.);
      }
    }
  }

  double s_diag2 = 0., e_diag = 0., s_all = 0., e_tr = 0.;
  if (diag_only) {
    // T(p, p) = -sum_k RgQ(p, k) Ainv(k, p) within the sub-block of p
    for (int i = t; i < nm; i += NT) {
      const int klo = i >= sz0 ? sz0 : 0, khi = i >= sz0 ? nm : sz0;
      const int pi = pidx(i);
      const cdbl *rq = RQg + (int64_t)pi * nm;
      const int sc = src[i];
      const int bi = (nb == 2 && i >= sz0) ? 1 : 0;
      cdbl acc = cmk(0., 0.);
      for (int kb = klo; kb < khi; kb += 16) {
        cdbl v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
          v[u] = kb + u < khi ? rq[pidx(kb + u)] : cmk(0., 0.);
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (kb + u < khi) cfma(acc, v[u], aget(bi, kb + u, sc));
      }
      const double w = 2. * (nmin + pi) + 1.;
      const double a2 = acc.x * acc.x + acc.y * acc.y;
      s_diag2 += w * a2;
      e_diag -= w * acc.x;
      e_tr -= acc.x;
    }
  } else {
    cdbl *T = tbuf + it.t_off + cls * nm2;
    if (nb == 2) {
      // the two sub-blocks do not couple: zero the cross entries
      for (int i = t >> 5; i < nm; i += NT >> 5) {
        const int bi = i >= sz0;
        for (int j = t & 31; j < nm; j += 32) {
          if ((j >= sz0) != bi)
            T[(int64_t)pidx(i) * nm + pidx(j)] = cmk(0., 0.);
        }
      }
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (b >= nb) continue;
      const int lo = off[b], hi = off[b] + sz[b];
      // columns in groups of 128 (two per lane), rows in groups of four
      for (int cb = lo; cb < hi; cb += 128) {
      const int c0 = cb + lane, c1 = cb + lane + 64;
      const bool v0 = c0 < hi, v1 = c1 < hi;
      const int s0 = v0 ? src[c0] : lo, s1 = v1 ? src[c1] : lo;
      for (int rg = lo + 4 * wv; rg < hi; rg += 4 * NW) {
        cdbl acc[4][2];
        int prow[4];
        cdbl rq[4][KH];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc[q][0] = cmk(0., 0.);
          acc[q][1] = cmk(0., 0.);
          const int r = rg + q < hi ? rg + q : rg;
          prow[q] = pidx(r);
#pragma unroll
          for (int h = 0; h < KH; ++h) {
            const int k = lo + lane + 64 * h;
            rq[q][h] = k < hi ? RQg[(int64_t)prow[q] * nm + pidx(k)]
                              : cmk(0., 0.);
          }
        }
#pragma unroll
        for (int h = 0; h < KH; ++h) {
          const int kb = lo + 64 * h;
          if (kb >= hi) break;
          const int ke = min(hi, kb + 64);
          for (int k = kb; k < ke; ++k) {
            const cdbl a0 = aget(b, k, s0);
            const cdbl a1 = aget(b, k, s1);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const cdbl bq = readlane_e(rq[q][h], k - kb);
              cfma(acc[q][0], bq, a0);
              cfma(acc[q][1], bq, a1);
            }
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = rg + q;
          if (r >= hi) break;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const bool valid = h == 0 ? v0 : v1;
            if (!valid) continue;
            const int c = h == 0 ? c0 : c1;
            const cdbl v = cmk(-acc[q][h].x, -acc[q][h].y);
            T[(int64_t)prow[q] * nm + pidx(c)] = v;
            const double a2 = v.x * v.x + v.y * v.y;
            s_all += a2;
            if (r == c) {
              const double w = 2. * (nmin + prow[q]) + 1.;
              s_diag2 += w * a2;
              e_diag += w * v.x;
              e_tr += v.x;
            }
          }
        }
      }
      }
    }
  }
  // deterministic block reduction: waves by shuffles, then thread 0
  double v[4] = {s_diag2, e_diag, s_all, e_tr};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wv][q] = v[q];
  }
  __syncthreads();
  if (t == 0) {
    for (int q = 0; q < 4; ++q) {
      double acc = 0.;
      for (int w = 0; w < NW; ++w) acc += red[w][q];
      qparts[(int64_t)task * kQParts + q] = acc;
    }
  }
}

// K4 for blocks up to order 16: the elimination one column per step with
// the matrix held in registers.  Thread (ty, tx) of a TRR x TCC grid owns
// rows ty + TRR a (a < RPT) and columns tx + TCC c (c < CPT).  A step
// exchanges only the pivot row, the old row k and column k through LDS
// (3 nm values); the pivot value of the next step travels with the
// (magnitude, row) reduction.  The inverse is stored
// to LDS once at the end for the product.
// Only the one-wave shape <64, 8, 2, 2> is launched and tested; a two-wave
// shape (<128, 16, 4, 2>, orders up to 32) returned wrong values in a timing
// experiment, cause not found, so another shape needs its own parity test.
// One workgroup per task.  dn (may be NULL): the task count is read from
// dn[0]; workgroups beyond it return at once.
template <int NT, int TCC, int RPT, int CPT>
__global__ __launch_bounds__(NT) void k_invert_solve_reg(
    const int32_t *tasks, const Item *items, const cdbl *qbuf, cdbl *tbuf,
    double *qparts, int diag_only, cdbl *ainv_out, const int32_t *dn) {
  constexpr int NW = NT / 64;
  constexpr int TRR = NT / TCC;
  extern __shared__ double smem[];
  __shared__ double wb_mag[NW][2];
  __shared__ int wb_row[NW][2];
  __shared__ cdbl wb_val[NW][2];
  __shared__ double red[NW][4];

  if (dn && (int)blockIdx.x >= dn[0]) return;
  const int task = tasks[blockIdx.x];
  const int item_index = task >> 1;
  const int cls = task & 1;
  const Item it = items[item_index];
  const int m = it.m;
  const int nmin = m > 0 ? m : 1;
  const int nm = it.N + 1 - nmin;
  const int64_t nm2 = (int64_t)nm * nm;
  const int nb = (m == 0 && nm > 1) ? 2 : 1;
  const int sz0 = (nb == 2) ? (nm + 1) / 2 : nm;
  const int off[2] = {0, sz0};
  const int sz[2] = {sz0, nm - sz0};
  auto pidx = [&](int i) {
    return nb == 2 ? (i < sz0 ? 2 * i : 2 * (i - sz0) + 1) : i;
  };
  auto blk = [&](int i) { return (nb == 2 && i >= sz0) ? 1 : 0; };

  cdbl *A = reinterpret_cast<cdbl *>(smem);
  cdbl *colk = A + nm2;
  cdbl *rowk = colk + nm;
  cdbl *rowold = rowk + nm;
  int *piv = reinterpret_cast<int *>(rowold + nm);
  int *src = piv + nm;

  const cdbl *Qg = qbuf + it.q_off + cls * nm2;
  const cdbl *RQg = qbuf + it.q_off + (2 + cls) * nm2;
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int tx = t % TCC, ty = t / TCC;

  cdbl R[RPT][CPT];
#pragma unroll
  for (int a = 0; a < RPT; ++a) {
    const int i = ty + TRR * a;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int j = tx + TCC * c;
      R[a][c] = (i < nm && j < nm && blk(i) == blk(j))
                    ? Qg[(int64_t)pidx(i) * nm + pidx(j)]
                    : cmk(0., 0.);
    }
  }

  // wave's best candidate (largest magnitude, then lowest row): a DPP
  // maximum of the magnitude key (bits of |a|^2 + 1, 0 = no candidate),
  // then of the complemented row among the lanes holding that key
  auto wave_best = [&](double &mag, int &row, cdbl &val) {
    const unsigned long long key =
        mag >= 0. ? (unsigned long long)__double_as_longlong(mag) + 1ull
                  : 0ull;
    const unsigned long long mx = wave_max_u64(key);
    if (mx == 0ull) {
      mag = -1.;
      row = nm;
      val = cmk(0., 0.);
      return;
    }
    const unsigned long long rk =
        key == mx ? (unsigned long long)(0xffffffffu - (unsigned)row) : 0ull;
    const unsigned long long rmx = wave_max_u64(rk);
    const unsigned long long hit = __ballot(key == mx && rk == rmx);
    const int src_lane = (int)__builtin_ctzll(hit);
    mag = __longlong_as_double((long long)(mx - 1ull));
    row = (int)(0xffffffffu - (unsigned)rmx);
    val = readlane_e(val, src_lane);
  };
  auto publish = [&](double *bm, int *br, cdbl *bv) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (b >= nb) {
        bm[b] = -1.;
        br[b] = nm;
        bv[b] = cmk(0., 0.);
      } else {
        wave_best(bm[b], br[b], bv[b]);
      }
      if (lane == 0) {
        wb_mag[wv][b] = bm[b];
        wb_row[wv][b] = br[b];
        wb_val[wv][b] = bv[b];
      }
    }
  };
  // pivot candidates of the first column of each sub-block
  {
    double bm[2] = {-1., -1.};
    int br[2] = {nm, nm};
    cdbl bv[2] = {cmk(0., 0.), cmk(0., 0.)};
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (b >= nb) continue;
      const int k = off[b];
      if (tx != k % TCC) continue;
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        if (c != k / TCC) continue;
#pragma unroll
        for (int a = 0; a < RPT; ++a) {
          const int i = ty + TRR * a;
          if (i < off[b] || i >= off[b] + sz[b]) continue;
          const cdbl v = R[a][c];
          const double mag = v.x * v.x + v.y * v.y;
          if (mag > bm[b] || (mag == bm[b] && i < br[b])) {
            bm[b] = mag;
            br[b] = i;
            bv[b] = v;
          }
        }
      }
    }
    publish(bm, br, bv);
  }
  __syncthreads();

  for (int s = 0; s < sz0; ++s) {
    int kb[2], pb[2];
    bool act[2];
    cdbl inv[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      act[b] = b < nb && s < sz[b];
      kb[b] = off[b] + s;
      double best = -1.;
      int row = nm;
      cdbl val = cmk(1., 0.);
      for (int w = 0; w < NW; ++w) {
        const double mg = wb_mag[w][b];
        const int rw = wb_row[w][b];
        if (mg > best || (mg == best && rw < row)) {
          best = mg;
          row = rw;
          val = wb_val[w][b];
        }
      }
      pb[b] = act[b] ? row : kb[b];
      inv[b] = act[b] ? crcp(val) : cmk(1., 0.);
    }
    // phase A: pivot row (scaled), old row k, column k after the swap
#pragma unroll
    for (int a = 0; a < RPT; ++a) {
      const int i = ty + TRR * a;
      if (i >= nm) continue;
      const int b = blk(i);
      if (!act[b]) continue;
      const int k = kb[b], p = pb[b];
      if (i == p || i == k) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          const int j = tx + TCC * c;
          if (j >= nm || blk(j) != b) continue;
          if (i == p) rowk[j] = cmul(R[a][c], inv[b]);
          if (i == k) rowold[j] = R[a][c];
        }
      }
      if (tx == k % TCC) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          if (c != k / TCC) continue;
          const int dst = (i == p) ? k : (i == k ? p : i);
          colk[dst] = R[a][c];
        }
      }
    }
    if (t == 0) {
#pragma unroll
      for (int b = 0; b < 2; ++b)
        if (act[b]) piv[kb[b]] = pb[b];
    }
    __syncthreads();
    // phase B: update, and candidates for the next pivot
    double bm[2] = {-1., -1.};
    int br[2] = {nm, nm};
    cdbl bv[2] = {cmk(0., 0.), cmk(0., 0.)};
#pragma unroll
    for (int a = 0; a < RPT; ++a) {
      const int i = ty + TRR * a;
      if (i >= nm) continue;
      const int b = blk(i);
      if (!act[b]) continue;
      const int k = kb[b], p = pb[b];
      const cdbl ci = colk[i];
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int j = tx + TCC * c;
        if (j >= nm || blk(j) != b) continue;
        cdbl v;
        if (i == k) {
          v = (j == k) ? inv[b] : rowk[j];
        } else {
          const cdbl base = (i == p) ? rowold[j] : R[a][c];
          if (j == k) {
            v = cmul(ci, inv[b]);
            v.x = -v.x;
            v.y = -v.y;
          } else {
            const cdbl g = rowk[j];
            v.x = base.x - (ci.x * g.x - ci.y * g.y);
            v.y = base.y - (ci.x * g.y + ci.y * g.x);
          }
        }
        R[a][c] = v;
        if (j == k + 1 && i > k) {
          const double mag = v.x * v.x + v.y * v.y;
          if (mag > bm[b] || (mag == bm[b] && i < br[b])) {
            bm[b] = mag;
            br[b] = i;
            bv[b] = v;
          }
        }
      }
    }
    publish(bm, br, bv);
    __syncthreads();
  }
  // the inverse to LDS for the product
#pragma unroll
  for (int a = 0; a < RPT; ++a) {
    const int i = ty + TRR * a;
    if (i >= nm) continue;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int j = tx + TCC * c;
      if (j < nm) A[(int64_t)i * nm + j] = R[a][c];
    }
  }
  __syncthreads();
  solve_tail<NT, 1>(it, task, cls, nm, nb, sz0, off, sz, A, nm,
                 nb == 2 ? A + (int64_t)sz0 * nm + sz0 : A, nm, piv, src, RQg,
                 tbuf, qparts, diag_only, red, ainv_out);
}

// K4 for orders 17..200: blocked Gauss-Jordan (ctm_gj.hpp).
// GLOBAL: the sub-blocks live in gscratch (gstride complex values per
// workgroup) instead of LDS, for blocks above the LDS capacity; the panel
// and update passes then stream the matrix through L2 once per W columns.
// dn (may be NULL): the task count is read from dn[0]; workgroups beyond
// it return at once.
template <int NT, int RS, int W, bool GLOBAL>
__global__ __launch_bounds__(NT) void k_invert_solve_blk(
    const int32_t *tasks, const Item *items, const cdbl *qbuf, cdbl *tbuf,
    double *qparts, int diag_only, cdbl *gscratch, int64_t gstride,
    cdbl *ainv_out, const int32_t *dn) {
  constexpr int NW = NT / 64;
  static_assert(NW >= 4, "two waves per sub-block at least");
  extern __shared__ double smem[];
  __shared__ int stp[2][2][W];
  __shared__ double red[NW][4];

  if (dn && (int)blockIdx.x >= dn[0]) return;
  const int task = tasks[blockIdx.x];
  const int item_index = task >> 1;
  const int cls = task & 1;
  const Item it = items[item_index];
  const int m = it.m;
  const int nmin = m > 0 ? m : 1;
  const int nm = it.N + 1 - nmin;
  const int64_t nm2 = (int64_t)nm * nm;
  const int nb = (m == 0 && nm > 1) ? 2 : 1;
  const int sz0 = (nb == 2) ? (nm + 1) / 2 : nm;
  const int off[2] = {0, sz0};
  const int sz[2] = {sz0, nm - sz0};
  auto pidx = [&](int i) {
    return nb == 2 ? (i < sz0 ? 2 * i : 2 * (i - sz0) + 1) : i;
  };
  // the sub-blocks stored one after the other, each sz[b] x sz[b] (for
  // m = 0 half the LDS of the full block, whose cross blocks are zero)
  cdbl *A = GLOBAL ? gscratch + (int64_t)blockIdx.x * gstride
                  : reinterpret_cast<cdbl *>(smem);
  // rows of an LDS sub-block padded to an odd number of elements: the
  // column accesses of the panel and update steps (lane i at row i) then
  // fall in distinct bank groups (an even stride that is a multiple of 16
  // elements puts every lane in one)
  const int ldb[2] = {GLOBAL ? sz[0] : blk_lds_stride(sz[0]),
                      GLOBAL ? sz[1] : blk_lds_stride(sz[1])};
  cdbl *Ab[2] = {A, A + (int64_t)sz0 * ldb[0]};
  const int64_t nstore = (int64_t)sz[0] * ldb[0] + (int64_t)sz[1] * ldb[1];
  int *piv = GLOBAL ? reinterpret_cast<int *>(smem)
                    : reinterpret_cast<int *>(A + nstore);
  int *src = piv + nm;
  const cdbl *Qg = qbuf + it.q_off + cls * nm2;
  const cdbl *RQg = qbuf + it.q_off + (2 + cls) * nm2;
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;

#pragma unroll
  for (int bb = 0; bb < 2; ++bb) {
    if (bb >= nb) continue;
    const int o = off[bb], n = sz[bb];
    for (int i = t >> 5; i < n; i += NT >> 5) {
      const int pi = pidx(o + i);
      for (int j = t & 31; j < n; j += 32)
        Ab[bb][(int64_t)i * ldb[bb] + j] = Qg[(int64_t)pi * nm + pidx(o + j)];
    }
  }
  __syncthreads();
  // waves [b * GW, (b + 1) * GW) serve sub-block b; the first is its panel
  // wave
  const int GW = NW / nb;
  const int b = wv / GW, gw = wv - b * GW;
  const int lo = off[b], hi = off[b] + sz[b];
  const int npan = (sz[b] + W - 1) / W;
  const int npmax = (sz[0] + W - 1) / W;
  cdbl *Av = Ab[b];
  const int ld = ldb[b];
  if (gw == 0)
    blk_panel<RS, W>(Av, ld, lo, hi, lo, min(W, sz[b]), lane, piv, stp[0][b]);
  __syncthreads();
  for (int P = 0; P < npmax; ++P) {
    if (P < npan) {
      const int c0 = lo + P * W, w = min(W, hi - c0);
      const int c1 = c0 + w, w1 = min(W, hi - c1);
      const int *steps = stp[P & 1][b];
      const bool ahead = P + 1 < npan;
      if (ahead && gw == 0) {
        blk_update<RS, W>(Av, ld, lo, hi, c0, w, steps, c1, c1 + w1, 1, 0, 0,
                          lane);
        blk_panel<RS, W>(Av, ld, lo, hi, c1, w1, lane, piv,
                         stp[(P + 1) & 1][b]);
      } else {
        // columns outside the panel (and the next panel, if any)
        const int nwk = ahead ? GW - 1 : GW, rank = ahead ? gw - 1 : gw;
        blk_update<RS, W>(Av, ld, lo, hi, c0, w, steps, lo + rank, hi, nwk,
                          c0, ahead ? c1 + w1 : c1, lane);
      }
    }
    __syncthreads();
  }
  solve_tail<NT, GLOBAL ? 4 : 2>(it, task, cls, nm, nb, sz0, off, sz, Ab[0],
                                 ldb[0], Ab[1],
                 ldb[1], piv, src, RQg, tbuf, qparts, diag_only, red,
                 ainv_out);
}

// ---------------------------------------------------------------------------
// K5: convergence test over one window of orders for each active grain.
// Items of grain g for the window are win_items[win_off[a] .. win_off[a+1])
// in ascending order.
__global__ void k_converge(const int32_t *active, int n_active,
                           const int32_t *win_off, const int32_t *win_items,
                           const Item *items, const double *qparts,
                           double tol, int last_order, ConvState *state) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_active) return;
  const int g = active[a];
  ConvState st = state[g];
  if (st.done) return;
  converge_items(st, win_off[a], win_off[a + 1], win_items, items, qparts,
                 tol, last_order);
  state[g] = st;
}

// Per-grain full-T sums in m order (TMatrix.hpp:2055-2110):
// Q_sca = sum_m f_m sum_c |T_c|^2, Q_ext = sum_m f_m sum_c tr Re T_c.
__global__ void k_full_sums(int n_grains, const int32_t *grain_items_off,
                            const int32_t *grain_items, const double *qparts,
                            double *qsca, double *qext) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_grains) return;
  double s = 0., e = 0.;
  for (int q = grain_items_off[g]; q < grain_items_off[g + 1]; ++q) {
    const int ii = grain_items[q];
    const double f = (q == grain_items_off[g]) ? 1. : 2.;
    const double *p0 = qparts + (int64_t)(2 * ii) * kQParts;
    const double *p1 = p0 + kQParts;
    s += f * (p0[2] + p1[2]);
    e += f * (p0[3] + p1[3]);
  }
  if (qsca) qsca[g] = s;
  if (qext) qext[g] = e;
}

__global__ void k_write_orders(int n_grains, const ConvState *state,
                               uint32_t *nmax_out, int32_t *status_out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_grains) return;
  if (nmax_out) nmax_out[g] = (uint32_t)state[g].nmax;
  if (status_out) status_out[g] = state[g].status;
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side).

hipError_t launch_gauss_legendre(const GlKey *keys, int n, double *gl,
                                 hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gauss_legendre, dim3(n), dim3(256), 0, s, keys, gl);
  return hipGetLastError();
}

hipError_t launch_wigner(const WigKey *keys, int n, const double *gl,
                         double *wig, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wigner, dim3(n), dim3(128), 0, s, keys, gl, wig);
  return hipGetLastError();
}

hipError_t launch_tables(const TableSet *sets, int n, int node_base,
                         int n_nodes, const GrainParams &gp, const double *gl,
                         double *tables, hipStream_t s, const int32_t *dn,
                         int grid) {
  if (!dn && (n <= 0 || n_nodes <= 0)) return hipSuccess;
  hipLaunchKernelGGL(k_tables, dim3(dn ? grid : (n_nodes + 127) / 128),
                     dim3(128), 0, s, sets, n, node_base, n_nodes, gp, gl,
                     tables, dn);
  return hipGetLastError();
}

hipError_t launch_assemble(int micro, const int32_t *order,
                           const int32_t *tile_prefix, int n_order,
                           int n_tiles, const Item *items,
                           const TableSet *sets, const WigKey *wkeys,
                           const GrainParams &gp, const double *tables,
                           const double *wig, cdbl *qbuf, hipStream_t s,
                           const int32_t *dn, int grid) {
  if (!dn && n_tiles <= 0) return hipSuccess;
  const dim3 g(dn ? grid : n_tiles);
  if (micro == 3)
    hipLaunchKernelGGL(k_assemble<3>, g, dim3(256), 0, s, order, tile_prefix,
                       n_order, items, sets, wkeys, gp, tables, wig, qbuf,
                       dn);
  else if (micro == 2)
    hipLaunchKernelGGL(k_assemble<2>, g, dim3(256), 0, s, order, tile_prefix,
                       n_order, items, sets, wkeys, gp, tables, wig, qbuf,
                       dn);
  else
    hipLaunchKernelGGL(k_assemble<1>, g, dim3(256), 0, s, order, tile_prefix,
                       n_order, items, sets, wkeys, gp, tables, wig, qbuf,
                       dn);
  return hipGetLastError();
}

hipError_t launch_assemble_pack4(const int32_t *order, int n_order, int m0,
                                 const Item *items, const TableSet *sets,
                                 const WigKey *wkeys, const GrainParams &gp,
                                 const double *tables, const double *wig,
                                 cdbl *qbuf, hipStream_t s,
                                 const int32_t *dn, int grid) {
  if (!dn && n_order <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_assemble_pack4, dim3(dn ? grid : (n_order + 3) / 4),
                     dim3(256), 0, s, order, n_order, m0, items, sets, wkeys,
                     gp, tables, wig, qbuf, dn);
  return hipGetLastError();
}

size_t invert_solve_lds_bytes(int nm) {
  return (size_t)nm * nm * sizeof(cdbl) + 3 * (size_t)nm * sizeof(cdbl) +
         2 * (size_t)nm * sizeof(int);
}

template <int NT, int TCC, int RPT, int CPT>
static hipError_t launch_solve_reg(const int32_t *tasks, int n, size_t lds,
                                   const Item *items, const cdbl *qbuf,
                                   cdbl *tbuf, double *qparts, int diag_only,
                                   cdbl *ainv_out, hipStream_t s,
                                   const int32_t *dn, int grid) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(k_invert_solve_reg<NT, TCC, RPT, CPT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_invert_solve_reg<NT, TCC, RPT, CPT>),
                     dim3(dn ? grid : n), dim3(NT), lds, s, tasks, items,
                     qbuf, tbuf, qparts, diag_only, ainv_out, dn);
  return hipGetLastError();
}

template <int NT, int RS, int W, bool GLOBAL = false>
static hipError_t launch_solve_blk(const int32_t *tasks, int n, size_t lds,
                                   const Item *items, const cdbl *qbuf,
                                   cdbl *tbuf, double *qparts, int diag_only,
                                   cdbl *ainv_out, hipStream_t s,
                                   const int32_t *dn, int grid,
                                   cdbl *gscratch = nullptr,
                                   int64_t gstride = 0) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(k_invert_solve_blk<NT, RS, W, GLOBAL>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_invert_solve_blk<NT, RS, W, GLOBAL>),
                     dim3(dn ? grid : n), dim3(NT), lds, s, tasks, items,
                     qbuf, tbuf, qparts, diag_only, gscratch, gstride,
                     ainv_out, dn);
  return hipGetLastError();
}

hipError_t launch_invert_solve(const int32_t *tasks, int n, int nm_max,
                               int64_t store_max, int rows_max,
                               const Item *items, const cdbl *qbuf, cdbl *tbuf,
                               double *qparts, int diag_only, cdbl *gscratch,
                               cdbl *ainv_out, hipStream_t s,
                               const int32_t *dn, int grid) {
  if (!dn && n <= 0) return hipSuccess;
  if (nm_max > kMaxLdsOrder) {
    // the matrix lives in gscratch (n * nm_max^2 complex values): blocked
    // elimination with four rows per lane (sub-blocks up to 256 rows), the
    // pivot bookkeeping in LDS
    if (rows_max > 4 * 64 || store_max > (int64_t)nm_max * nm_max)
      return hipErrorInvalidValue;
    return launch_solve_blk<512, 4, 4, true>(
        tasks, n, 2 * (size_t)nm_max * sizeof(int), items, qbuf, tbuf, qparts,
        diag_only, ainv_out, s, dn, grid, gscratch,
        (int64_t)nm_max * nm_max);
  }
  // the blocked kernel stores the sub-blocks compactly (store_max complex
  // values at most) and needs two rows per lane above 64 rows
  const size_t lds_blk =
      (size_t)store_max * sizeof(cdbl) + 2 * (size_t)nm_max * sizeof(int);
  // register-resident elimination: grid and elements per thread by size
  const size_t lds = invert_solve_lds_bytes(nm_max);
  if (nm_max <= 16)
    return launch_solve_reg<64, 8, 2, 2>(tasks, n, lds, items, qbuf, tbuf,
                                         qparts, diag_only, ainv_out, s, dn,
                                       grid);
  // above order 16: blocked elimination in LDS with panels of 4 columns
  // (measured faster than 2, 6, 8 or 16), four waves per block up to order
  // 48, eight above, two rows per lane for sub-blocks above 64 rows
  if (nm_max <= 48)
    return launch_solve_blk<256, 1, 4>(tasks, n, lds_blk, items, qbuf, tbuf,
                                       qparts, diag_only, ainv_out, s, dn,
                                       grid);
  if (rows_max <= 32)  // m = 0 blocks up to order 64
    return launch_solve_blk<256, 1, 4>(tasks, n, lds_blk, items, qbuf, tbuf,
                                       qparts, diag_only, ainv_out, s, dn,
                                       grid);
  if (rows_max <= 64)
    return launch_solve_blk<512, 1, 4>(tasks, n, lds_blk, items, qbuf, tbuf,
                                       qparts, diag_only, ainv_out, s, dn,
                                       grid);
  return launch_solve_blk<512, 2, 4>(tasks, n, lds_blk, items, qbuf, tbuf,
                                     qparts, diag_only, ainv_out, s, dn,
                                     grid);
}

hipError_t launch_converge(const int32_t *active, int n_active,
                           const int32_t *win_off, const int32_t *win_items,
                           const Item *items, const double *qparts, double tol,
                           int last_order, ConvState *state, hipStream_t s) {
  if (n_active <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_converge, dim3((n_active + 127) / 128), dim3(128), 0, s,
                     active, n_active, win_off, win_items, items, qparts, tol,
                     last_order, state);
  return hipGetLastError();
}

hipError_t launch_full_sums(int n_grains, const int32_t *off,
                            const int32_t *gi, const double *qparts,
                            double *qsca, double *qext, hipStream_t s) {
  if (n_grains <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_full_sums, dim3((n_grains + 127) / 128), dim3(128), 0,
                     s, n_grains, off, gi, qparts, qsca, qext);
  return hipGetLastError();
}

hipError_t launch_write_orders(int n_grains, const ConvState *state,
                               uint32_t *nmax_out, int32_t *status_out,
                               hipStream_t s) {
  if (n_grains <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_write_orders, dim3((n_grains + 127) / 128), dim3(128),
                     0, s, n_grains, state, nmax_out, status_out);
  return hipGetLastError();
}

}  // namespace ctm
