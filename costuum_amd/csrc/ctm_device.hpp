// ctm_device.hpp -- shared device-side types of the gfx950 T-matrix path.
//
// Data layout in HBM (all offsets in doubles unless stated):
//   GL cache     per distinct G: cos(theta)[G], weight[G] of the first G
//                nodes (theta > pi/2) of the 2G-point Gauss-Legendre rule.
//   Wigner cache per distinct (N, m): d^n_{0m}[nm][G], d'[nm][G] for
//                n = max(m,1)..N at those nodes (grain-independent).
//   Table set    per (grain, order N): node scalars [11][G] (w r^2, dr/r,
//                1/sin, 1/sin^2, 1/kr, 1/(m kr) re, im, j_0(kr), y_0(kr),
//                j_0(m kr) re, im) followed by Bessel tables [4][N][G]
//                (j_n, y_n at kr; re/im of j_n at m kr; n = 1..N).  The
//                derivatives [z f_n]'/z = f_{n-1} - n f_n / z are formed by
//                the readers from neighbouring orders (bessel_kr,
//                bessel_mkr), which halves the tables' HBM traffic.
//   Q area       per work item (grain, N, m): four nm x nm complex blocks
//                Q_0, Q_1, RgQ_0, RgQ_1 (parity classes 0 and 1).
//   T area       per work item: T_0, T_1 (nm x nm complex each).
//
// Parity classes: for m > 0 the EBCM Q-matrix row (half h, order n) couples
// only to columns (h', n') with (n + h) = (n' + h') mod 2, so Q splits into
// two independent nm x nm blocks; class c holds the rows with
// (n + h) mod 2 == c, indexed by p = n - max(m, 1).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ctm {

struct cdbl {
  double x, y;
};

__host__ __device__ inline cdbl cmk(double a, double b) { return {a, b}; }
__host__ __device__ inline cdbl cadd(cdbl a, cdbl b) {
  return {a.x + b.x, a.y + b.y};
}
__host__ __device__ inline cdbl csub(cdbl a, cdbl b) {
  return {a.x - b.x, a.y - b.y};
}
__host__ __device__ inline cdbl cmul(cdbl a, cdbl b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline cdbl cscale(cdbl a, double s) {
  return {a.x * s, a.y * s};
}
__host__ __device__ inline cdbl crcp(cdbl a) {
  const double d = a.x * a.x + a.y * a.y;
  return {a.x / d, -a.y / d};
}
// acc += a * b (complex)
__host__ __device__ inline void cfma(cdbl &acc, cdbl a, cdbl b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc += s * b (real times complex)
__host__ __device__ inline void rfma(cdbl &acc, double s, cdbl b) {
  acc.x = fma(s, b.x, acc.x);
  acc.y = fma(s, b.y, acc.y);
}

// Single-precision complex (the FP32 elimination of the mixed path).
struct cflt {
  float x, y;
};

__host__ __device__ inline cflt cmkf(float a, float b) { return {a, b}; }
__host__ __device__ inline cflt cmul(cflt a, cflt b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline cflt crcp(cflt a) {
  const float d = a.x * a.x + a.y * a.y;
  return {a.x / d, -a.y / d};
}
__host__ __device__ inline void cfma(cflt &acc, cflt a, cflt b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
}

// Number of node scalars and Bessel arrays in a table set.
constexpr int kNodeScalars = 11;
constexpr int kBesselArrays = 4;

// j_n(kr), y_n(kr) and [z j_n]'/z, [z y_n]'/z at z = kr for order n >= 1 at
// node ig of a table set (S: node scalars, B: Bessel tables, NG = N G).
__device__ __forceinline__ void bessel_kr(const double *S, const double *B,
                                          int64_t NG, int G, int n, int ig,
                                          double &j, double &dj, double &y,
                                          double &dy) {
  const int64_t o = (int64_t)(n - 1) * G + ig;
  j = B[o];
  y = B[NG + o];
  const double jp = n > 1 ? B[o - G] : S[7 * G + ig];
  const double yp = n > 1 ? B[NG + o - G] : S[8 * G + ig];
  const double zinv = S[4 * G + ig];
  dj = jp - n * zinv * j;
  dy = yp - n * zinv * y;
}

// j_n(m kr) and [z j_n]'/z at z = m kr for order n >= 1.
__device__ __forceinline__ void bessel_mkr(const double *S, const double *B,
                                           int64_t NG, int G, int n, int ig,
                                           cdbl &jm, cdbl &djm) {
  const int64_t o = (int64_t)(n - 1) * G + ig;
  jm = cmk(B[2 * NG + o], B[3 * NG + o]);
  const cdbl jp = n > 1 ? cmk(B[2 * NG + o - G], B[3 * NG + o - G])
                        : cmk(S[9 * G + ig], S[10 * G + ig]);
  const cdbl zinv = cmk(S[5 * G + ig], S[6 * G + ig]);
  djm = csub(jp, cmul(cscale(zinv, (double)n), jm));
}

struct GlKey {
  int32_t G;
  int32_t pad;
  int64_t off;
};

struct WigKey {
  int32_t N, m, G, pad;
  int64_t gl_off;
  int64_t off;
};

struct TableSet {
  int32_t grain, N, G;
  int32_t node0;  // nodes of the sets before this one in its list
  int64_t gl_off;
  int64_t off;
};

// One (grain, order N, m) T-matrix block computation.
struct Item {
  int32_t grain, N, m, G;
  int32_t set;   // table set index
  int32_t wig;   // Wigner key index
  int64_t q_off; // complex offset of the Q area
  int64_t t_off; // complex offset of the T area
};

struct TileTask {
  int32_t item, tr, tc, pad;
};

// Per-(item, class) partial sums written by the solve kernel.
//   [0] sum (2n+1) |T(p,p)|^2   [1] sum (2n+1) Re T(p,p)   (m = 0 test)
//   [2] sum |T|^2 over the block [3] sum Re T(p,p)          (full-T sums)
constexpr int kQParts = 4;

// Per-grain convergence state (TMatrixResource.hpp:112-124,743-775).
struct ConvState {
  double old_s, old_e, dqs, dqe;
  int32_t first;     // 1 until the first order was evaluated
  int32_t done;      // 1 once converged or out of orders
  int32_t nmax;      // converged (or last) order
  int32_t item;      // phase-1 item holding the m = 0 block at nmax
  int32_t status;    // CTM_CONVERGED / CTM_NOT_CONVERGED
  int32_t pad[3];
};

// The reference's sequential convergence test (TMatrixResource.hpp:743-775)
// over the m = 0 items q0 .. q1 - 1 of one grain in ascending order (their
// indices win_items[q], or q itself when win_items is NULL), continuing from
// the grain's state st; stops at the first order that passes or at
// last_order.
__device__ inline void converge_items(ConvState &st, int q0, int q1,
                                      const int32_t *win_items,
                                      const Item *items,
                                      const double *qparts, double tol,
                                      int last_order) {
  for (int q = q0; q < q1; ++q) {
    const int ii = win_items ? win_items[q] : q;
    const Item it = items[ii];
    const double *p0 = qparts + (int64_t)(2 * ii) * kQParts;
    const double *p1 = p0 + kQParts;
    const double qs = p0[0] + p1[0];
    const double qe = p0[1] + p1[1];
    if (st.first) {
      st.first = 0;
      st.dqs = -1.;
      st.dqe = -1.;
    } else {
      st.dqs = fabs((st.old_s - qs) / qs);
      st.dqe = fabs((st.old_e - qe) / qe);
    }
    st.old_s = qs;
    st.old_e = qe;
    st.nmax = it.N;
    st.item = ii;
    if (st.dqs > 0. && st.dqs <= tol && st.dqe <= tol) {
      st.done = 1;
      st.status = 0;
      break;
    }
    if (it.N >= last_order) {
      st.done = 1;
      st.status = 1;
      break;
    }
  }
}

}  // namespace ctm
