// ctm_tangent.hip -- forward-mode derivatives of the batch's T-matrices with
// respect to (R_V, d, Re m, Im m) at frozen orders.
//
// The reference has no derivatives; these differentiate its assembly
// (TMatrixResource.hpp:623-705, 993-1097) and T = -RgQ Q^-1 exactly:
//   tables   d/dp of the node scalars and of j_n, [z j_n]'/z, y_n, [z y_n]'/z
//            at z = k r(theta) R_V and z = m k r(theta) R_V, from
//            j_n' = [z j_n]'/z - j_n / z and, by the Riccati-Bessel equation,
//            ([z j_n]'/z)' = (n (n + 1) / z^2 - 1) j_n - ([z j_n]'/z) / z;
//   assembly dQ = sum over nodes of dL R + L dR (dual-number staging of the
//            same row and column factors), plus the explicit k^2 dm term;
//   solve    dT = -(dRgQ + T dQ) Q^-1 with the stored Q^-1.
// Observables that are linear in T follow by feeding dT to the observable
// kernels.
#include "ctm_tangent.hpp"

namespace ctm {

namespace {

// value and derivative
struct dv {
  double v, d;
};
struct dz {
  cdbl v, d;
};
__device__ inline dv mkd(double v, double d) { return dv{v, d}; }
__device__ inline dv dmul(dv a, dv b) { return dv{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline dv dmulc(dv a, double c) { return dv{a.v * c, a.d * c}; }
__device__ inline dv dadd(dv a, dv b) { return dv{a.v + b.v, a.d + b.d}; }
__device__ inline dv dneg(dv a) { return dv{-a.v, -a.d}; }
__device__ inline dz zmk(dv re, dv im) {
  return dz{cmk(re.v, im.v), cmk(re.d, im.d)};
}
__device__ inline dz zmul(dz a, dz b) {
  return dz{cmul(a.v, b.v), cadd(cmul(a.d, b.v), cmul(a.v, b.d))};
}
__device__ inline dz zscale(dz a, dv f) {
  return dz{cscale(a.v, f.v), cadd(cscale(a.d, f.v), cscale(a.v, f.d))};
}
__device__ inline dz zscalec(dz a, double f) {
  return dz{cscale(a.v, f), cscale(a.d, f)};
}
__device__ inline dz zadd(dz a, dz b) { return dz{cadd(a.v, b.v), cadd(a.d, b.d)}; }

}  // namespace

// ---------------------------------------------------------------------------
// Tangent tables: one block per (set, direction), one thread per node.  A
// set's tangent tables hold [7][G] node-scalar tangents and [8][N][G]
// tangents of j, [z j]'/z, y, [z y]'/z at kr and re/im of j, [z j]'/z at
// m kr; they sit at twice the set's offset in the base tables (7 G + 8 N G
// is at most twice the base set's 11 G + 4 N G), direction dir at
// dir * 2 * table_size.
constexpr int kTangentScalars = 7;

__device__ __forceinline__ int64_t tangent_set_off(int dir,
                                                   int64_t table_size,
                                                   int64_t off) {
  return (int64_t)dir * 2 * table_size + 2 * off;
}

__global__ void k_tables_tangent(const TableSet *sets, const GrainParams gp,
                                 const double *gl, const double *tables,
                                 int64_t table_size, double *ttables) {
  const TableSet s = sets[blockIdx.x];
  const int dir = blockIdx.y;
  const int N = s.N, G = s.G;
  const double rv = gp.radius[s.grain];
  const double d = gp.axis_ratio[s.grain];
  const double lam = gp.wavelength[s.grain];
  const cdbl mr = cmk(gp.mr_re[s.grain], gp.mr_im[s.grain]);
  const double *cth = gl + s.gl_off;
  const double *wgt = cth + G;
  const double *S = tables + s.off;
  const double *B = S + (int64_t)kNodeScalars * G;
  double *dS = ttables + tangent_set_off(dir, table_size, s.off);
  double *dB = dS + (int64_t)kTangentScalars * G;
  const int64_t NG = (int64_t)N * G;
  const double k = 2. * M_PI / lam;
  const double d2 = d * d;
  for (int ig = threadIdx.x; ig < G; ig += blockDim.x) {
    const double c = cth[ig];
    const double w = wgt[ig];
    const double s2 = (1. - c) * (1. + c);
    const double sn = sqrt(s2);
    const double r2a2 = 1. / (s2 + d2 * (c * c));
    const double r2 = cbrt(d) * cbrt(d) * r2a2;
    // d ln(r^2) / dd for r^2 = d^(2/3) / (sin^2 + d^2 cos^2)
    const double dlnr2 = 2. / (3. * d) - 2. * d * c * c * r2a2;
    const double kr = k * sqrt(r2) * rv;
    const cdbl zm = cscale(mr, kr);
    double dzr = 0.;
    cdbl dzm = cmk(0., 0.);
    double dwr2 = 0., ddr = 0.;
    if (dir == 0) {
      dzr = kr / rv;
      dzm = cscale(mr, dzr);
    } else if (dir == 1) {
      dzr = 0.5 * kr * dlnr2;
      dzm = cscale(mr, dzr);
      dwr2 = w * r2 * dlnr2;
      ddr = c * sn * ((d2 - 1.) * (-2. * d * c * c) * r2a2 * r2a2 +
                      r2a2 * 2. * d);
    } else if (dir == 2) {
      dzm = cmk(kr, 0.);
    } else {
      dzm = cmk(0., kr);
    }
    dS[0 * G + ig] = dwr2;
    dS[1 * G + ig] = ddr;
    dS[2 * G + ig] = 0.;
    dS[3 * G + ig] = 0.;
    dS[4 * G + ig] = -dzr / (kr * kr);
    const cdbl dinv = cscale(cmul(dzm, crcp(cmul(zm, zm))), -1.);
    dS[5 * G + ig] = dinv.x;
    dS[6 * G + ig] = dinv.y;
    const double zi = 1. / kr;
    const cdbl zmi = crcp(zm);
    for (int i = 0; i < N; ++i) {
      const int n = i + 1;
      const double nn = n * (n + 1.);
      const int64_t o = (int64_t)i * G + ig;
      // real argument: j and y
      double j, dj, y, dy;
      bessel_kr(S, B, NG, G, n, ig, j, dj, y, dy);
      const double q = nn * zi * zi - 1.;
      dB[o] = (dj - j * zi) * dzr;
      dB[NG + o] = (q * j - dj * zi) * dzr;
      dB[2 * NG + o] = (dy - y * zi) * dzr;
      dB[3 * NG + o] = (q * y - dy * zi) * dzr;
      // complex argument: j(m kr)
      cdbl jm, djm;
      bessel_mkr(S, B, NG, G, n, ig, jm, djm);
      const cdbl qm = csub(cscale(cmul(zmi, zmi), nn), cmk(1., 0.));
      const cdbl tj = cmul(csub(djm, cmul(jm, zmi)), dzm);
      const cdbl tdj = cmul(csub(cmul(qm, jm), cmul(djm, zmi)), dzm);
      dB[4 * NG + o] = tj.x;
      dB[5 * NG + o] = tj.y;
      dB[6 * NG + o] = tdj.x;
      dB[7 * NG + o] = tdj.y;
    }
  }
}

// ---------------------------------------------------------------------------
// Tangent assembly: pair tiles of 16 rows and kTangentTileCols columns
// (twice that for m = 0), four waves split by pair parity as in the forward
// kernel, each thread one row and two columns (cc and cc + half the tile
// width) so a row's staged factors serve two pairs.  Rows and columns stage
// value and derivative of every factor; the accumulators hold the
// derivatives of J_a, J_b, RgJ_a, RgJ_b.  NEEDV (the refractive-index
// directions, blockIdx.y + 2) also accumulates their values for the
// d(k^2 m) term; along R_V and d that term vanishes and the value products
// are skipped.
constexpr int kTT = 16;
constexpr int kTCh = 4;
constexpr int kTLS = 34;
constexpr int kTRS = 10;

template <bool NEEDV>
__global__ __launch_bounds__(256) void k_assemble_tangent(
    const int32_t *order, const int32_t *tile_prefix, int n_order,
    const Item *items, const TableSet *sets, const WigKey *wkeys,
    const GrainParams gp, const double *tables, const double *ttables,
    int64_t table_size, const double *wig, cdbl *dq, int64_t q_size) {
  constexpr int kTC = kTangentTileCols;
  __shared__ double Lv[kTCh][kTT][kTLS], Ld[kTCh][kTT][kTLS];
  __shared__ double Rv[kTCh][2 * kTC][kTRS], Rd[kTCh][2 * kTC][kTRS];

  const int dir = blockIdx.y + (NEEDV ? 2 : 0);
  const int bid = blockIdx.x;
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
  const int cw = m0 ? 2 * kTC : kTC;  // tile columns
  const int ntc = (nm + cw - 1) / cw;
  const int lt = bid - tile_prefix[lo];
  const int tr = lt / ntc, tc = lt - (lt / ntc) * ntc;
  const double *S = tables + ts.off;
  const double *B = S + (int64_t)kNodeScalars * G;
  const double *dS = ttables + tangent_set_off(dir, table_size, ts.off);
  const double *dB = dS + (int64_t)kTangentScalars * G;
  const int64_t NG = (int64_t)N * G;
  const double *Wd = wig + wk.off;
  const double *Wdd = Wd + (int64_t)nm * G;

  const int t = threadIdx.x;
  const int wv = t >> 6, lane = t & 63;
  const int rr = 2 * (lane >> 3) + (wv & 1);
  // first column of the thread; the second is cc + cw / 2 (same parity)
  const int cc = m0 ? kTT * (wv >> 1) + 2 * (lane & 7) + (wv & 1)
                    : 2 * (lane & 7) + (wv >> 1);
  const int half = cw / 2;
  const int p1 = tr * kTT + rr;
  const bool odd = !m0 && (((wv & 1) ^ (wv >> 1)) != 0);
  const int nstage = kTT * kTCh + cw * kTCh;
  const double mm = (double)m;

  dz Ja[2], Jb[2], RJa[2], RJb[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    Ja[c] = dz{cmk(0., 0.), cmk(0., 0.)};
    Jb[c] = Ja[c];
    RJa[c] = Ja[c];
    RJb[c] = Ja[c];
  }

  for (int ig0 = 0; ig0 < G; ig0 += kTCh) {
    __syncthreads();
    for (int e = t; e < nstage; e += 256) {
      if (e < kTT * kTCh) {
        const int r = e / kTCh, kk = e % kTCh;
        const int ig = ig0 + kk;
        const int p = tr * kTT + r;
        double *L = Lv[kk][r];
        double *LD = Ld[kk][r];
        if (ig < G && p < nm) {
          const int n1 = nmin + p;
          const double d1 = Wd[(int64_t)p * G + ig];
          const double dd1 = Wdd[(int64_t)p * G + ig];
          const int64_t bo = (int64_t)(n1 - 1) * G + ig;
          double bj, bdj, by, bdy;
          bessel_kr(S, B, NG, G, n1, ig, bj, bdj, by, bdy);
          const dv j1 = mkd(bj, dB[bo]);
          const dv dj1 = mkd(bdj, dB[NG + bo]);
          const dv y1 = mkd(by, dB[2 * NG + bo]);
          const dv dy1 = mkd(bdy, dB[3 * NG + bo]);
          const dv wr2 = mkd(S[ig], dS[ig]);
          const dv dr = mkd(S[G + ig], dS[G + ig]);
          const double s1 = S[2 * G + ig], s2 = S[3 * G + ig];
          const dv krinv = mkd(S[4 * G + ig], dS[4 * G + ig]);
          const dz imkr = zmk(mkd(S[5 * G + ig], dS[5 * G + ig]),
                              mkd(S[6 * G + ig], dS[6 * G + ig]));
          const double nn1 = n1 * (n1 + 1.);
          const double a1 = sqrt((2. * n1 + 1.) / nn1);
          const dz h1 = zmk(j1, y1), dh1 = zmk(dj1, dy1);
          const double m2s2 = mm * mm * s2;
          const dv q = dmulc(dmul(dr, krinv), nn1 * d1);
          const dv fA1 = dmulc(wr2, m2s2 * d1 * a1);
          const dv fA2 = dmulc(wr2, dd1 * a1);
          const dv fA3 = dmulc(dmul(wr2, dr), dd1 * a1);
          const dv fA5 = dmulc(wr2, -a1);
          const dz A1 = zscale(h1, fA1);
          const dz A2 = zscale(h1, fA2);
          const dz A3 = zmul(imkr, zscale(h1, fA3));
          const dz A4 = zscale(dh1, dneg(fA1));
          const dz A5 = zscale(zadd(zscalec(dh1, dd1), zscale(h1, q)), fA5);
          const dv L10 = dmul(fA1, j1), L11 = dmul(fA2, j1);
          const dz A3r = zscale(imkr, dmul(fA3, j1));
          const dv L14 = dneg(dmul(fA1, dj1));
          const dv L15 = dmul(fA5, dadd(dmulc(dj1, dd1), dmul(q, j1)));
          const dv dsi = dmulc(wr2, s1 * mm);
          const dv fB5 = dmulc(dsi, d1 * a1);
          const dv fB4 = dmulc(dsi, dd1 * a1);
          const dv fB1 = dmulc(dsi, a1);
          const dv fB3 = dmulc(dmul(dsi, dr), d1 * a1);
          const dz B5 = zscale(h1, fB5);
          const dz B4 = zscale(h1, fB4);
          const dz B2 = zscale(dh1, fB5);
          const dz B1 = zscale(zadd(zscalec(dh1, dd1), zscale(h1, q)), fB1);
          const dz B3 = zmul(imkr, zscale(dh1, fB3));
          const dv L26 = dmul(fB5, j1), L27 = dmul(fB4, j1);
          const dv L28 = dmul(fB5, dj1);
          const dv L29 = dmul(fB1, dadd(dmulc(dj1, dd1), dmul(q, j1)));
          const dz B3r = zscale(imkr, dmul(fB3, dj1));
          const dz zc[10] = {A1, A2, A3, A4, A5, B5, B4, B2, B1, B3};
          const int zpos[10] = {0, 2, 4, 6, 8, 16, 18, 20, 22, 24};
#pragma unroll
          for (int q2 = 0; q2 < 10; ++q2) {
            L[zpos[q2]] = zc[q2].v.x; L[zpos[q2] + 1] = zc[q2].v.y;
            LD[zpos[q2]] = zc[q2].d.x; LD[zpos[q2] + 1] = zc[q2].d.y;
          }
          const dv rc[8] = {L10, L11, L14, L15, L26, L27, L28, L29};
          const int rpos[8] = {10, 11, 14, 15, 26, 27, 28, 29};
#pragma unroll
          for (int q2 = 0; q2 < 8; ++q2) {
            L[rpos[q2]] = rc[q2].v;
            LD[rpos[q2]] = rc[q2].d;
          }
          L[12] = A3r.v.x; L[13] = A3r.v.y; LD[12] = A3r.d.x; LD[13] = A3r.d.y;
          L[30] = B3r.v.x; L[31] = B3r.v.y; LD[30] = B3r.d.x; LD[31] = B3r.d.y;
        } else {
#pragma unroll
          for (int i = 0; i < 32; ++i) {
            L[i] = 0.;
            LD[i] = 0.;
          }
        }
      } else {
        const int tt = e - kTT * kTCh;
        const int c = tt / kTCh, kk = tt % kTCh;
        const int ig = ig0 + kk;
        const int p = tc * cw + c;
        double *R = Rv[kk][c];
        double *RD = Rd[kk][c];
        if (ig < G && p < nm) {
          const int n2 = nmin + p;
          const double d2 = Wd[(int64_t)p * G + ig];
          const double dd2 = Wdd[(int64_t)p * G + ig];
          const int64_t bo = (int64_t)(n2 - 1) * G + ig;
          cdbl bjm, bdjm;
          bessel_mkr(S, B, NG, G, n2, ig, bjm, bdjm);
          const dz jm = zmk(mkd(bjm.x, dB[4 * NG + bo]),
                            mkd(bjm.y, dB[5 * NG + bo]));
          const dz djm = zmk(mkd(bdjm.x, dB[6 * NG + bo]),
                             mkd(bdjm.y, dB[7 * NG + bo]));
          const double nn2 = n2 * (n2 + 1.);
          const double a2 = sqrt((2. * n2 + 1.) / nn2);
          const dz f[5] = {zscalec(djm, a2 * d2), zscalec(djm, a2 * dd2),
                           zscalec(jm, a2 * nn2 * d2), zscalec(jm, a2 * d2),
                           zscalec(jm, a2 * dd2)};
#pragma unroll
          for (int q2 = 0; q2 < 5; ++q2) {
            R[2 * q2] = f[q2].v.x; R[2 * q2 + 1] = f[q2].v.y;
            RD[2 * q2] = f[q2].d.x; RD[2 * q2 + 1] = f[q2].d.y;
          }
        } else {
#pragma unroll
          for (int i = 0; i < 10; ++i) {
            R[i] = 0.;
            RD[i] = 0.;
          }
        }
      }
    }
    __syncthreads();
    const int kmax = (G - ig0) < kTCh ? (G - ig0) : kTCh;
    for (int kk = 0; kk < kmax; ++kk) {
      const double *L = Lv[kk][rr] + (odd ? 16 : 0);
      const double *LD = Ld[kk][rr] + (odd ? 16 : 0);
      // A += l r (dual): d += l.v r.d + l.d r.v, and v += l.v r.v if NEEDV.
      // Along Re m and Im m (NEEDV) a row factor has a derivative only
      // where it carries 1/(m kr) (A3, A3r, B3, B3r: the factors of R3);
      // the other l.d are zero and are not read (LD = 0 skips them).
      auto zz = [&](dz &A, int i, const dz &r, bool ld_used) {
        const cdbl lv = cmk(L[i], L[i + 1]);
        if (NEEDV) cfma(A.v, lv, r.v);
        cfma(A.d, lv, r.d);
        if (!NEEDV || ld_used) cfma(A.d, cmk(LD[i], LD[i + 1]), r.v);
      };
      auto rz = [&](dz &A, int i, const dz &r) {
        const double lv = L[i];
        if (NEEDV) rfma(A.v, lv, r.v);
        rfma(A.d, lv, r.d);
        if (!NEEDV) rfma(A.d, LD[i], r.v);
      };
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double *R = Rv[kk][cc + c * half];
        const double *RD = Rd[kk][cc + c * half];
        dz r[5];
#pragma unroll
        for (int q2 = 0; q2 < 5; ++q2)
          r[q2] = dz{cmk(R[2 * q2], R[2 * q2 + 1]),
                     cmk(RD[2 * q2], RD[2 * q2 + 1])};
        if (!odd) {
          zz(Ja[c], 0, r[0], false);
          zz(Ja[c], 2, r[1], false);
          zz(Ja[c], 4, r[2], true);    // A3
          zz(Jb[c], 6, r[3], false);
          zz(Jb[c], 8, r[4], false);
          rz(RJa[c], 10, r[0]);
          rz(RJa[c], 11, r[1]);
          zz(RJa[c], 12, r[2], true);  // A3r
          rz(RJb[c], 14, r[3]);
          rz(RJb[c], 15, r[4]);
        } else {
          zz(Ja[c], 0, r[4], false);
          zz(Ja[c], 2, r[3], false);
          zz(Jb[c], 4, r[1], false);
          zz(Jb[c], 6, r[0], false);
          zz(Jb[c], 8, r[2], true);    // B3
          rz(RJa[c], 10, r[4]);
          rz(RJa[c], 11, r[3]);
          rz(RJb[c], 12, r[1]);
          rz(RJb[c], 13, r[0]);
          zz(RJb[c], 14, r[2], true);  // B3r
        }
      }
    }
  }
  if (p1 >= nm) return;
  auto timesi = [](dz a) {
    return dz{cmk(-a.v.y, a.v.x), cmk(-a.d.y, a.d.x)};
  };
  const double lam = gp.wavelength[it.grain];
  const double k = 2. * M_PI / lam;
  const double k2 = k * k;
  const cdbl k2mr = cscale(cmk(gp.mr_re[it.grain], gp.mr_im[it.grain]), k2);
  // d(k^2 m) / dp: k^2 along Re m, i k^2 along Im m
  const cdbl dk2m = dir == 2 ? cmk(k2, 0.) : (dir == 3 ? cmk(0., k2) : cmk(0., 0.));
  auto dq_of = [&](dz a, dz b) {
    return cadd(cadd(cmul(dk2m, a.v), cmul(k2mr, a.d)), cscale(b.d, k2));
  };
  const int n1 = nmin + p1;
  const int cx = n1 & 1;
  const int64_t nm2 = (int64_t)nm * nm;
  cdbl *Q = dq + (int64_t)dir * q_size + it.q_off;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int p2 = tc * cw + cc + c * half;
    if (p2 >= nm) continue;
    dz ja = Ja[c], jb = Jb[c], rja = RJa[c], rjb = RJb[c];
    if (odd) {
      ja = timesi(ja);
      jb = timesi(jb);
      rja = timesi(rja);
      rjb = timesi(rjb);
    }
    const int64_t e = (int64_t)p1 * nm + p2;
    Q[cx * nm2 + e] = dq_of(ja, jb);
    Q[(1 - cx) * nm2 + e] = dq_of(jb, ja);
    Q[(2 + cx) * nm2 + e] = dq_of(rja, rjb);
    Q[(2 + 1 - cx) * nm2 + e] = dq_of(rjb, rja);
  }
}

// ---------------------------------------------------------------------------
// Tangent solve, in two launches so that M is complete before it is used.
// Block (task, dir); task = 2 item + class.
__global__ __launch_bounds__(256) void k_tangent_m(const Item *items,
                                                   const cdbl *tbuf, cdbl *dq,
                                                   int64_t q_size) {
  // all four directions per thread: T(r, k) is loaded once for them
  const int task = blockIdx.x;
  const Item it = items[task >> 1];
  const int cls = task & 1;
  const int nm = it.N + 1 - (it.m > 0 ? it.m : 1);
  const int64_t nm2 = (int64_t)nm * nm;
  const cdbl *T = tbuf + it.t_off + cls * nm2;
  // per direction: dQ, and dRgQ replaced by M = dRgQ + T dQ
  const cdbl *dQ[kTangentDirs];
  cdbl *M[kTangentDirs];
#pragma unroll
  for (int d = 0; d < kTangentDirs; ++d) {
    cdbl *Q = dq + (int64_t)d * q_size + it.q_off;
    dQ[d] = Q + cls * nm2;
    M[d] = Q + (2 + cls) * nm2;
  }
  // m = 0: only indices of equal parity couple (TMatrixResource.hpp:622);
  // the assembly leaves the other entries unwritten, so they are skipped
  // here and zero in M
  const int step = it.m == 0 ? 2 : 1;
  for (int64_t e = threadIdx.x; e < nm2; e += blockDim.x) {
    const int r = (int)(e / nm), c = (int)(e - (int64_t)(e / nm) * nm);
    if (step == 2 && ((r ^ c) & 1)) {
#pragma unroll
      for (int d = 0; d < kTangentDirs; ++d) M[d][e] = cmk(0., 0.);
      continue;
    }
    cdbl acc[kTangentDirs];
#pragma unroll
    for (int d = 0; d < kTangentDirs; ++d) acc[d] = M[d][e];
    for (int k = step == 2 ? (r & 1) : 0; k < nm; k += step) {
      const cdbl tv = T[(int64_t)r * nm + k];
#pragma unroll
      for (int d = 0; d < kTangentDirs; ++d)
        cfma(acc[d], tv, dQ[d][(int64_t)k * nm + c]);
    }
#pragma unroll
    for (int d = 0; d < kTangentDirs; ++d) M[d][e] = acc[d];
  }
}
__global__ __launch_bounds__(256) void k_tangent_t(const Item *items,
                                                   const cdbl *tbuf,
                                                   const cdbl *ainv,
                                                   const cdbl *dq,
                                                   int64_t q_size, cdbl *dt,
                                                   int64_t t_size,
                                                   double *tparts) {
  // all four directions per thread: Q^-1(k, c) is loaded once for them
  __shared__ double red[4][kTangentDirs][2];
  const int task = blockIdx.x;
  const Item it = items[task >> 1];
  const int cls = task & 1;
  const int nm = it.N + 1 - (it.m > 0 ? it.m : 1);
  const int64_t nm2 = (int64_t)nm * nm;
  const cdbl *T = tbuf + it.t_off + cls * nm2;
  const cdbl *Ai = ainv + it.t_off + cls * nm2;
  const cdbl *M[kTangentDirs];
  cdbl *D[kTangentDirs];
#pragma unroll
  for (int d = 0; d < kTangentDirs; ++d) {
    M[d] = dq + (int64_t)d * q_size + it.q_off + (2 + cls) * nm2;
    D[d] = dt + (int64_t)d * t_size + it.t_off + cls * nm2;
  }
  double s_sca[kTangentDirs], s_ext[kTangentDirs];
#pragma unroll
  for (int d = 0; d < kTangentDirs; ++d) s_sca[d] = s_ext[d] = 0.;
  const int step = it.m == 0 ? 2 : 1;
  for (int64_t e = threadIdx.x; e < nm2; e += blockDim.x) {
    const int r = (int)(e / nm), c = (int)(e - (int64_t)(e / nm) * nm);
    cdbl acc[kTangentDirs];
#pragma unroll
    for (int d = 0; d < kTangentDirs; ++d) acc[d] = cmk(0., 0.);
    if (!(step == 2 && ((r ^ c) & 1)))
      for (int k = step == 2 ? (r & 1) : 0; k < nm; k += step) {
        const cdbl av = Ai[(int64_t)k * nm + c];
#pragma unroll
        for (int d = 0; d < kTangentDirs; ++d)
          cfma(acc[d], M[d][(int64_t)r * nm + k], av);
      }
    const cdbl tv = T[e];
#pragma unroll
    for (int d = 0; d < kTangentDirs; ++d) {
      const cdbl v = cmk(-acc[d].x, -acc[d].y);
      D[d][e] = v;
      s_sca[d] += 2. * (tv.x * v.x + tv.y * v.y);
      if (r == c) s_ext[d] += v.x;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < kTangentDirs; ++d) {
    for (int o = 32; o > 0; o >>= 1) {
      s_sca[d] += __shfl_xor(s_sca[d], o);
      s_ext[d] += __shfl_xor(s_ext[d], o);
    }
    if (lane == 0) {
      red[wv][d][0] = s_sca[d];
      red[wv][d][1] = s_ext[d];
    }
  }
  __syncthreads();
  if (threadIdx.x < kTangentDirs) {
    const int d = threadIdx.x;
    double a = 0., b = 0.;
    for (int w = 0; w < 4; ++w) {
      a += red[w][d][0];
      b += red[w][d][1];
    }
    tparts[((int64_t)task * kTangentDirs + d) * 2] = a;
    tparts[((int64_t)task * kTangentDirs + d) * 2 + 1] = b;
  }
}

__global__ void k_tangent_sums(int n_grains, const int32_t *gi_off,
                               const double *tparts, double *dqsca,
                               double *dqext) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_grains * kTangentDirs) return;
  const int g = idx / kTangentDirs, dir = idx % kTangentDirs;
  double s = 0., e = 0.;
  for (int i = gi_off[g]; i < gi_off[g + 1]; ++i) {
    const double f = i == gi_off[g] ? 1. : 2.;
    for (int cls = 0; cls < 2; ++cls) {
      const double *p = tparts + ((int64_t)(2 * i + cls) * kTangentDirs + dir) * 2;
      s += f * p[0];
      e += f * p[1];
    }
  }
  dqsca[idx] = s;
  dqext[idx] = e;
}

// Derivatives of the observables that are quadratic in T (Z, and Q_abs
// with the scattered light): S is linear in T and every observable is at
// most quadratic in S, so O(T + s dT) - O(T - s dT) = 2 s dO exactly (the
// quadratic terms cancel) for any s.  Per grain, s = max |T| / max |dT|
// keeps both evaluations of the size of O(T), so the difference loses no
// more than a few ulps of it.
__global__ __launch_bounds__(256) void k_pm_scale(const int32_t *gi_off,
                                                  const Item *items,
                                                  const cdbl *tbuf,
                                                  const cdbl *dt,
                                                  double *scale) {
  const int g = blockIdx.x;
  double mt = 0., md = 0.;
  for (int i = gi_off[g]; i < gi_off[g + 1]; ++i) {
    const Item it = items[i];
    const int64_t nm = it.N + 1 - (it.m > 0 ? it.m : 1);
    for (int64_t e = threadIdx.x; e < 2 * nm * nm; e += 256) {
      const cdbl a = tbuf[it.t_off + e], b = dt[it.t_off + e];
      mt = fmax(mt, fmax(fabs(a.x), fabs(a.y)));
      md = fmax(md, fmax(fabs(b.x), fabs(b.y)));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mt = fmax(mt, __shfl_xor(mt, o));
    md = fmax(md, __shfl_xor(md, o));
  }
  __shared__ double red[4][2];
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6][0] = mt;
    red[threadIdx.x >> 6][1] = md;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      mt = fmax(mt, red[w][0]);
      md = fmax(md, red[w][1]);
    }
    mt = fmax(mt, red[0][0]);
    md = fmax(md, red[0][1]);
    scale[g] = (mt > 0. && md > 0.) ? mt / md : 1.;
  }
}

// out = T + sign s_g dT over the T areas of the grains' items (one
// workgroup per item).
__global__ __launch_bounds__(256) void k_pm_form(const Item *items,
                                                 const cdbl *tbuf,
                                                 const cdbl *dt,
                                                 const double *scale,
                                                 double sign, cdbl *out) {
  const Item it = items[blockIdx.x];
  const int64_t nm = it.N + 1 - (it.m > 0 ? it.m : 1);
  const double s = sign * scale[it.grain];
  for (int64_t e = threadIdx.x; e < 2 * nm * nm; e += 256) {
    const cdbl a = tbuf[it.t_off + e], b = dt[it.t_off + e];
    out[it.t_off + e] = cmk(fma(s, b.x, a.x), fma(s, b.y, a.y));
  }
}

// out = (plus - minus) / (2 s_g) for an array of n_grains rows of `width`
// values.
__global__ void k_pm_combine(int64_t n, int64_t width, const double *plus,
                             const double *minus, const double *scale,
                             double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (plus[i] - minus[i]) / (2. * scale[i / width]);
}

hipError_t launch_pm_scale(int n_grains, const int32_t *gi_off,
                           const Item *items, const cdbl *tbuf, const cdbl *dt,
                           double *scale, hipStream_t s) {
  if (n_grains <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pm_scale, dim3(n_grains), dim3(256), 0, s, gi_off,
                     items, tbuf, dt, scale);
  return hipGetLastError();
}

hipError_t launch_pm_form(int n_items, const Item *items, const cdbl *tbuf,
                          const cdbl *dt, const double *scale, double sign,
                          cdbl *out, hipStream_t s) {
  if (n_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pm_form, dim3(n_items), dim3(256), 0, s, items, tbuf,
                     dt, scale, sign, out);
  return hipGetLastError();
}

hipError_t launch_pm_combine(int64_t n, int64_t width, const double *plus,
                             const double *minus, const double *scale,
                             double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pm_combine, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, n, width, plus, minus, scale, out);
  return hipGetLastError();
}

hipError_t launch_tables_tangent(const TableSet *sets, int n_sets,
                                 const GrainParams &gp, const double *gl,
                                 const double *tables, int64_t table_size,
                                 double *ttables, hipStream_t s) {
  if (n_sets <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tables_tangent, dim3(n_sets, kTangentDirs), dim3(64),
                     0, s, sets, gp, gl, tables, table_size, ttables);
  return hipGetLastError();
}

hipError_t launch_assemble_tangent(const int32_t *order,
                                   const int32_t *tile_prefix, int n_order,
                                   int n_tiles, const Item *items,
                                   const TableSet *sets, const WigKey *wkeys,
                                   const GrainParams &gp, const double *tables,
                                   const double *ttables, int64_t table_size,
                                   const double *wig, cdbl *dq, int64_t q_size,
                                   hipStream_t s) {
  if (n_tiles <= 0) return hipSuccess;
  // R_V and d (no value products), then Re m and Im m
  hipLaunchKernelGGL(k_assemble_tangent<false>, dim3(n_tiles, 2), dim3(256),
                     0, s, order, tile_prefix, n_order, items, sets, wkeys, gp,
                     tables, ttables, table_size, wig, dq, q_size);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_assemble_tangent<true>, dim3(n_tiles, 2), dim3(256),
                     0, s, order, tile_prefix, n_order, items, sets, wkeys, gp,
                     tables, ttables, table_size, wig, dq, q_size);
  return hipGetLastError();
}

hipError_t launch_tangent_solve(int n_tasks, const Item *items,
                                const cdbl *tbuf, const cdbl *ainv, cdbl *dq,
                                int64_t q_size, cdbl *dt, int64_t t_size,
                                double *tparts, hipStream_t s) {
  if (n_tasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tangent_m, dim3(n_tasks), dim3(256), 0, s, items,
                     tbuf, dq, q_size);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_tangent_t, dim3(n_tasks), dim3(256), 0, s, items,
                     tbuf, ainv, dq, q_size, dt, t_size, tparts);
  return hipGetLastError();
}

hipError_t launch_tangent_sums(int n_grains, const int32_t *gi_off,
                               const double *tparts, double *dqsca,
                               double *dqext, hipStream_t s) {
  const int n = n_grains * kTangentDirs;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tangent_sums, dim3((n + 127) / 128), dim3(128), 0, s,
                     n_grains, gi_off, tparts, dqsca, dqext);
  return hipGetLastError();
}

}  // namespace ctm
