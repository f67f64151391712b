// ctm_observables.hip -- gfx950 kernels for the observables of a batch of
// T-matrices: forward-scattering extinction, absorption (optionally minus the
// scattered light, by the reference's GL x GL quadrature over incoming
// directions) and the scattering matrix on get_table's (theta, phi) grid.
//
// Every observable is built from the S matrix of an (outgoing, incoming)
// direction pair (ExtinctionCoefficientTask.hpp:517-619 and its copies in
// AbsorptionCoefficientTask.hpp:612-715 and ScatteringMatrixTask.hpp:588-691):
//   S_k(o, i, dphi) = (1/k) sum_m g_m^k(dphi) X_m^k(o, i)
// where, with pi/tau of the outgoing (a) and incoming (b) direction and
// M = c_{nn' n} T_m,
//   X^0 = a^T M b,  X^3 = a'^T M b',  X^1 = a^T M b',  X^2 = -a'^T M b,
//   a = (pi, tau), a' = (tau, pi) over the two halves of the m block,
// g^0 = g^3 = 1 (m = 0) or 2 cos(m dphi), g^1 = g^2 = 0 (m = 0) or
// 2 sin(m dphi).  The reference evaluates S pair by pair in O(N^3); here the
// m-block products are shared by all pairs of a grain: K6 forms X per
// (grain, m) work item, K7 sums over m per (grain, direction).
#include "ctm_device.hpp"
#include "ctm_kernels.hpp"
#include "ctm_observables.hpp"

namespace ctm {

// pi/tau of d^n_{0m} at one direction (SpecialFunctions.hpp:398-495):
// pi = m d/sin, tau = d(d)/dtheta, for n = 1..nmax.
__device__ void pitau_dev(double cosx, double sinx, double sinx_inv, int nmax,
                          int m, double *pi, double *tau, int stride) {
  if (fabs(1. - fabs(cosx)) <= 1.e-10) {
    if (m == 1) {
      int sign = 1;
      for (int n = 1; n <= nmax; ++n) {
        double dn = 0.5 * sqrt(n * (n + 1.));
        if (cosx < 0.) dn *= sign;
        pi[(n - 1) * stride] = m * dn;
        if (cosx < 0.) dn = -dn;
        tau[(n - 1) * stride] = dn;
        sign = -sign;
      }
    } else {
      for (int n = 0; n < nmax; ++n) {
        pi[n * stride] = 0.;
        tau[n * stride] = 0.;
      }
    }
    return;
  }
  if (m == 0) {
    double dprev = 1., dcur = cosx;
    for (int n = 0; n < nmax; ++n) {
      const double np1 = n + 1., np2 = n + 2.;
      const double tnp3 = 2. * np1 + 1.;
      const double dnext = (tnp3 * cosx * dcur - np1 * dprev) / np2;
      const double deriv = sinx_inv * np1 * np2 * (dnext - dprev) / tnp3;
      pi[n * stride] = 0. * (dcur * sinx_inv);
      tau[n * stride] = deriv;
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
    for (int n = 0; n + 1 < m && n < nmax; ++n) {
      pi[n * stride] = 0.;
      tau[n * stride] = 0.;
    }
    for (int n = m - 1; n < nmax; ++n) {
      const double np1 = n + 1., np2 = n + 2.;
      const double tnp3 = 2. * np1 + 1.;
      const double s1 = sqrt(np1 * np1 - m2);
      const double s2 = sqrt(np2 * np2 - m2);
      const double dnext = (tnp3 * cosx * dcur - s1 * dprev) / s2;
      const double deriv =
          sinx_inv * (np1 * s2 * dnext - np2 * s1 * dprev) / tnp3;
      pi[n * stride] = m * (dcur * sinx_inv);
      tau[n * stride] = deriv;
      dprev = dcur;
      dcur = dnext;
    }
  }
}

// Grain-independent pi/tau table PT[m][angle][n][2] for m, n <= nmax_all.
__global__ void k_pitau(const double *ang, int n_ang, int nmax_all,
                        double *PT) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (nmax_all + 1) * n_ang) return;
  const int m = idx / n_ang, a = idx - m * n_ang;
  double *base = PT + ((int64_t)(m * n_ang + a) * nmax_all) * 2;
  pitau_dev(ang[3 * a], ang[3 * a + 1], ang[3 * a + 2], nmax_all, m, base,
            base + 1, 2);
}

// c_{nn n} = i^(nn - n - 1) a_nn a_n (NBasedResources.hpp:115-137).
__device__ inline cdbl cfac(int nn, int n) {
  const double a = sqrt((2. * nn + 1.) / (nn * (nn + 1.))) *
                   sqrt((2. * n + 1.) / (n * (n + 1.)));
  const int e = (((nn - n - 1) % 4) + 4) % 4;
  return e == 0 ? cmk(a, 0.) : e == 1 ? cmk(0., a) : e == 2 ? cmk(-a, 0.)
                                                           : cmk(0., -a);
}

// K6: X_m^k for every direction pair of one (grain, m) item.  The incoming
// directions are processed in chunks of CK: Y(h n; i) = sum_nn c_{nn n}
// T(h n; h' nn) b(h' nn; i) for b and b' (for the parity-class layout h' is
// fixed by the class), one thread per row (h n) with the chunk's
// accumulators in registers.  When the 2 nm rows leave threads idle,
// several chunks run side by side (G groups of 2 nm threads; G chunks per
// round).  Then one thread per direction pair of the round's incoming
// directions (pairs listed by incoming direction) contracts Y with the
// outgoing pi/tau for all four k.  pi/tau of the outgoing directions
// (stored [n][direction] so neighbouring pairs read neighbouring words),
// of the round's incoming directions and the n-factors a_n are staged in
// LDS.
constexpr int kObsXQ = 8;  // elements of a T row per batch of loads

// One workgroup per listed item.  dn (may be NULL): the item count is read
// from dn[0]; workgroups beyond it return at once.
template <int CK, int NT, int GCAP, bool PTG>
__global__ __launch_bounds__(NT) void k_obs_x(const ObsSpec spec,
                                               const int32_t *item_list,
                                               const Item *items,
                                               const cdbl *tbuf,
                                               const cdbl *abuf,
                                               const int64_t *aoff,
                                               const double *PT,
                                               int nmax_all, int nm_cap,
                                               cdbl *X, const int32_t *dn) {
  extern __shared__ double smem[];
  if (dn && (int)blockIdx.x >= dn[0]) return;
  const int ii = item_list[blockIdx.x];
  const Item it = items[ii];
  const int m = it.m;
  const int nmin = m > 0 ? m : 1;
  const int nm = it.N + 1 - nmin;
  const int64_t nm2 = (int64_t)nm * nm;
  const int n2m = 2 * nm;
  const cdbl *T = tbuf + it.t_off;
  const cdbl *A = abuf ? abuf + aoff[ii] : nullptr;
  // LDS: Ysh [G CK][2][2 nm] cdbl (G 2 nm <= rows_cap),
  // PTo [nm][n_out][2], Bv [G CK][nm][2], an [nm]
  const int rows_cap = obs_x_rows_cap(nm_cap, NT, GCAP);
  cdbl *Ysh = reinterpret_cast<cdbl *>(smem);
  // PTG: the outgoing pi/tau are read from the (L2-resident) table instead
  double *PTo = smem + 2 * (CK * 2 * rows_cap);
  double *Bv = PTo + (PTG ? 0 : (int64_t)spec.n_out * nm_cap * 2);
  double *an = Bv + (int64_t)CK * (rows_cap / 2) * 2;
  const int t = threadIdx.x;
  const int n_ang = spec.n_ang;
  const int n_out = spec.n_out;
  const double *PTm = PT + (int64_t)m * n_ang * nmax_all * 2;
  for (int e = t; !PTG && e < n_out * nm; e += NT) {
    const int p = e / n_out, o = e - (e / n_out) * n_out;
    const double *src =
        PTm + ((int64_t)spec.out_angle[o] * nmax_all + (nmin + p - 1)) * 2;
    PTo[(p * n_out + o) * 2] = src[0];
    PTo[(p * n_out + o) * 2 + 1] = src[1];
  }
  for (int p = t; p < nm; p += NT) {
    const int n = nmin + p;
    an[p] = sqrt((2. * n + 1.) / (n * (n + 1.)));
  }
  cdbl *Xo = X + (int64_t)ii * spec.n_pairs * 4;
  // chunk groups per round
  const int nchunk = (spec.n_in + CK - 1) / CK;
  if (nchunk == 0) return;
  const int gmax = n2m >= NT ? 1 : min(min(nchunk, NT / n2m), GCAP);
  const int rounds = (nchunk + gmax - 1) / gmax;
  const int G = (nchunk + rounds - 1) / rounds;
  const int gi = G == 1 ? 0 : t / n2m;

  for (int r = 0; r < rounds; ++r) {
    const int in0 = r * G * CK;
    const int in1 = min(spec.n_in, in0 + G * CK);
    __syncthreads();
    for (int e = t; e < G * CK * nm; e += NT) {
      const int ic = e / nm, p = e - (e / nm) * nm;
      double bp = 0., bt = 0.;
      if (in0 + ic < in1) {
        const double *src =
            PTm +
            ((int64_t)spec.in_angle[in0 + ic] * nmax_all + (nmin + p - 1)) * 2;
        bp = src[0];
        bt = src[1];
      }
      Bv[(ic * nm + p) * 2] = bp;
      Bv[(ic * nm + p) * 2 + 1] = bt;
    }
    __syncthreads();
    const int row0 = G == 1 ? t : t - gi * n2m;
    const int rstep = G == 1 ? NT : n2m;
    for (int row = row0; gi < G && row < n2m; row += rstep) {
      const int h = row >= nm ? 1 : 0;
      const int p = row - h * nm;
      const int n = nmin + p;
      const double a_n = an[p];
      const double *Bg = Bv + gi * CK * nm * 2;
      cdbl acc[CK][2];
#pragma unroll
      for (int ic = 0; ic < CK; ++ic) {
        acc[ic][0] = cmk(0., 0.);
        acc[ic][1] = cmk(0., 0.);
      }
      const int cls = (h + n) & 1;
      const int nh2 = A ? 2 : 1;
      for (int hh = 0; hh < nh2; ++hh) {
        for (int qb = 0; qb < nm; qb += kObsXQ) {
          // the row's next kObsXQ elements of T (one cache line), loaded
          // together: one memory latency per batch
          cdbl tq[kObsXQ];
#pragma unroll
          for (int u = 0; u < kObsXQ; ++u) {
            const int q = min(qb + u, nm - 1);
            const int h2 = A ? hh : ((cls + nmin + q) & 1);
            tq[u] = A ? A[(int64_t)row * n2m + h2 * nm + q]
                      : T[cls * nm2 + (int64_t)p * nm + q];
          }
#pragma unroll
          for (int u = 0; u < kObsXQ; ++u) {
            const int q = qb + u;
            if (q >= nm) break;
            const int nn = nmin + q;
            const int h2 = A ? hh : ((cls + nn) & 1);
            const cdbl tv = tq[u];
            // c_{nn n} T = i^e a_nn a_n T, e = nn - n - 1 mod 4: odd e
            // swaps the parts, e = 1, 2 negate the real part and e = 2, 3
            // the imaginary part
            const double f = a_n * an[q];
            const int e = (nn - n - 1) & 3;
            const double u0 = (e & 1) ? tv.y : tv.x;
            const double u1 = (e & 1) ? tv.x : tv.y;
            const cdbl cm = cmk(((e + 1) & 2) ? -f * u0 : f * u0,
                                (e & 2) ? -f * u1 : f * u1);
#pragma unroll
            for (int ic = 0; ic < CK; ++ic) {
              const double bp = Bg[(ic * nm + q) * 2];
              const double bt = Bg[(ic * nm + q) * 2 + 1];
              // b = (pi, tau) over h2, b' = (tau, pi)
              const double b0 = h2 ? bt : bp;
              const double b1 = h2 ? bp : bt;
              rfma(acc[ic][0], b0, cm);
              rfma(acc[ic][1], b1, cm);
            }
          }
        }
      }
#pragma unroll
      for (int ic = 0; ic < CK; ++ic) {
        const int icg = gi * CK + ic;
        Ysh[(icg * 2 + 0) * n2m + row] = acc[ic][0];
        Ysh[(icg * 2 + 1) * n2m + row] = acc[ic][1];
      }
    }
    __syncthreads();
    // pair contractions: X^0 = a.Y_b, X^1 = a.Y_b', X^2 = -a'.Y_b,
    // X^3 = a'.Y_b' with a = (pi, tau), a' = (tau, pi)
    const int q0 = spec.in_pair_off[in0], q1 = spec.in_pair_off[in1];
    for (int qi = q0 + t; qi < q1; qi += NT) {
      const int pr = spec.pair_by_in[qi];
      const int ic = spec.pair_in[pr] - in0;
      const int o = spec.pair_out[pr];
      const cdbl *Yb = Ysh + (ic * 2 + 0) * n2m;
      const cdbl *Ybp = Ysh + (ic * 2 + 1) * n2m;
      cdbl x0 = cmk(0., 0.), x1 = cmk(0., 0.), x2 = cmk(0., 0.),
           x3 = cmk(0., 0.);
      // rows p (half 0) and nm + p (half 1) share pi/tau of order p
      const double *pto =
          PTm + ((int64_t)spec.out_angle[o] * nmax_all + (nmin - 1)) * 2;
      for (int p = 0; p < nm; ++p) {
        const double pi = PTG ? pto[2 * p] : PTo[(p * n_out + o) * 2];
        const double ta = PTG ? pto[2 * p + 1] : PTo[(p * n_out + o) * 2 + 1];
        const cdbl y0 = Yb[p], y1 = Yb[nm + p];
        const cdbl z0 = Ybp[p], z1 = Ybp[nm + p];
        rfma(x0, pi, y0);
        rfma(x0, ta, y1);
        rfma(x1, pi, z0);
        rfma(x1, ta, z1);
        rfma(x2, ta, y0);
        rfma(x2, pi, y1);
        rfma(x3, ta, z0);
        rfma(x3, pi, z1);
      }
      Xo[(int64_t)pr * 4 + 0] = x0;
      Xo[(int64_t)pr * 4 + 1] = x1;
      Xo[(int64_t)pr * 4 + 2] = cmk(-x2.x, -x2.y);
      Xo[(int64_t)pr * 4 + 3] = x3;
    }
  }
}

// K7a: extinction and absorption per (grain, outgoing direction j), four
// directions per workgroup (kObsQJ).  Forward pair j: S = (1/k) sum_m f_m
// X_m(j).  Absorption pairs (j, i) with every quadrature azimuth:
// S(phi) = (1/k) sum_m g_m(-phi) X_m(j, i).
constexpr int kObsQJ = 4;
constexpr int kObsQPhi = 4;  // azimuths per thread (batch plan: 4 measured 0.1 ms per
                             // step faster than 2; host path: 2 best of 2, 4, 5, 10)
constexpr int kObsQM = 4;    // m values per batch of X loads

// Only grains with nq_lo <= (number of m blocks) <= nq_hi are summed (the
// others return at once): the sums of a class of grains can run as soon
// as the X of its blocks exist.
__global__ __launch_bounds__(256) void k_obs_q(
    const ObsSpec spec, const int32_t *gi_off, const GrainParams gp,
    const cdbl *X, double *qext, double *qextpol, double *qextcpol,
    double *qabs, double *qabspol, int nq_lo, int nq_hi) {
  const int njb = (spec.n_theta + kObsQJ - 1) / kObsQJ;
  const int g = blockIdx.x / njb;
  const int j0 = (blockIdx.x - g * njb) * kObsQJ;
  const int nj = min(kObsQJ, spec.n_theta - j0);
  const int i0 = gi_off[g], i1 = gi_off[g + 1];
  if (i1 - i0 < nq_lo || i1 - i0 > nq_hi) return;
  const double lam = gp.wavelength[g];
  const double k = 2. * M_PI / lam;
  const double kinv = 1. / k;
  const double rv = gp.radius[g];
  const double norm = M_PI * rv * rv;
  const double pre = 2. * M_PI / k;
  const int t = threadIdx.x;
  const int wv = t >> 6, lane = t & 63;
  __shared__ double red[4][kObsQJ][2];
  __shared__ double fwd[kObsQJ][2][2];
  if (wv < nj) {
    // forward pair j0 + wv: the lanes take every 64th m, then a shuffle
    // reduction
    double v[4] = {0., 0., 0., 0.};
    const int pr = spec.fwd_pair0 + j0 + wv;
    for (int q = i0 + lane; q < i1; q += 64) {
      const double f = (q == i0) ? 1. : 2.;
      const cdbl *Xp = X + ((int64_t)q * spec.n_pairs + pr) * 4;
      v[0] += f * Xp[0].x;
      v[1] += f * Xp[0].y;
      v[2] += f * Xp[3].x;
      v[3] += f * Xp[3].y;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
      for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
    if (lane == 0) {
      fwd[wv][0][0] = v[0] * kinv;
      fwd[wv][0][1] = v[1] * kinv;
      fwd[wv][1][0] = v[2] * kinv;
      fwd[wv][1][1] = v[3] * kinv;
    }
  }
  double zs[kObsQJ], zp[kObsQJ];
#pragma unroll
  for (int q = 0; q < kObsQJ; ++q) zs[q] = zp[q] = 0.;
  if (spec.ngauss > 0) {
    // one thread per (direction, incoming node, group of kObsQPhi
    // azimuths): X of a pair is loaded once for the group's azimuths
    const int ng = spec.ngauss;
    const int npg = (ng + kObsQPhi - 1) / kObsQPhi;
    const int per_j = ng * npg;
    const int64_t xs = (int64_t)spec.n_pairs * 4;
    const int nq = i1 - i0;
    for (int e = t; e < nj * per_j; e += 256) {
      const int jj = e / per_j, r = e - jj * per_j;
      const int ti = r / npg, pg = r - ti * npg;
      const int pj0 = pg * kObsQPhi;
      const int pr = spec.abs_pair0 + (j0 + jj) * ng + ti;
      cdbl S[kObsQPhi][4], ex[kObsQPhi], st[kObsQPhi];
#pragma unroll
      for (int u = 0; u < kObsQPhi; ++u) {
        const int pj = min(pj0 + u, ng - 1);
        st[u] = cmk(spec.phi_cos[pj], -spec.phi_sin[pj]);
        ex[u] = cmk(1., 0.);
#pragma unroll
        for (int c = 0; c < 4; ++c) S[u][c] = cmk(0., 0.);
      }
      // e^{i m dphi} by recurrence, as the reference
      const cdbl *Xb = X + ((int64_t)i0 * spec.n_pairs + pr) * 4;
      // X of kObsQM consecutive m loaded together (one memory latency per
      // batch)
      for (int mb = 0; mb < nq; mb += kObsQM) {
        cdbl xv[kObsQM][4];
#pragma unroll
        for (int d = 0; d < kObsQM; ++d)
#pragma unroll
          for (int c = 0; c < 4; ++c)
            xv[d][c] = mb + d < nq ? Xb[(int64_t)(mb + d) * xs + c]
                                   : cmk(0., 0.);
#pragma unroll
        for (int d = 0; d < kObsQM; ++d) {
          const int m = mb + d;
          if (m >= nq) break;
#pragma unroll
          for (int u = 0; u < kObsQPhi; ++u) {
            const double fc = (m == 0) ? 1. : 2. * ex[u].x;
            const double fs = (m == 0) ? 0. : 2. * ex[u].y;
            ex[u] = cmul(ex[u], st[u]);
            rfma(S[u][0], fc, xv[d][0]);
            rfma(S[u][3], fc, xv[d][3]);
            rfma(S[u][1], fs, xv[d][1]);
            rfma(S[u][2], fs, xv[d][2]);
          }
        }
      }
      double a = 0., b = 0.;
#pragma unroll
      for (int u = 0; u < kObsQPhi; ++u) {
        if (pj0 + u >= ng) break;
        double n2[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const cdbl sq = cscale(S[u][q], kinv);
          n2[q] = sq.x * sq.x + sq.y * sq.y;
        }
        const double w = spec.theta_w[ti] * spec.phi_w[pj0 + u];
        a += 0.5 * (n2[0] + n2[1] + n2[2] + n2[3]) * w;
        b += 0.5 * (n2[0] + n2[1] - n2[2] - n2[3]) * w;
      }
#pragma unroll
      for (int q = 0; q < kObsQJ; ++q) {
        if (q == jj) {
          zs[q] += a;
          zp[q] += b;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kObsQJ; ++q) {
    for (int o = 32; o > 0; o >>= 1) {
      zs[q] += __shfl_xor(zs[q], o);
      zp[q] += __shfl_xor(zp[q], o);
    }
    if (lane == 0) {
      red[wv][q][0] = zs[q];
      red[wv][q][1] = zp[q];
    }
  }
  __syncthreads();
  if (t < nj) {
    const double zsum = red[0][t][0] + red[1][t][0] + red[2][t][0] +
                        red[3][t][0];
    const double zpol = red[0][t][1] + red[1][t][1] + red[2][t][1] +
                        red[3][t][1];
    const double e_ = pre * (fwd[t][0][1] + fwd[t][1][1]);
    const double ep = pre * (fwd[t][0][1] - fwd[t][1][1]);
    const double ec = pre * (fwd[t][1][0] - fwd[t][0][0]);
    const int64_t o = (int64_t)g * spec.n_theta + j0 + t;
    if (qext) qext[o] = e_ / norm;
    if (qextpol) qextpol[o] = ep / norm;
    if (qextcpol) qextcpol[o] = ec / norm;
    if (qabs) qabs[o] = (e_ - zsum) / norm;
    if (qabspol) qabspol[o] = (ep - zpol) / norm;
  }
}

// K7b: scattering matrix Z on the (theta, phi) grid for incidence
// theta = pi/2, phi = 0 (ScatteringMatrixTask.hpp:698-751); one thread per
// (grain, theta, phi); Z row-major 4x4 per point.
__global__ void k_obs_z(const ObsSpec spec, int n_grains,
                        const int32_t *gi_off, const GrainParams gp,
                        const cdbl *X, double *Z) {
  const int nphi = spec.n_sphi;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n_grains * spec.n_stheta * nphi;
  if (idx >= total) return;
  const int g = (int)(idx / ((int64_t)spec.n_stheta * nphi));
  const int rem = (int)(idx - (int64_t)g * spec.n_stheta * nphi);
  const int it = rem / nphi, ip = rem - (rem / nphi) * nphi;
  const int i0 = gi_off[g], i1 = gi_off[g + 1];
  const double k = 2. * M_PI / gp.wavelength[g];
  const int pr = spec.sca_pair0 + it;
  cdbl S[4] = {cmk(0., 0.), cmk(0., 0.), cmk(0., 0.), cmk(0., 0.)};
  cdbl ex = cmk(1., 0.);
  const cdbl st = cmk(spec.sphi_cos[ip], spec.sphi_sin[ip]);
  for (int q = i0; q < i1; ++q) {
    const int m = q - i0;
    const cdbl *Xp = X + ((int64_t)q * spec.n_pairs + pr) * 4;
    const double fc = (m == 0) ? 1. : 2. * ex.x;
    const double fs = (m == 0) ? 0. : 2. * ex.y;
    ex = cmul(ex, st);
    rfma(S[0], fc, Xp[0]);
    rfma(S[3], fc, Xp[3]);
    rfma(S[1], fs, Xp[1]);
    rfma(S[2], fs, Xp[2]);
  }
  const double kinv = 1. / k;
  for (int q = 0; q < 4; ++q) S[q] = cscale(S[q], kinv);
  // products S_a conj(S_b)
  auto pc = [&](int a, int b) {
    return cmk(S[a].x * S[b].x + S[a].y * S[b].y,
               S[a].y * S[b].x - S[a].x * S[b].y);
  };
  const cdbl s00 = pc(0, 0), s11 = pc(1, 1), s22 = pc(2, 2), s33 = pc(3, 3);
  const cdbl s01 = pc(0, 1), s32 = pc(3, 2), s02 = pc(0, 2), s31 = pc(3, 1);
  const cdbl s03 = pc(0, 3), s12 = pc(1, 2), s21 = pc(2, 1), s20 = pc(2, 0);
  const cdbl s30 = pc(3, 0);
  double *z = Z + idx * 16;
  z[0] = 0.5 * (s00.x + s11.x + s22.x + s33.x);
  z[1] = 0.5 * (s00.x - s11.x + s22.x - s33.x);
  z[2] = -s01.x - s32.x;
  z[3] = -(s01.y - s32.y);           // Re(i (S0 S1* - S3 S2*))
  z[4] = 0.5 * (s00.x + s11.x - s22.x - s33.x);
  z[5] = 0.5 * (s00.x - s11.x - s22.x + s33.x);
  z[6] = -s01.x + s32.x;
  z[7] = -(s01.y + s32.y);
  z[8] = -s02.x - s31.x;
  z[9] = -s02.x + s31.x;
  z[10] = s03.x + s12.x;
  z[11] = (s03.y + s21.y);           // Re(-i (S0 S3* + S2 S1*))
  z[12] = -(s20.y + s31.y);          // Re(i (S2 S0* + S3 S1*))
  z[13] = -(s20.y - s31.y);
  z[14] = (s30.y - s12.y);           // Re(-i (S3 S0* - S1 S2*))
  z[15] = s30.x - s12.x;
}

hipError_t launch_pitau(const double *ang, int n_ang, int nmax_all, double *PT,
                        hipStream_t s) {
  const int n = (nmax_all + 1) * n_ang;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pitau, dim3((n + 127) / 128), dim3(128), 0, s, ang,
                     n_ang, nmax_all, PT);
  return hipGetLastError();
}

size_t obs_x_lds_bytes(int nm_cap, int n_out, int ck, int nt, int gcap,
                       bool ptg) {
  const size_t rows_cap = obs_x_rows_cap(nm_cap, nt, gcap);
  return (size_t)ck * 2 * rows_cap * sizeof(cdbl) +
         (ptg ? 0 : (size_t)n_out * nm_cap * 2 * sizeof(double)) +
         (size_t)ck * (rows_cap / 2) * 2 * sizeof(double) +
         (size_t)nm_cap * sizeof(double);
}

template <int CK, int NT, int GCAP, bool PTG>
static hipError_t launch_obs_x_ck(const ObsSpec &spec,
                                  const int32_t *item_list, int n_items,
                                  int nm_cap, const Item *items,
                                  const cdbl *tbuf, const cdbl *abuf,
                                  const int64_t *aoff, const double *PT,
                                  int nmax_all, cdbl *X, hipStream_t s,
                                  const int32_t *dn, int grid) {
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(k_obs_x<CK, NT, GCAP, PTG>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL((k_obs_x<CK, NT, GCAP, PTG>), dim3(dn ? grid : n_items),
                     dim3(NT),
                     obs_x_lds_bytes(nm_cap, spec.n_out, CK, NT, GCAP, PTG),
                     s, spec,
                     item_list, items, tbuf, abuf, aoff, PT, nmax_all, nm_cap,
                     X, dn);
  return hipGetLastError();
}

hipError_t launch_obs_x(const ObsSpec &spec, const int32_t *item_list,
                        int n_items, int nm_cap, int ck, const Item *items,
                        const cdbl *tbuf, const cdbl *abuf,
                        const int64_t *aoff, const double *PT, int nmax_all,
                        cdbl *X, hipStream_t s, const int32_t *dn,
                        int grid, bool lean) {
  if (!dn && n_items <= 0) return hipSuccess;
  // lean: the smallest LDS footprint for the largest blocks (one chunk of
  // eight directions, outgoing pi/tau from the table), for launches that
  // share the GPU with other kernels: a workgroup needing most of a CU's
  // LDS waits until the CU has drained
  if (lean && ck == 8 &&
      obs_x_lds_bytes(nm_cap, spec.n_out, 8, 256, 1, true) <= kObsXMaxLds)
    return launch_obs_x_ck<8, 256, 1, true>(spec, item_list, n_items, nm_cap,
                                            items, tbuf, abuf, aoff, PT,
                                            nmax_all, X, s, dn, grid);
  if (lean && ck == 4 && nm_cap > 32 &&
      obs_x_lds_bytes(nm_cap, spec.n_out, 4, 256, 256, true) <= kObsXMaxLds)
    return launch_obs_x_ck<4, 256, 256, true>(spec, item_list, n_items,
                                              nm_cap, items, tbuf, abuf,
                                              aoff, PT, nmax_all, X, s, dn,
                                              grid);
  // ck 8: the largest blocks, 512 threads (two 8-direction chunks side by
  // side) if the LDS allows, else 256; the outgoing pi/tau are staged in
  // LDS unless they do not fit, then read from the table (measured slower
  // where both fit)
  if (ck == 8) {
    if (obs_x_lds_bytes(nm_cap, spec.n_out, 8, 512, 2, false) <= kObsXMaxLds)
      return launch_obs_x_ck<8, 512, 2, false>(spec, item_list, n_items,
                                               nm_cap, items, tbuf, abuf,
                                               aoff, PT, nmax_all, X, s, dn,
                                               grid);
    if (obs_x_lds_bytes(nm_cap, spec.n_out, 8, 256, 1, false) <= kObsXMaxLds)
      return launch_obs_x_ck<8, 256, 1, false>(spec, item_list, n_items,
                                               nm_cap, items, tbuf, abuf,
                                               aoff, PT, nmax_all, X, s, dn,
                                               grid);
    return launch_obs_x_ck<8, 256, 1, true>(spec, item_list, n_items, nm_cap,
                                            items, tbuf, abuf, aoff, PT,
                                            nmax_all, X, s, dn, grid);
  }
  // blocks up to order 16 on 128 threads: more workgroups per CU for the
  // many small items (measured faster than 256 and 64)
  if (nm_cap <= 16 &&
      obs_x_lds_bytes(nm_cap, spec.n_out, 4, 128, 128, false) <= kObsXMaxLds)
    return launch_obs_x_ck<4, 128, 128, false>(spec, item_list, n_items,
                                               nm_cap, items, tbuf, abuf,
                                               aoff, PT, nmax_all, X, s, dn,
                                               grid);
  if (obs_x_lds_bytes(nm_cap, spec.n_out, 4, 256, 256, false) <= kObsXMaxLds)
    return launch_obs_x_ck<4, 256, 256, false>(spec, item_list, n_items,
                                               nm_cap, items, tbuf, abuf,
                                               aoff, PT, nmax_all, X, s, dn,
                                               grid);
  return launch_obs_x_ck<4, 256, 256, true>(spec, item_list, n_items, nm_cap,
                                            items, tbuf, abuf, aoff, PT,
                                            nmax_all, X, s, dn, grid);
}

hipError_t launch_obs_q(const ObsSpec &spec, int n_grains,
                        const int32_t *gi_off, const GrainParams &gp,
                        const cdbl *X, double *qext, double *qextpol,
                        double *qextcpol, double *qabs, double *qabspol,
                        hipStream_t s, int nq_lo, int nq_hi) {
  if (n_grains <= 0 || spec.n_theta <= 0) return hipSuccess;
  const int njb = (spec.n_theta + kObsQJ - 1) / kObsQJ;
  hipLaunchKernelGGL(k_obs_q, dim3(n_grains * njb), dim3(256), 0, s,
                     spec, gi_off, gp, X, qext, qextpol, qextcpol, qabs,
                     qabspol, nq_lo, nq_hi);
  return hipGetLastError();
}

hipError_t launch_obs_z(const ObsSpec &spec, int n_grains,
                        const int32_t *gi_off, const GrainParams &gp,
                        const cdbl *X, double *Z, hipStream_t s) {
  const int64_t total = (int64_t)n_grains * spec.n_stheta * spec.n_sphi;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_obs_z, dim3((unsigned)((total + 127) / 128)), dim3(128),
                     0, s, spec, n_grains, gi_off, gp, X, Z);
  return hipGetLastError();
}

}  // namespace ctm

namespace ctm {

// K9: S, Z and K for arbitrary (alpha, beta) particle orientation and
// laboratory directions (TMatrix.hpp:1109-1229,1306-1482,1891-2010): the
// directions are rotated to the particle frame (B, A_L, A_P matrices), S is
// summed over m in the particle frame, then rotated back.  One workgroup
// per query: m runs sequentially, the (n, n') double sum of each m is spread
// over the threads.
constexpr int kQueryMB = 8;  // orders m whose pi/tau are computed together

__global__ __launch_bounds__(256) void k_query(const Query *queries,
                                               const Item *items,
                                               const int32_t *gi_off,
                                               const GrainParams gp,
                                               const cdbl *tbuf,
                                               const cdbl *abuf,
                                               const int64_t *aoff,
                                               int nstride, double *out) {
  // pi/tau of both directions for kQueryMB orders m at a time (dynamic
  // LDS): ptab[(mm * 4 + d) * nstride + n - 1], d = 0, 1 pi/tau incoming,
  // 2, 3 pi/tau outgoing; nstride >= the largest order of the launch
  extern __shared__ double ptab[];
  __shared__ double geo[32];
  __shared__ double red[4][8][2];
  const Query q = queries[blockIdx.x];
  const int g = q.grain;
  const int i0 = gi_off[g];
  const int N = items[i0].N;
  const int t = threadIdx.x;
  if (t == 0) {
    const double ca = cos(q.alpha), sa = sin(q.alpha);
    const double cb = cos(q.beta), sb = sin(q.beta);
    const double cti = cos(q.theta_in), sti = sin(q.theta_in);
    const double cto = cos(q.theta_out), sto = sin(q.theta_out);
    const double cpi = cos(q.phi_in), spi = sin(q.phi_in);
    const double cpo = cos(q.phi_out), spo = sin(q.phi_out);
    double pf[2][6];  // per direction: cos th, sin th, 1/sin, cos ph, sin ph
    double AL[2][3][2], AP[2][2][3];
    for (int d = 0; d < 2; ++d) {
      const double ct = d ? cto : cti, st = d ? sto : sti;
      const double cp = d ? cpo : cpi, sp = d ? spo : spi;
      const double cpr = cp * ca + sp * sa;
      const double spr = sp * ca - cp * sa;
      const double ctp = ct * cb + st * sb * cpr;
      const double stp = sin(acos(ctp));
      double cpp, spp, stpi;
      if (stp != 0.) {
        stpi = 1. / stp;
        cpp = (cb * st * cpr - sb * ct) * stpi;
        spp = st * spr * stpi;
      } else {
        cpp = 1.;
        spp = 0.;
        stpi = 9000.;
      }
      pf[d][0] = ctp; pf[d][1] = stp; pf[d][2] = stpi;
      pf[d][3] = cpp; pf[d][4] = spp;
      AL[d][0][0] = ct * cp; AL[d][0][1] = -sp;
      AL[d][1][0] = ct * sp; AL[d][1][1] = cp;
      AL[d][2][0] = -st;     AL[d][2][1] = 0.;
      AP[d][0][0] = ctp * cpp; AP[d][0][1] = ctp * spp; AP[d][0][2] = -stp;
      AP[d][1][0] = -spp;      AP[d][1][1] = cpp;       AP[d][1][2] = 0.;
    }
    const double B[3][3] = {{ca * cb, sa * cb, -sb},
                            {-sa, ca, 0.},
                            {ca * sb, sa * sb, cb}};
    double R[2][2][2];
    for (int d = 0; d < 2; ++d) {
      double C[3][2];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j) {
          double s = 0.;
          for (int k = 0; k < 3; ++k) s += B[i][k] * AL[d][k][j];
          C[i][j] = s;
        }
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
          double s = 0.;
          for (int k = 0; k < 3; ++k) s += AP[d][i][k] * C[k][j];
          R[d][i][j] = s;
        }
    }
    const double det = 1. / (R[1][0][0] * R[1][1][1] - R[1][0][1] * R[1][1][0]);
    const double tmp = R[1][0][0];
    R[1][0][0] = R[1][1][1] * det;
    R[1][0][1] = -R[1][0][1] * det;
    R[1][1][0] = -R[1][1][0] * det;
    R[1][1][1] = tmp * det;
    for (int d = 0; d < 2; ++d)
      for (int i = 0; i < 5; ++i) geo[d * 5 + i] = pf[d][i];
    for (int d = 0; d < 2; ++d)
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) geo[10 + d * 4 + i * 2 + j] = R[d][i][j];
    // a direction along the symmetry axis (the test of pitau_dev) has
    // pi = tau = 0 for every m but 1, so only m = 1 adds to S
    geo[18] = (fabs(1. - fabs(pf[0][0])) <= 1.e-10 ||
               fabs(1. - fabs(pf[1][0])) <= 1.e-10) ? 1. : 0.;
  }
  __syncthreads();
  // e^{i(phi_out - phi_in)} in the particle frame
  const double cpi = geo[3], spi = geo[4], cpo = geo[8], spo = geo[9];
  const cdbl step = cmk(cpo * cpi + spo * spi, spo * cpi - cpo * spi);
  // orders m with a contribution: all, or only m = 1 (the terms skipped
  // are exact zeros)
  const bool axis = geo[18] != 0.;
  const int m_lo = axis ? 1 : 0, m_hi = axis ? min(1, N) : N;
  cdbl ex = axis ? step : cmk(1., 0.);  // e^{i m_lo dphi}
  cdbl S[4] = {cmk(0., 0.), cmk(0., 0.), cmk(0., 0.), cmk(0., 0.)};
  for (int mb = m_lo; mb <= m_hi; mb += kQueryMB) {
    // the recurrences of kQueryMB orders m and both directions in parallel
    __syncthreads();
    if (t < 2 * kQueryMB) {
      const int mm = t >> 1, d = t & 1;  // d 0: incoming, 1: outgoing
      if (mb + mm <= m_hi)
        pitau_dev(geo[d * 5], geo[d * 5 + 1], geo[d * 5 + 2], N, mb + mm,
                  ptab + (mm * 4 + 2 * d) * nstride,
                  ptab + (mm * 4 + 2 * d + 1) * nstride, 1);
    }
    __syncthreads();
  for (int m = mb; m <= m_hi && m < mb + kQueryMB; ++m) {
    const int nmin = m > 0 ? m : 1;
    const int nm = N + 1 - nmin;
    const double *pt = ptab + (m - mb) * 4 * nstride;
    const double fcos = (m == 0) ? 1. : 2. * ex.x;
    const double fsin = (m == 0) ? 0. : 2. * ex.y;
    ex = cmul(ex, step);
    const Item &it = items[i0 + m];
    const int64_t nm2 = (int64_t)nm * nm;
    const cdbl *T = tbuf + it.t_off;
    const cdbl *A = abuf ? abuf + aoff[i0 + m] : nullptr;
    for (int e = t; e < nm * nm; e += 256) {
      const int pn = e / nm, pnn = e - (e / nm) * nm;
      const int n = nmin + pn, nn = nmin + pnn;
      const double pin = pt[nn - 1], tin = pt[nstride + nn - 1];
      const double pout = pt[2 * nstride + n - 1];
      const double tout = pt[3 * nstride + n - 1];
      // T(h n; h' nn) for the four half blocks
      cdbl Th[2][2];
      for (int h = 0; h < 2; ++h)
        for (int h2 = 0; h2 < 2; ++h2) {
          if (A) {
            Th[h][h2] = A[(int64_t)(h * nm + pn) * 2 * nm + h2 * nm + pnn];
          } else {
            const int c1 = (h + n) & 1, c2 = (h2 + nn) & 1;
            Th[h][h2] = (c1 == c2) ? T[c1 * nm2 + (int64_t)pn * nm + pnn]
                                   : cmk(0., 0.);
          }
        }
      // c_{nn n} = i^(nn - n - 1) a_nn a_n
      const double a = sqrt((2. * nn + 1.) / (nn * (nn + 1.))) *
                       sqrt((2. * n + 1.) / (n * (n + 1.)));
      const int ph = (nn - n - 1) & 3;
      const cdbl c = ph == 0 ? cmk(a, 0.) : ph == 1 ? cmk(0., a)
                   : ph == 2 ? cmk(-a, 0.) : cmk(0., -a);
      if (m == 0) {
        const cdbl f = cscale(c, tout * tin);
        cfma(S[0], f, Th[1][1]);
        cfma(S[3], f, Th[0][0]);
      } else {
        const double pp = pout * pin, pt = pout * tin, tp = tout * pin,
                     tt = tout * tin;
        // T11 = Th[0][0], T12 = Th[0][1], T21 = Th[1][0], T22 = Th[1][1]
        cdbl s0 = cadd(cadd(cscale(Th[0][0], pp), cscale(Th[1][0], tp)),
                       cadd(cscale(Th[0][1], pt), cscale(Th[1][1], tt)));
        cdbl s1 = cadd(cadd(cscale(Th[0][0], pt), cscale(Th[1][0], tt)),
                       cadd(cscale(Th[0][1], pp), cscale(Th[1][1], tp)));
        cdbl s2 = cadd(cadd(cscale(Th[0][0], tp), cscale(Th[1][0], pp)),
                       cadd(cscale(Th[0][1], tt), cscale(Th[1][1], pt)));
        cdbl s3 = cadd(cadd(cscale(Th[0][0], tt), cscale(Th[1][0], pt)),
                       cadd(cscale(Th[0][1], tp), cscale(Th[1][1], pp)));
        cfma(S[0], cscale(c, fcos), s0);
        cfma(S[1], cscale(c, fsin), s1);
        cfma(S[2], cscale(c, -fsin), s2);
        cfma(S[3], cscale(c, fcos), s3);
      }
    }
  }
  }
  // block reduction of S
  for (int k = 0; k < 4; ++k) {
    double x = S[k].x, y = S[k].y;
    for (int o = 32; o > 0; o >>= 1) {
      x += __shfl_xor(x, o);
      y += __shfl_xor(y, o);
    }
    if ((t & 63) == 0) {
      red[k][t >> 6][0] = x;
      red[k][t >> 6][1] = y;
    }
  }
  __syncthreads();
  if (t == 0) {
    const double kinv = gp.wavelength[g] / (2. * M_PI);
    cdbl s[4];
    for (int k = 0; k < 4; ++k) {
      double x = 0., y = 0.;
      for (int w = 0; w < 4; ++w) {
        x += red[k][w][0];
        y += red[k][w][1];
      }
      s[k] = cmk(x * kinv, y * kinv);
    }
    const double *Ri = geo + 10, *Ro = geo + 14;
    // S' = R_out^-1 S R_in
    const cdbl c11 = cadd(cscale(s[0], Ri[0]), cscale(s[1], Ri[2]));
    const cdbl c12 = cadd(cscale(s[0], Ri[1]), cscale(s[1], Ri[3]));
    const cdbl c21 = cadd(cscale(s[2], Ri[0]), cscale(s[3], Ri[2]));
    const cdbl c22 = cadd(cscale(s[2], Ri[1]), cscale(s[3], Ri[3]));
    cdbl F[4];
    F[0] = cadd(cscale(c11, Ro[0]), cscale(c21, Ro[1]));
    F[1] = cadd(cscale(c12, Ro[0]), cscale(c22, Ro[1]));
    F[2] = cadd(cscale(c11, Ro[2]), cscale(c21, Ro[3]));
    F[3] = cadd(cscale(c12, Ro[2]), cscale(c22, Ro[3]));
    double *o = out + (int64_t)blockIdx.x * kQueryOut;
    for (int k = 0; k < 4; ++k) {
      o[2 * k] = F[k].x;
      o[2 * k + 1] = F[k].y;
    }
    auto pc = [&](int a, int b) {
      return cmk(F[a].x * F[b].x + F[a].y * F[b].y,
                 F[a].y * F[b].x - F[a].x * F[b].y);
    };
    const cdbl s00 = pc(0, 0), s11 = pc(1, 1), s22 = pc(2, 2), s33 = pc(3, 3);
    const cdbl s01 = pc(0, 1), s32 = pc(3, 2), s02 = pc(0, 2), s31 = pc(3, 1);
    const cdbl s03 = pc(0, 3), s12 = pc(1, 2), s21 = pc(2, 1), s20 = pc(2, 0);
    const cdbl s30 = pc(3, 0);
    double *z = o + 8;
    z[0] = 0.5 * (s00.x + s11.x + s22.x + s33.x);
    z[1] = 0.5 * (s00.x - s11.x + s22.x - s33.x);
    z[2] = -s01.x - s32.x;
    z[3] = -(s01.y - s32.y);
    z[4] = 0.5 * (s00.x + s11.x - s22.x - s33.x);
    z[5] = 0.5 * (s00.x - s11.x - s22.x + s33.x);
    z[6] = -s01.x + s32.x;
    z[7] = -(s01.y + s32.y);
    z[8] = -s02.x - s31.x;
    z[9] = -s02.x + s31.x;
    z[10] = s03.x + s12.x;
    z[11] = (s03.y + s21.y);
    z[12] = -(s20.y + s31.y);
    z[13] = -(s20.y - s31.y);
    z[14] = (s30.y - s12.y);
    z[15] = s30.x - s12.x;
    // extinction matrix (TMatrix.hpp:1984-2007) with S of the same
    // direction pair
    const double pre = 2. * M_PI / (2. * M_PI / gp.wavelength[g]);
    double *K = o + 24;
    K[0] = pre * (F[0].y + F[3].y);
    K[5] = K[0];
    K[10] = K[0];
    K[15] = K[0];
    K[1] = pre * (F[0].y - F[3].y);
    K[4] = K[1];
    K[2] = -pre * (F[1].y + F[2].y);
    K[8] = K[2];
    K[3] = pre * (F[2].x - F[1].x);
    K[12] = K[3];
    K[6] = pre * (F[2].y - F[1].y);
    K[9] = -K[6];
    K[7] = -pre * (F[1].x + F[2].x);
    K[13] = -K[7];
    K[11] = pre * (F[3].x - F[0].x);
    K[14] = -K[11];
  }
}

hipError_t launch_query(const Query *q, int n, int nmax, const Item *items,
                        const int32_t *gi_off, const GrainParams &gp,
                        const cdbl *tbuf, const cdbl *abuf,
                        const int64_t *aoff, double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(k_query),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr = true;
  }
  const int nstride = nmax > 0 ? nmax : 1;
  const size_t lds = (size_t)kQueryMB * 4 * nstride * sizeof(double);
  hipLaunchKernelGGL(k_query, dim3(n), dim3(256), lds, s, q, items, gi_off,
                     gp, tbuf, abuf, aoff, nstride, out);
  return hipGetLastError();
}

}  // namespace ctm
