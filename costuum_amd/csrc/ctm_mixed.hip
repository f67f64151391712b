// ctm_mixed.hip -- mixed-precision solve of the all-m phase.
//
// The reference inverts every Q block in FP64 (Matrix::plu_inverse,
// Matrix.hpp:261-384) and forms T = -RgQ Q^-1 (TMatrixResource.hpp:
// 1102-1119).  Here, for blocks of order 17..99:
//   1. k_invert32_blk: the class block Q_c is equilibrated with powers of
//      two (row scales, then column scales of the row-scaled block: the
//      EBCM Q entries span far more than FP32's range) and eliminated in
//      FP32 by the same blocked Gauss-Jordan as the FP64 solve (ctm_gj.hpp,
//      same pivot rule on the FP32 magnitudes).  Its inverse X (FP32) and
//      the scale exponents go to HBM.
//   2. k_refine: FP64 products against the FP64 Q and RgQ: T = -RgQ X, then
//      sweeps R = -RgQ - T Q, T += R X (classical iterative refinement of
//      T Q = -RgQ with X as the approximate inverse).  Each sweep shrinks
//      the error by rho ~ cond(Q) 2^-24; a sweep's correction d_k estimates
//      the error before it, so the error after it is about d_k^2 / d_(k-1).
//      The task stops when that estimate is below kRefineTol of max |T|; it
//      is flagged for the FP64 solve when the sweeps stall above
//      kRefineStall of max |T| or kRefineMaxIter sweeps do not reach the
//      bar.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctm_gj.hpp"
#include "ctm_mixed.hpp"

namespace ctm {

namespace {

// e with 2^(e - 1) <= v < 2^e (v > 0), 0 for v = 0 or non-finite v
__device__ inline int exp2_above(double v) {
  int e = 0;
  if (v > 0. && v < INFINITY) (void)frexp(v, &e);
  return e;
}

__device__ inline double wave_max_f64(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// One workgroup per task.  dn (may be NULL): the task count is read from
// dn[0]; workgroups beyond it return at once.
template <int NT, int RS, int W>
__global__ __launch_bounds__(NT) void k_invert32_blk(const int32_t *tasks,
                                                     const Item *items,
                                                     const cdbl *qbuf,
                                                     cflt *xbuf,
                                                     int32_t *xexp,
                                                     const int32_t *dn) {
  constexpr int NW = NT / 64;
  static_assert(NW >= 4, "two waves per sub-block at least");
  extern __shared__ double smem[];
  __shared__ int stp[2][2][W];

  if (dn && (int)blockIdx.x >= dn[0]) return;
  const int task = tasks[blockIdx.x];
  const int cls = task & 1;
  const Item it = items[task >> 1];
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
  cflt *A = reinterpret_cast<cflt *>(smem);
  // odd row strides: conflict-free column accesses (as in the FP64 solve)
  const int ldb[2] = {blk_lds_stride(sz[0]), blk_lds_stride(sz[1])};
  cflt *Ab[2] = {A, A + (int64_t)sz0 * ldb[0]};
  const int64_t nstore = (int64_t)sz[0] * ldb[0] + (int64_t)sz[1] * ldb[1];
  int *piv = reinterpret_cast<int *>(A + nstore);
  int *src = piv + nm;
  int *er = src + nm;  // row exponents (local index)
  int *ec = er + nm;   // column exponents (local index)
  const cdbl *Qg = qbuf + it.q_off + cls * nm2;
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;

  // row scales: one wave per row of its sub-block
  for (int L = wv; L < nm; L += NW) {
    const int b = (nb == 2 && L >= sz0) ? 1 : 0;
    const int o = off[b], n = sz[b];
    const int64_t p = pidx(L);
    double mx = 0.;
    for (int j = lane; j < n; j += 64) {
      const cdbl v = Qg[p * nm + pidx(o + j)];
      mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
    }
    mx = wave_max_f64(mx);
    if (lane == 0) er[L] = -exp2_above(mx);
  }
  __syncthreads();
  // column scales of the row-scaled block: one wave per column
  for (int L = wv; L < nm; L += NW) {
    const int b = (nb == 2 && L >= sz0) ? 1 : 0;
    const int o = off[b], n = sz[b];
    const int pc = pidx(L);
    double mx = 0.;
    for (int i = lane; i < n; i += 64) {
      const cdbl v = Qg[(int64_t)pidx(o + i) * nm + pc];
      mx = fmax(mx, ldexp(fmax(fabs(v.x), fabs(v.y)), er[o + i]));
    }
    mx = wave_max_f64(mx);
    if (lane == 0) ec[L] = -exp2_above(mx);
  }
  __syncthreads();
  // the equilibrated sub-blocks in FP32
#pragma unroll
  for (int bb = 0; bb < 2; ++bb) {
    if (bb >= nb) continue;
    const int o = off[bb], n = sz[bb];
    for (int i = t >> 5; i < n; i += NT >> 5) {
      const int64_t pi = pidx(o + i);
      for (int j = t & 31; j < n; j += 32) {
        const cdbl v = Qg[pi * nm + pidx(o + j)];
        const int e = er[o + i] + ec[o + j];
        Ab[bb][(int64_t)i * ldb[bb] + j] =
            cmkf((float)ldexp(v.x, e), (float)ldexp(v.y, e));
      }
    }
  }
  __syncthreads();
  // blocked Gauss-Jordan, as in the FP64 kernel (k_invert_solve_blk)
  const int GW = NW / nb;
  const int b = wv / GW, gw = wv - b * GW;
  const int lo = off[b], hi = off[b] + sz[b];
  const int npan = (sz[b] + W - 1) / W;
  const int npmax = (sz[0] + W - 1) / W;
  cflt *Av = Ab[b];
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
        const int nwk = ahead ? GW - 1 : GW, rank = ahead ? gw - 1 : gw;
        blk_update<RS, W>(Av, ld, lo, hi, c0, w, steps, lo + rank, hi, nwk,
                          c0, ahead ? c1 + w1 : c1, lane);
      }
    }
    __syncthreads();
  }
  // column unpermutation: (D_r Q D_c)^-1 (k, c) = A(k, src(c)) within a
  // sub-block; Q^-1 (k, c) = 2^(ec[k]) (D_r Q D_c)^-1 (k, c) 2^(er[c])
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
  cflt *X = xbuf + it.t_off + cls * nm2;
  for (int k = t >> 5; k < nm; k += NT >> 5) {
    const int bk = (nb == 2 && k >= sz0) ? 1 : 0;
    for (int c = t & 31; c < nm; c += 32) {
      const int bc = (nb == 2 && c >= sz0) ? 1 : 0;
      cflt v = cmkf(0.f, 0.f);
      if (bk == bc)
        v = Ab[bk][(int64_t)(k - off[bk]) * ldb[bk] + (src[c] - off[bk])];
      X[(int64_t)pidx(k) * nm + pidx(c)] = v;
    }
  }
  int32_t *xe = xexp + it.t_off + (int64_t)cls * 2 * nm;
  for (int L = t; L < nm; L += NT) {
    xe[pidx(L)] = ec[L];
    xe[nm + pidx(L)] = er[L];
  }
}

// C = A B for one task's nm x nm class blocks, thread (ty, tx) of a TY x
// TX grid holding rows ty + TY a (a < RY) and columns tx + TX b (b < RX);
// A and B pass through LDS in chunks of KC columns / rows.  For m = 0 the
// entries coupling rows and columns of different parity are zero
// (m0mask), whatever the buffers hold there.
template <int NT, int TX, int RY, int RX, int KC, class LA, class LB>
__device__ __forceinline__ void block_gemm(int nm, bool m0mask, LA loadA,
                                           LB loadB, cdbl (*As)[NT / TX * RY + 1],
                                           cdbl (*Bs)[TX * RX + 1],
                                           cdbl (&acc)[RY][RX]) {
  constexpr int TY = NT / TX, MR = TY * RY, MC = TX * RX;
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
#pragma unroll
  for (int a = 0; a < RY; ++a)
#pragma unroll
    for (int b = 0; b < RX; ++b) acc[a][b] = cmk(0., 0.);
  for (int k0 = 0; k0 < nm; k0 += KC) {
    __syncthreads();
    for (int e = t; e < KC * MR; e += NT) {
      const int r = e / KC, kk = e - r * KC, k = k0 + kk;
      As[kk][r] = (r < nm && k < nm && !(m0mask && ((r ^ k) & 1)))
                      ? loadA(r, k)
                      : cmk(0., 0.);
    }
    for (int e = t; e < KC * MC; e += NT) {
      const int kk = e / MC, c = e - kk * MC, k = k0 + kk;
      Bs[kk][c] = (c < nm && k < nm && !(m0mask && ((k ^ c) & 1)))
                      ? loadB(k, c)
                      : cmk(0., 0.);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      cdbl av[RY], bv[RX];
#pragma unroll
      for (int a = 0; a < RY; ++a) av[a] = As[kk][ty + TY * a];
#pragma unroll
      for (int b = 0; b < RX; ++b) bv[b] = Bs[kk][tx + TX * b];
#pragma unroll
      for (int a = 0; a < RY; ++a)
#pragma unroll
        for (int b = 0; b < RX; ++b) cfma(acc[a][b], av[a], bv[b]);
    }
  }
}

// Uniform maximum of v over the workgroup (red: NT / 64 doubles of LDS).
template <int NT>
__device__ __forceinline__ double block_max(double v, double *red) {
  v = wave_max_f64(v);
  const int t = threadIdx.x;
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < NT / 64; ++w) m = fmax(m, red[w]);
  return m;
}

// One workgroup per task; flags[i] belongs to tasks[i].  dn (may be NULL):
// the task count is read from dn[0]; workgroups beyond it return at once.
template <int NT, int TX, int RY, int RX>
__global__ __launch_bounds__(NT) void k_refine(
    const int32_t *tasks, const Item *items, const cdbl *qbuf,
    const cflt *xbuf, const int32_t *xexp, cdbl *tbuf, cdbl *rbuf,
    double *qparts, int32_t *flags, const int32_t *dn) {
  constexpr int TY = NT / TX, MR = TY * RY, MC = TX * RX, KC = 8;
  constexpr int NW = NT / 64;
  __shared__ cdbl As[KC][MR + 1];
  __shared__ cdbl Bs[KC][MC + 1];
  __shared__ int xr[MR > MC ? MR : MC], xc[MR > MC ? MR : MC];
  __shared__ double red[NW];
  __shared__ double red4[NW][4];

  if (dn && (int)blockIdx.x >= dn[0]) return;
  const int task = tasks[blockIdx.x];
  const int cls = task & 1;
  const Item it = items[task >> 1];
  const int m = it.m;
  const bool m0 = m == 0;
  const int nmin = m > 0 ? m : 1;
  const int nm = it.N + 1 - nmin;
  const int64_t nm2 = (int64_t)nm * nm;
  const cdbl *Q = qbuf + it.q_off + cls * nm2;
  const cdbl *RQ = qbuf + it.q_off + (2 + cls) * nm2;
  const cflt *X = xbuf + it.t_off + cls * nm2;
  const int32_t *xe = xexp + it.t_off + (int64_t)cls * 2 * nm;
  cdbl *T = tbuf + it.t_off + cls * nm2;
  cdbl *R = rbuf + it.t_off + cls * nm2;
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  for (int i = t; i < nm; i += NT) {
    xr[i] = xe[i];
    xc[i] = xe[nm + i];
  }
  __syncthreads();
  auto ldX = [&](int k, int j) {
    const cflt v = X[(int64_t)k * nm + j];
    const int e = xr[k] + xc[j];
    return cmk(ldexp((double)v.x, e), ldexp((double)v.y, e));
  };
  auto ldQ = [&](int k, int j) { return Q[(int64_t)k * nm + j]; };
  auto ldRQ = [&](int i, int k) { return RQ[(int64_t)i * nm + k]; };
  auto ldT = [&](int i, int k) { return T[(int64_t)i * nm + k]; };
  auto ldR = [&](int i, int k) { return R[(int64_t)i * nm + k]; };
  auto cross = [&](int i, int j) { return m0 && ((i ^ j) & 1); };

  cdbl acc[RY][RX];
  // T0 = -RgQ X
  block_gemm<NT, TX, RY, RX, KC>(nm, m0, ldRQ, ldX, As, Bs, acc);
  double tm = 0.;
#pragma unroll
  for (int a = 0; a < RY; ++a)
#pragma unroll
    for (int b = 0; b < RX; ++b) {
      const int i = ty + TY * a, j = tx + TX * b;
      if (i < nm && j < nm) {
        const cdbl v = cross(i, j) ? cmk(0., 0.) : cmk(-acc[a][b].x,
                                                       -acc[a][b].y);
        T[(int64_t)i * nm + j] = v;
        tm = fmax(tm, fmax(fabs(v.x), fabs(v.y)));
      }
    }
  double tmax = block_max<NT>(tm, red);
  double dprev = tmax;
  int ok = 0;
  for (int sweep = 0; sweep < kRefineMaxIter; ++sweep) {
    // P = T Q (T of the previous step, written by this workgroup)
    block_gemm<NT, TX, RY, RX, KC>(nm, m0, ldT, ldQ, As, Bs, acc);
#pragma unroll
    for (int a = 0; a < RY; ++a)
#pragma unroll
      for (int b = 0; b < RX; ++b) {
        const int i = ty + TY * a, j = tx + TX * b;
        if (i < nm && j < nm) {
          const cdbl rq = RQ[(int64_t)i * nm + j];
          R[(int64_t)i * nm + j] =
              cross(i, j) ? cmk(0., 0.)
                          : cmk(-rq.x - acc[a][b].x, -rq.y - acc[a][b].y);
        }
      }
    // D = R X;  T += D
    block_gemm<NT, TX, RY, RX, KC>(nm, m0, ldR, ldX, As, Bs, acc);
    double dm = 0.;
    tm = 0.;
#pragma unroll
    for (int a = 0; a < RY; ++a)
#pragma unroll
      for (int b = 0; b < RX; ++b) {
        const int i = ty + TY * a, j = tx + TX * b;
        if (i < nm && j < nm && !cross(i, j)) {
          cdbl v = T[(int64_t)i * nm + j];
          v.x += acc[a][b].x;
          v.y += acc[a][b].y;
          T[(int64_t)i * nm + j] = v;
          dm = fmax(dm, fmax(fabs(acc[a][b].x), fabs(acc[a][b].y)));
          tm = fmax(tm, fmax(fabs(v.x), fabs(v.y)));
        }
      }
    const double dmax = block_max<NT>(dm, red);
    tmax = block_max<NT>(tm, red);
    // predicted error after this sweep: d_k (d_k / d_(k-1))
    const double est = dmax * fmin(1., dmax / dprev);
    if (est <= kRefineTol * tmax) {
      ok = 1;
      break;
    }
    if (!(dmax <= 0.5 * dprev)) {  // stalled (or not finite)
      ok = dmax <= kRefineStall * tmax ? 1 : 0;
      break;
    }
    dprev = dmax;
  }
  // per-task sums of the final T (the solve kernels' qparts)
  double s_diag2 = 0., e_diag = 0., s_all = 0., e_tr = 0.;
#pragma unroll
  for (int a = 0; a < RY; ++a)
#pragma unroll
    for (int b = 0; b < RX; ++b) {
      const int i = ty + TY * a, j = tx + TX * b;
      if (i < nm && j < nm) {
        const cdbl v = T[(int64_t)i * nm + j];
        const double a2 = v.x * v.x + v.y * v.y;
        s_all += a2;
        if (i == j) {
          const double w = 2. * (nmin + i) + 1.;
          s_diag2 += w * a2;
          e_diag += w * v.x;
          e_tr += v.x;
        }
      }
    }
  double v4[4] = {s_diag2, e_diag, s_all, e_tr};
#pragma unroll
  for (int q = 0; q < 4; ++q)
    for (int o = 32; o > 0; o >>= 1) v4[q] += __shfl_xor(v4[q], o);
  if ((t & 63) == 0)
    for (int q = 0; q < 4; ++q) red4[t >> 6][q] = v4[q];
  __syncthreads();
  if (t == 0) {
    for (int q = 0; q < 4; ++q) {
      double s = 0.;
      for (int w = 0; w < NW; ++w) s += red4[w][q];
      qparts[(int64_t)task * kQParts + q] = s;
    }
    flags[blockIdx.x] = ok ? 0 : 1;
  }
}

template <int NT, int RS, int W>
hipError_t launch_inv32(const int32_t *tasks, int n, size_t lds,
                        const Item *items, const cdbl *qbuf, cflt *xbuf,
                        int32_t *xexp, hipStream_t s, const int32_t *dn,
                        int grid) {
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(k_invert32_blk<NT, RS, W>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_invert32_blk<NT, RS, W>), dim3(dn ? grid : n),
                     dim3(NT), lds, s, tasks, items, qbuf, xbuf, xexp, dn);
  return hipGetLastError();
}

template <int NT, int TX, int RY, int RX>
hipError_t launch_ref(const int32_t *tasks, int n, const Item *items,
                      const cdbl *qbuf, const cflt *xbuf,
                      const int32_t *xexp, cdbl *tbuf, cdbl *rbuf,
                      double *qparts, int32_t *flags, hipStream_t s,
                      const int32_t *dn, int grid) {
  hipLaunchKernelGGL((k_refine<NT, TX, RY, RX>), dim3(dn ? grid : n),
                     dim3(NT), 0, s, tasks, items, qbuf, xbuf, xexp, tbuf,
                     rbuf, qparts, flags, dn);
  return hipGetLastError();
}

}  // namespace

size_t invert32_lds_bytes(int64_t store_max, int nm_max) {
  return (size_t)store_max * sizeof(cflt) + 4 * (size_t)nm_max * sizeof(int);
}

hipError_t launch_invert32(const int32_t *tasks, int n, int nm_max,
                           int64_t store_max, int rows_max, const Item *items,
                           const cdbl *qbuf, cflt *xbuf, int32_t *xexp,
                           hipStream_t s, const int32_t *dn, int grid) {
  if (!dn && n <= 0) return hipSuccess;
  if (nm_max < 1 || nm_max > kMixedMaxOrder || rows_max > 2 * 64 ||
      store_max > (int64_t)nm_max * blk_lds_stride(nm_max))
    return hipErrorInvalidValue;
  const size_t lds = invert32_lds_bytes(store_max, nm_max);
  if (nm_max <= 48 || rows_max <= 32)
    return launch_inv32<256, 1, 4>(tasks, n, lds, items, qbuf, xbuf, xexp, s,
                                   dn, grid);
  if (rows_max <= 64)
    return launch_inv32<512, 1, 4>(tasks, n, lds, items, qbuf, xbuf, xexp, s,
                                   dn, grid);
  return launch_inv32<512, 2, 4>(tasks, n, lds, items, qbuf, xbuf, xexp, s,
                                 dn, grid);
}

hipError_t launch_refine(const int32_t *tasks, int n, int nm_max,
                         const Item *items, const cdbl *qbuf,
                         const cflt *xbuf, const int32_t *xexp, cdbl *tbuf,
                         cdbl *rbuf, double *qparts, int32_t *flags,
                         hipStream_t s, const int32_t *dn, int grid) {
  if (!dn && n <= 0) return hipSuccess;
  if (nm_max < 1 || nm_max > 128) return hipErrorInvalidValue;
  if (nm_max <= 32)
    return launch_ref<256, 16, 2, 2>(tasks, n, items, qbuf, xbuf, xexp, tbuf,
                                     rbuf, qparts, flags, s, dn, grid);
  if (nm_max <= 64)
    return launch_ref<256, 16, 4, 4>(tasks, n, items, qbuf, xbuf, xexp, tbuf,
                                     rbuf, qparts, flags, s, dn, grid);
  return launch_ref<1024, 32, 4, 4>(tasks, n, items, qbuf, xbuf, xexp, tbuf,
                                    rbuf, qparts, flags, s, dn, grid);
}

}  // namespace ctm
