// ctm_average.hip -- gfx950 kernels for the orientation-averaged (ensemble)
// T-matrix of Mishchenko (1991) eq. 3.27 (AlignmentAverageTask.hpp:115-215,
// TMatrixCalculator.hpp:229-307):
//   <T>(m; h1 n1, h2 n2) = (-1)^m sum_N p_N C^N(n1 m, n2 -m)
//                           B_N(h1 h2; n1, n2)
//   B_N(h1 h2; n1, n2) = sum_{m1 <= min(n1,n2)} f_m1 (-1)^m1
//                           C^N(n1 m1, n2 -m1) T_m1(h1 n1; h2 n2)
// with (h1, h2) in {(0,0), (1,1)} for N + n1 + n2 even and {(0,1), (1,0)}
// otherwise.  B_N does not depend on m, so the reference's inner
// recomputation over m (O(N^6) as written) becomes O(N^4) here.
#include "ctm_average.hpp"

namespace ctm {

// K8a: Clebsch-Gordan coefficients <n1 m n2 -m | N 0> for m = 0..min(n1,n2)
// by the reference's ladder recursion (SpecialFunctions.hpp:1131-1244): the
// top row M = N from C(m1 = n1) = 1 downwards, normalised, then rows
// M = N-1 .. 0 from the row above.  One 64-thread workgroup per (n1, n2, N);
// two LDS rows of 2 n1 + 1 entries.
__global__ __launch_bounds__(64) void k_cg(const CgKey *keys, double *cg) {
  __shared__ double rows[2][2 * kMaxAvgOrder + 1];
  const CgKey key = keys[blockIdx.x];
  const int n1 = key.n1, n2 = key.n2, N = key.N;
  const int ns = 2 * n1 + 1;
  const int t = threadIdx.x;
  double *top = rows[N & 1];
  for (int i = t; i < ns; i += 64) top[i] = 0.;
  __syncthreads();
  if (t == 0) {
    top[n1 + n1] = 1.;
    double norm = 1.;
    const int lo = (-n1 > N - n2 - 1) ? -n1 : N - n2 - 1;
    for (int m1 = n1 - 1; m1 >= lo; --m1) {
      const double cp2 = sqrt((n2 - N + m1 + 1.) * (n2 + N - m1));
      const double cp1 = sqrt((n1 - m1) * (n1 + m1 + 1.));
      const double next = -cp2 * top[m1 + 1 + n1] / cp1;
      top[m1 + n1] = next;
      norm += next * next;
    }
    const double inorm = 1. / sqrt(norm);
    for (int m1 = n1; m1 >= lo; --m1) top[m1 + n1] *= inorm;
  }
  __syncthreads();
  for (int M = N - 1; M >= 0; --M) {
    const double *up = rows[(M + 1) & 1];
    double *cur = rows[M & 1];
    const double cmi = 1. / sqrt((N + M + 1.) * (N - M));
    const int a = (-n1 > M - n2) ? -n1 : M - n2;
    const int b = (n1 < M + n2) ? n1 : M + n2;
    for (int i = t; i < ns; i += 64) {
      const int m1 = i - n1;
      double v = 0.;
      if (m1 >= a && m1 <= b) {
        const int m2 = M - m1;
        double t1 = 0., t2 = 0.;
        if (m1 + 1 <= n1) t1 = sqrt((n1 + m1 + 1.) * (n1 - m1)) * up[i + 1];
        if (m2 + 1 <= n2) t2 = sqrt((n2 + m2 + 1.) * (n2 - m2)) * up[i];
        v = cmi * (t1 + t2);
      }
      cur[i] = v;
    }
    __syncthreads();
  }
  const int mmax = n1 < n2 ? n1 : n2;
  for (int m = t; m <= mmax; m += 64) cg[key.off + m] = rows[0][m + n1];
}

// K8b: ensemble T-matrix for one (grain, n1, n2): B_N for every N with a
// non-zero coefficient, then <T> for every m <= min(n1, n2), written to the
// full block layout (row (h n1), column (h' n2) of the (2 nm)^2 block m).
__global__ __launch_bounds__(64) void k_average(
    const AvgTask *tasks, const Item *items, const int32_t *gi_off,
    const cdbl *tbuf, const double *pN, int ncoef, const int32_t *dist,
    const int32_t *cg_index, const double *cg, int nmax_cg, cdbl *abuf,
    const int64_t *aoff) {
  // B_N for N = lo..hi (at most 2 nmax_cg + 1 rows of two values)
  extern __shared__ cdbl Bsh[];
  const AvgTask task = tasks[blockIdx.x];
  const int g = task.grain, n1 = task.n1, n2 = task.n2;
  const int i0 = gi_off[g];
  const int Ng = items[i0].N;
  const int M = n1 < n2 ? n1 : n2;
  const int lo = n1 > n2 ? n1 - n2 : n2 - n1;
  const int hi = n1 + n2;
  const int t = threadIdx.x;
  if (dist && dist[g] < 0) {
    // alignment disabled for this grain: the ensemble T-matrix is the
    // grain's own (AlignmentAverageTask.hpp:195-214), written in the full
    // (2 nm)^2 layout; pairs of different class are zero
    for (int m = t; m <= M; m += 64) {
      const int nmin = m > 0 ? m : 1;
      const int nm = Ng + 1 - nmin;
      const int64_t nm2 = (int64_t)nm * nm;
      const cdbl *T = tbuf + items[i0 + m].t_off;
      cdbl *A = abuf + aoff[i0 + m];
      const int p1 = n1 - nmin, p2 = n2 - nmin;
      for (int h1 = 0; h1 < 2; ++h1)
        for (int h2 = 0; h2 < 2; ++h2) {
          const int c1 = (h1 + n1) & 1, c2 = (h2 + n2) & 1;
          A[(int64_t)(h1 * nm + p1) * (2 * nm) + h2 * nm + p2] =
              c1 == c2 ? T[c1 * nm2 + (int64_t)p1 * nm + p2] : cmk(0., 0.);
        }
    }
    return;
  }
  const double *p = pN + (int64_t)(dist ? dist[g] : 0) * ncoef;
  // C^N(n1 m, n2 -m) for m = 0..M lives at cg[cg_index[(n1, n2, N)] + m]
  auto cgp = [&](int N) -> const double * {
    const int idx = cg_index[((int64_t)n1 * (nmax_cg + 1) + n2) *
                                 (2 * nmax_cg + 1) + N];
    return idx >= 0 ? cg + idx : nullptr;
  };
  for (int N = lo + t; N <= hi; N += 64) {
    cdbl b0 = cmk(0., 0.), b1 = cmk(0., 0.);
    const double *C = (N < ncoef && p[N] != 0.) ? cgp(N) : nullptr;
    if (C) {
      const bool even = ((N + hi) & 1) == 0;
      for (int m1 = 0; m1 <= M; ++m1) {
        const Item &it = items[i0 + m1];
        const int nmin = m1 > 0 ? m1 : 1;
        const int nm = Ng + 1 - nmin;
        const int64_t nm2 = (int64_t)nm * nm;
        const cdbl *T = tbuf + it.t_off;
        const double f = ((m1 == 0) ? 1. : 2.) * ((m1 & 1) ? -1. : 1.) * C[m1];
        // blocks (h1, h2): even -> (0,0) into b0, (1,1) into b1;
        // odd -> (0,1) into b0, (1,0) into b1
        const int p1 = n1 - nmin, p2 = n2 - nmin;
#pragma unroll
        for (int blk = 0; blk < 2; ++blk) {
          const int h1 = blk;
          const int h2 = even ? blk : 1 - blk;
          const int c1 = (h1 + n1) & 1, c2 = (h2 + n2) & 1;
          if (c1 != c2) continue;
          const cdbl v = T[c1 * nm2 + (int64_t)p1 * nm + p2];
          cdbl &acc = blk == 0 ? b0 : b1;
          acc.x = fma(f, v.x, acc.x);
          acc.y = fma(f, v.y, acc.y);
        }
      }
    }
    Bsh[2 * (N - lo)] = b0;
    Bsh[2 * (N - lo) + 1] = b1;
  }
  __syncthreads();
  for (int m = t; m <= M; m += 64) {
    const int nmin = m > 0 ? m : 1;
    const int nm = Ng + 1 - nmin;
    const int n2m = 2 * nm;
    cdbl acc[2][2] = {{cmk(0., 0.), cmk(0., 0.)}, {cmk(0., 0.), cmk(0., 0.)}};
    for (int N = lo; N <= hi; ++N) {
      if (N >= ncoef || p[N] == 0.) continue;
      const double *C = cgp(N);
      if (!C) continue;
      const double w = ((m & 1) ? -1. : 1.) * p[N] * C[m];
      const bool even = ((N + hi) & 1) == 0;
      const cdbl b0 = Bsh[2 * (N - lo)], b1 = Bsh[2 * (N - lo) + 1];
      if (even) {
        rfma(acc[0][0], w, b0);
        rfma(acc[1][1], w, b1);
      } else {
        rfma(acc[0][1], w, b0);
        rfma(acc[1][0], w, b1);
      }
    }
    cdbl *A = abuf + aoff[i0 + m];
    const int p1 = n1 - nmin, p2 = n2 - nmin;
    for (int h1 = 0; h1 < 2; ++h1)
      for (int h2 = 0; h2 < 2; ++h2)
        A[(int64_t)(h1 * nm + p1) * n2m + h2 * nm + p2] = acc[h1][h2];
  }
}

// Grid-stride zero fill, two complex values per thread and pass.
__global__ __launch_bounds__(256) void k_zero(cdbl *a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
    a[i] = cmk(0., 0.);
}

hipError_t launch_zero(cdbl *a, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, a, n);
  return hipGetLastError();
}

// One thread per output value; the weighted sum runs over the group in
// order, as the reference's shape average does
// (ShapeAveragingTask.hpp:127-166).  Consecutive threads read consecutive
// values of one row, so every load is coalesced.
__global__ __launch_bounds__(256) void k_weighted_reduce(
    int64_t total, int group, int64_t width, const double *in,
    const double *w, int normalise, double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t o = i / width, k = i - o * width;
  const double *wg = w + o * group;
  const double *x = in + o * group * width + k;
  double acc = 0., norm = 0.;
  for (int s = 0; s < group; ++s) {
    acc += wg[s] * x[(int64_t)s * width];
    norm += wg[s];
  }
  out[i] = normalise ? acc * (1. / norm) : acc;
}

hipError_t launch_weighted_reduce(int n_groups, int group, int64_t width,
                                  const double *in, const double *w,
                                  int normalise, double *out, hipStream_t s) {
  const int64_t total = (int64_t)n_groups * width;
  if (total <= 0) return hipSuccess;
  const int64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(k_weighted_reduce, dim3((unsigned)blocks), dim3(256), 0,
                     s, total, group, width, in, w, normalise, out);
  return hipGetLastError();
}

hipError_t launch_cg(const CgKey *keys, int n, double *cg, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cg, dim3(n), dim3(64), 0, s, keys, cg);
  return hipGetLastError();
}

hipError_t launch_average(const AvgTask *tasks, int n, const Item *items,
                          const int32_t *gi_off, const cdbl *tbuf,
                          const double *pN, int ncoef, const int32_t *dist,
                          const int32_t *cg_index, const double *cg,
                          int nmax_cg, cdbl *abuf, const int64_t *aoff,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t lds = (size_t)(2 * nmax_cg + 1) * 2 * sizeof(cdbl);
  hipLaunchKernelGGL(k_average, dim3(n), dim3(64), lds, s, tasks, items,
                     gi_off, tbuf, pN, ncoef, dist, cg_index, cg, nmax_cg,
                     abuf, aoff);
  return hipGetLastError();
}

}  // namespace ctm
