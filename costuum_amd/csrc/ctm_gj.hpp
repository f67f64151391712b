// ctm_gj.hpp -- the panel steps of the blocked Gauss-Jordan inversion,
// shared by the FP64 solve (ctm_kernels.hip) and the FP32 elimination of the
// mixed-precision path (ctm_mixed.hip).  E is the complex element type in
// LDS: cdbl or cflt.
#pragma once

#include "ctm_device.hpp"

namespace ctm {

// Row stride (elements) of an LDS sub-block of n rows: n rounded up to odd.
__host__ __device__ inline int blk_lds_stride(int n) { return n | 1; }

template <typename E>
__device__ __forceinline__ E ezero();
template <>
__device__ __forceinline__ cdbl ezero<cdbl>() {
  return cmk(0., 0.);
}
template <>
__device__ __forceinline__ cflt ezero<cflt>() {
  return cmkf(0.f, 0.f);
}

// K4 for orders 49..99 with the matrix in LDS: blocked Gauss-Jordan.  The
// columns of each sub-block are eliminated in panels of W.  One wave holds a
// panel in registers (RS rows per lane) and runs the W elimination steps on
// it alone -- pivot reduction, row exchange and update need no barrier, and
// it records the pivot rows.  The panel's W steps compose to one
// transformation that is the identity outside the panel's columns, whose
// panel columns are the eliminated panel itself (F); every other column j
// of the sub-block then becomes, after the panel's row exchanges (sigma),
//   A(i, j) = [i not a panel row] A(sigma i, j) + sum_q F(i, q) A(sigma q, j),
// one rank-W update per column, done by one wave per column (loads of the
// column precede its stores in the wave's order, so no barrier).  Look-ahead:
// the wave that eliminates the next panel first applies this update to that
// panel's columns, while the other waves of the sub-block update the rest;
// one barrier separates panels.  Same pivot rule as k_invert_solve_reg (largest
// |a_ik|^2 over rows i >= k, first row on ties); results equal it to
// rounding.  Each sub-block is stored compactly (sz x sz, one after the
// other), so an m = 0 block holds half the LDS of its full square.
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) |
                              (long long)(unsigned)lo);
}

// Lane l's value of a complex element (uniform result).
__device__ __forceinline__ cdbl readlane_e(cdbl v, int l) {
  return cmk(readlane_f64(v.x, l), readlane_f64(v.y, l));
}
__device__ __forceinline__ cflt readlane_e(cflt v, int l) {
  return cmkf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), l)),
              __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), l)));
}

// Pivot key: the bits of |a|^2 (non-negative, so ordered as integers) plus
// one; 0 marks rows that cannot pivot.
__device__ __forceinline__ unsigned long long pivot_key(cdbl v) {
  const double mag = v.x * v.x + v.y * v.y;
  return (unsigned long long)__double_as_longlong(mag) + 1ull;
}
__device__ __forceinline__ unsigned long long pivot_key(cflt v) {
  const float mag = v.x * v.x + v.y * v.y;
  return (unsigned long long)__float_as_uint(mag) + 1ull;
}

// Wave maximum of a 64-bit key: DPP steps within each row of 16 lanes
// (quad swaps, half-row and row mirrors), then the four rows' maxima by
// readlane.  The result is uniform.
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_max_u64(
    unsigned long long v) {
  const int lo = (int)(unsigned)v, hi = (int)(v >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
  const unsigned long long o =
      ((unsigned long long)(unsigned)hi2 << 32) | (unsigned)lo2;
  return o > v ? o : v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(
    unsigned long long v) {
  v = dpp_max_u64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v = dpp_max_u64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v = dpp_max_u64<0x141>(v);  // row_half_mirror
  v = dpp_max_u64<0x140>(v);  // row_mirror
  const int lo = (int)(unsigned)v, hi = (int)(v >> 32);
  unsigned long long m = 0ull;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const unsigned long long x =
        ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(hi, 16 * r)
         << 32) |
        (unsigned)__builtin_amdgcn_readlane(lo, 16 * r);
    m = x > m ? x : m;
  }
  return m;
}

template <int RS, int W, typename E>
__device__ __forceinline__ void blk_panel(E *Ab, int ld, int lo, int hi,
                                          int c0, int w, int lane, int *piv,
                                          int *steps) {
  E P[RS][W];
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int i = lo + lane + 64 * r;
#pragma unroll
    for (int j = 0; j < W; ++j)
      P[r][j] = (i < hi && j < w) ? Ab[(int64_t)(i - lo) * ld + (c0 - lo) + j]
                                  : ezero<E>();
  }
#pragma unroll
  for (int s = 0; s < W; ++s) {
    if (s >= w) break;
    const int k = c0 + s;
    // pivot: key = bits(|a|^2) + 1 orders non-negative magnitudes; 0 = none
    unsigned long long key[RS], mx = 0ull;
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int i = lo + lane + 64 * r;
      key[r] = (i >= k && i < hi) ? pivot_key(P[r][s]) : 0ull;
      mx = key[r] > mx ? key[r] : mx;
    }
    mx = wave_max_u64(mx);
    int p = k;
#pragma unroll
    for (int r = RS - 1; r >= 0; --r) {
      const unsigned long long bal = __ballot(key[r] == mx);
      if (bal) p = lo + 64 * r + __builtin_ctzll(bal);
    }
    p = __builtin_amdgcn_readfirstlane(p);
    const int rk = k - lo, rp = p - lo;
    E vk[W], vp[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      E a = P[0][j], b = P[0][j];
#pragma unroll
      for (int r = 1; r < RS; ++r) {
        if ((rk >> 6) == r) a = P[r][j];
        if ((rp >> 6) == r) b = P[r][j];
      }
      vk[j] = readlane_e(a, rk & 63);
      vp[j] = readlane_e(b, rp & 63);
    }
    if (lane == 0) {
      piv[k] = p;
      steps[s] = p;
    }
    const E inv = crcp(vp[s]);
    E g[W];
#pragma unroll
    for (int j = 0; j < W; ++j) g[j] = cmul(vp[j], inv);
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int i = lo + lane + 64 * r;
      if (i == k) {
#pragma unroll
        for (int j = 0; j < W; ++j) P[r][j] = (j == s) ? inv : g[j];
      } else {
        // row p now holds the old row k
        if (i == p) {
#pragma unroll
          for (int j = 0; j < W; ++j) P[r][j] = vk[j];
        }
        const E ci = P[r][s];
#pragma unroll
        for (int j = 0; j < W; ++j) {
          if (j == s) {
            E v = cmul(ci, inv);
            v.x = -v.x;
            v.y = -v.y;
            P[r][j] = v;
          } else {
            P[r][j].x = P[r][j].x - (ci.x * g[j].x - ci.y * g[j].y);
            P[r][j].y = P[r][j].y - (ci.x * g[j].y + ci.y * g[j].x);
          }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int i = lo + lane + 64 * r;
    if (i >= hi) continue;
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (j < w) Ab[(int64_t)(i - lo) * ld + (c0 - lo) + j] = P[r][j];
  }
}

// The rank-W update of columns j = jbeg, jbeg + jstep, ... < jend of the
// sub-block [lo, hi), skipping [skip0, skip1), by one wave, with the panel
// at columns c0..c0+w-1 and its pivot rows steps[0..w).
template <int RS, int W, typename E>
__device__ __forceinline__ void blk_update(E *Ab, int ld, int lo, int hi,
                                           int c0, int w, const int *steps,
                                           int jbeg, int jend, int jstep,
                                           int skip0, int skip1, int lane) {
  int sg[RS];
  E F[RS][W];
  int sq[W];
#pragma unroll
  for (int q = 0; q < W; ++q) {
    int pos = c0 + q;
    for (int s = w - 1; s >= 0; --s) {
      const int k = c0 + s, p = steps[s];
      pos = pos == k ? p : (pos == p ? k : pos);
    }
    sq[q] = q < w ? pos : c0;
  }
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int i = lo + lane + 64 * r;
    int pos = i;
    for (int s = w - 1; s >= 0; --s) {
      const int k = c0 + s, p = steps[s];
      pos = pos == k ? p : (pos == p ? k : pos);
    }
    sg[r] = i < hi ? pos : lo;
#pragma unroll
    for (int q = 0; q < W; ++q)
      F[r][q] = (i < hi && q < w) ? Ab[(int64_t)(i - lo) * ld + (c0 - lo) + q]
                                  : ezero<E>();
  }
  for (int j = jbeg; j < jend; j += jstep) {
    if (j >= skip0 && j < skip1) continue;
    E po[W];
#pragma unroll
    for (int q = 0; q < W; ++q) po[q] = Ab[(int64_t)(sq[q] - lo) * ld + (j - lo)];
    E acc[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int i = lo + lane + 64 * r;
      const bool prow = i >= c0 && i < c0 + w;
      acc[r] = prow ? ezero<E>() : Ab[(int64_t)(sg[r] - lo) * ld + (j - lo)];
#pragma unroll
      for (int q = 0; q < W; ++q) cfma(acc[r], F[r][q], po[q]);
    }
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int i = lo + lane + 64 * r;
      if (i < hi) Ab[(int64_t)(i - lo) * ld + (j - lo)] = acc[r];
    }
  }
}


}  // namespace ctm
