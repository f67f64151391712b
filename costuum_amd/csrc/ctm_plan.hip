// ctm_plan.hip -- device list builders of the batch plan (see ctm_plan.hpp).
//
// The host path lists every phase on the CPU after reading the convergence
// flags back (ctm_api.hip: WorkList, build_tasks).  These kernels build the
// same lists from device state, so the phases of a batch follow one another
// without a host round trip: the convergence windows of the next round come
// from the device's convergence state (TaskManager.hpp:1146-1197 keeps that
// loop on the host threads), and the all-m blocks from the converged orders.
#include <hip/hip_runtime.h>

#include "ctm_plan.hpp"

namespace ctm {

namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanChunk = 1024;  // elements per block of the tile scan

// Sum of n^2 for n = 1 .. k (0 for k <= 0).
__device__ __forceinline__ int64_t sum_sq(int64_t k) {
  return k > 0 ? k * (k + 1) * (2 * k + 1) / 6 : 0;
}
__device__ __forceinline__ int64_t sum_n(int64_t k) {
  return k > 0 ? k * (k + 1) / 2 : 0;
}

// Exclusive prefix of v over the workgroup (NT threads, NT / 64 <= 64
// waves); *total receives the sum.  sh: 2 NT / 64 int64 of LDS.
template <int NT>
__device__ int64_t block_scan_excl(int64_t v, int64_t *sh, int64_t *total) {
  constexpr int NW = NT / 64;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  __syncthreads();  // sh may still be read by a previous scan
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (w == 0) {
    int64_t s = lane < NW ? sh[lane] : 0;
    for (int o = 1; o < NW; o <<= 1) {
      const int64_t y = __shfl_up(s, o, 64);
      if (lane >= o) s += y;
    }
    if (lane < NW) sh[NW + lane] = s;
  }
  __syncthreads();
  const int64_t before = w > 0 ? sh[NW + w - 1] : 0;
  if (total) *total = sh[2 * NW - 1];
  return before + x - v;
}

// Histogram area (ints): part bins and cursors by (part, gmax - G), order
// bins with solve and observable cursors by (last - nm), list starts.
struct HistView {
  int32_t *part_bin, *part_cur, *nm_bin, *sol_cur, *obs_cur;
  int32_t *part_start, *bucket_start, *obs_start;
  int nbins_part, nbins_nm;
};
__device__ __forceinline__ HistView hist_view(const PlanLists &L) {
  HistView h;
  h.nbins_part = kPlanParts * (L.gmax + 1);
  h.nbins_nm = L.last_order + 1;
  h.part_bin = L.hist;
  h.part_cur = h.part_bin + h.nbins_part;
  h.nm_bin = h.part_cur + h.nbins_part;
  h.sol_cur = h.nm_bin + h.nbins_nm;
  h.obs_cur = h.sol_cur + h.nbins_nm;
  h.part_start = h.obs_cur + h.nbins_nm;
  h.bucket_start = h.part_start + kPlanParts;
  h.obs_start = h.bucket_start + kPlanBuckets;
  return h;
}

// ---------------------------------------------------------------------------
// Round state.

// First window of every grain and a fresh convergence state
// (TMatrixResource.hpp:112-124); thread 0 also clears the run status.
__global__ void k_plan_first(const PlanLists L, const GrainParams gp) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) {
    PlanStatus z{};
    *L.status = z;
  }
  if (g >= L.n) return;
  ConvState c{};
  c.old_s = c.old_e = 0.;
  c.dqs = c.dqe = -2.;
  c.first = 1;
  c.done = 0;
  c.nmax = L.min_order;
  c.item = -1;
  c.status = 1;
  L.state[g] = c;
  const int end = first_window_end(gp.radius[g], gp.axis_ratio[g],
                                   gp.wavelength[g], gp.mr_re[g],
                                   L.win_scale, L.win_add);
  const int lim = end > L.min_order + 1 ? end : L.min_order + 1;
  L.lo[g] = L.min_order;
  L.hi[g] = lim < L.last_order ? lim : L.last_order;
}

// Next window of every grain still open; an empty one (lo > hi) for the
// converged grains.
__global__ void k_plan_next(const PlanLists L, int final_round) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= L.n) return;
  if (L.state[g].done) {
    L.lo[g] = 1;
    L.hi[g] = 0;
    return;
  }
  const int hi = L.hi[g];
  L.lo[g] = hi + 1;
  L.hi[g] = next_window_end(hi, L.last_order, final_round != 0);
}

// ---------------------------------------------------------------------------
// Listing.

// Per-grain counts (sets, items, nodes, table doubles, Q and T complex
// values) in closed form, their prefix sums, the capacity checks and the
// reset of the phase record and histograms.  One workgroup.
__global__ __launch_bounds__(kScanThreads) void k_plan_scan(
    const PlanLists L, int mode, int phase, int want_x) {
  __shared__ int64_t sh[2 * kScanThreads / 64];
  __shared__ int64_t tot[6];
  __shared__ int skip_sh, xover_sh;
  PlanPhase &P = L.phase[phase];
  const int t = threadIdx.x;
  const int n = L.n, glfac = L.glfac;
  const int per = (n + kScanThreads - 1) / kScanThreads;
  const int g0 = t * per, g1 = min(n, g0 + per);
  auto counts = [&](int g, int64_t c[6]) {
    for (int k = 0; k < 6; ++k) c[k] = 0;
    const ConvState st = L.state[g];
    if (mode == kPlanWindow) {
      const int lo = L.lo[g], hi = L.hi[g];
      if (st.done || lo > hi) return;
      const int64_t s1 = sum_n(hi) - sum_n(lo - 1);
      const int64_t s2 = sum_sq(hi) - sum_sq(lo - 1);
      c[0] = hi - lo + 1;
      c[1] = hi - lo + 1;
      c[2] = glfac * s1;
      c[3] = glfac * ((int64_t)kNodeScalars * s1 + (int64_t)kBesselArrays * s2);
      c[4] = 4 * s2;
      c[5] = 0;
    } else {
      const int64_t N = st.nmax, G = glfac * N;
      c[0] = 1;
      c[1] = N + 1;
      c[2] = G;
      c[3] = (int64_t)kNodeScalars * G + (int64_t)kBesselArrays * N * G;
      c[4] = 4 * (N * N + sum_sq(N));
      c[5] = c[4] / 2;
    }
  };
  int64_t loc[6] = {0, 0, 0, 0, 0, 0};
  for (int g = g0; g < g1; ++g) {
    int64_t c[6];
    counts(g, c);
    for (int k = 0; k < 6; ++k) loc[k] += c[k];
  }
  int64_t base[6];
  for (int k = 0; k < 6; ++k) {
    int64_t total;
    base[k] = block_scan_excl<kScanThreads>(loc[k], sh, &total);
    if (t == 0) tot[k] = total;
  }
  __syncthreads();
  if (t == 0) {
    PlanStatus &S = *L.status;
    int over = 0;
    if (tot[2] > 0x7fffffff) over |= kPlanOverflowNodes;
    if (tot[3] > L.caps.tables) over |= kPlanOverflowTables;
    if (tot[4] > L.caps.q) over |= kPlanOverflowQ;
    if (tot[5] > L.caps.t) over |= kPlanOverflowT;
    const int64_t xneed = want_x ? tot[1] * (int64_t)L.n_pairs * 4 : 0;
    const int xover = xneed > L.caps.x;
    S.overflow |= over | (xover ? kPlanOverflowX : 0);
    S.need_tables = max(S.need_tables, tot[3]);
    S.need_q = max(S.need_q, tot[4]);
    S.need_t = max(S.need_t, tot[5]);
    S.need_x = max(S.need_x, xneed);
    PlanPhase z{};
    z.skip = over != 0;
    z.skip_x = !want_x || xover || z.skip;
    if (!z.skip) {
      z.sets_nodes[0] = (int32_t)tot[0];
      z.sets_nodes[1] = (int32_t)tot[2];
      z.n_items = (int32_t)tot[1];
    }
    z.tables = tot[3];
    z.q = tot[4];
    z.t = tot[5];
    P = z;
    skip_sh = z.skip;
    xover_sh = !want_x || xover || z.skip;
  }
  __syncthreads();
  // offsets per grain (and the first item of each grain)
  int64_t run[6];
  for (int k = 0; k < 6; ++k) run[k] = base[k];
  for (int g = g0; g < g1; ++g) {
    int64_t c[6];
    counts(g, c);
    for (int k = 0; k < 6; ++k) {
      L.goff[(int64_t)g * 6 + k] = run[k];
      run[k] += c[k];
    }
    L.gitem[g] = skip_sh ? 0 : (int32_t)run[1] - (int32_t)c[1];
    L.gobs[g] = xover_sh ? 0 : (int32_t)run[1] - (int32_t)c[1];
  }
  if (t == kScanThreads - 1) {
    for (int k = 0; k < 6; ++k) L.goff[(int64_t)n * 6 + k] = tot[k];
    L.gitem[n] = skip_sh ? 0 : (int32_t)tot[1];
    L.gobs[n] = xover_sh ? 0 : (int32_t)tot[1];
  }
  // histograms and cursors
  const int64_t nh = plan_hist_ints(L.gmax, L.last_order);
  for (int64_t i = t; i < nh; i += kScanThreads) L.hist[i] = 0;
}

// One thread per item: the item (and, for its first order or m = 0, the
// grain's table set), and its (part, G) and order histogram entries,
// counted in LDS per workgroup and added to the global histograms once per
// workgroup and bin (many items share a bin).
constexpr int kHistLds = 2048;  // LDS bins: (part, G) then order
__global__ __launch_bounds__(256) void k_plan_items(const PlanLists L,
                                                    int mode, int phase) {
  __shared__ int32_t lh[kHistLds];
  const PlanPhase &P = L.phase[phase];
  const int n_items = P.n_items;
  const HistView h = hist_view(L);
  const int glfac = L.glfac;
  const int nbp = h.nbins_part, nb = h.nbins_part + h.nbins_nm;
  const bool in_lds = nb <= kHistLds;
  const int t = threadIdx.x;
  for (int i0 = blockIdx.x * blockDim.x; i0 < n_items;
       i0 += gridDim.x * blockDim.x) {
    if (in_lds) {
      for (int k = t; k < nb; k += blockDim.x) lh[k] = 0;
      __syncthreads();
    }
    const int i = i0 + t;
    if (i < n_items) {
      // the grain whose item range holds i
      int lo = 0, hi = L.n - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (L.gitem[mid] <= i) lo = mid; else hi = mid - 1;
      }
      const int g = lo;
      const int j = i - L.gitem[g];
      const int64_t *go = L.goff + (int64_t)g * 6;
      Item it{};
      it.grain = g;
      if (mode == kPlanWindow) {
        const int wlo = L.lo[g];
        const int N = wlo + j;
        it.N = N;
        it.m = 0;
        it.G = glfac * N;
        it.set = (int32_t)(go[0] + j);
        it.q_off = go[4] + 4 * (sum_sq(N - 1) - sum_sq(wlo - 1));
        it.t_off = -1;
        TableSet ts{};
        ts.grain = g;
        ts.N = N;
        ts.G = it.G;
        ts.node0 = (int32_t)(go[2] + glfac * (sum_n(N - 1) - sum_n(wlo - 1)));
        ts.gl_off = L.gl_off[N];
        ts.off = go[3] + glfac * ((int64_t)kNodeScalars *
                                      (sum_n(N - 1) - sum_n(wlo - 1)) +
                                  (int64_t)kBesselArrays *
                                      (sum_sq(N - 1) - sum_sq(wlo - 1)));
        L.sets[it.set] = ts;
      } else {
        const int N = L.state[g].nmax;
        const int m = j;
        it.N = N;
        it.m = m;
        it.G = glfac * N;
        it.set = (int32_t)go[0];
        const int64_t ql =
            m == 0 ? 0 : 4 * ((int64_t)N * N + sum_sq(N) - sum_sq(N + 1 - m));
        it.q_off = go[4] + ql;
        it.t_off = go[5] + ql / 2;
        if (m == 0) {
          TableSet ts{};
          ts.grain = g;
          ts.N = N;
          ts.G = it.G;
          ts.node0 = (int32_t)go[2];
          ts.gl_off = L.gl_off[N];
          ts.off = go[3];
          L.sets[it.set] = ts;
        }
      }
      it.wig = L.wig_base[it.N] + it.m;
      L.items[i] = it;
      const int nm = it.N + 1 - (it.m > 0 ? it.m : 1);
      const int part = block_part(nm, it.m == 0, L.micro_ratio);
      const int bp = part * (L.gmax + 1) + (L.gmax - it.G);
      const int bn = L.last_order - nm;
      if (in_lds) {
        atomicAdd(&lh[bp], 1);
        atomicAdd(&lh[nbp + bn], 1);
      } else {
        atomicAdd(&h.part_bin[bp], 1);
        atomicAdd(&h.nm_bin[bn], 1);
      }
    }
    if (in_lds) {
      __syncthreads();
      for (int k = t; k < nb; k += blockDim.x)
        if (lh[k]) atomicAdd(k < nbp ? &h.part_bin[k] : &h.nm_bin[k - nbp], lh[k]);
      __syncthreads();
    }
  }
}

// One workgroup: prefix sums of the histograms (cursors), the item count
// of every part, solve bucket and observable bucket, and their starts.
__global__ __launch_bounds__(kScanThreads) void k_plan_bins(
    const PlanLists L, int phase, int want_x) {
  __shared__ int64_t sh[2 * kScanThreads / 64];
  PlanPhase &P = L.phase[phase];
  const HistView h = hist_view(L);
  const int t = threadIdx.x;
  // (part, G) bins: up to four per thread
  {
    const int nb = h.nbins_part;
    const int per = (nb + kScanThreads - 1) / kScanThreads;
    const int b0 = t * per, b1 = min(nb, b0 + per);
    int64_t s = 0;
    for (int b = b0; b < b1; ++b) s += h.part_bin[b];
    int64_t run = block_scan_excl<kScanThreads>(s, sh, nullptr);
    for (int b = b0; b < b1; ++b) {
      h.part_cur[b] = (int32_t)run;
      run += h.part_bin[b];
    }
  }
  // order bins
  {
    const int nb = h.nbins_nm;
    const int per = (nb + kScanThreads - 1) / kScanThreads;
    const int b0 = t * per, b1 = min(nb, b0 + per);
    int64_t s = 0;
    for (int b = b0; b < b1; ++b) s += h.nm_bin[b];
    int64_t run = block_scan_excl<kScanThreads>(s, sh, nullptr);
    for (int b = b0; b < b1; ++b) {
      h.sol_cur[b] = (int32_t)run;
      h.obs_cur[b] = (int32_t)run;
      run += h.nm_bin[b];
    }
  }
  __syncthreads();
  if (t == 0) {
    const int last = L.last_order;
    const int n_items = P.n_items;
    // part p's items start at its first bin's cursor
    for (int p = 0; p < kPlanParts; ++p) {
      const int first = h.part_cur[p * (L.gmax + 1)];
      const int next = p + 1 < kPlanParts ? h.part_cur[(p + 1) * (L.gmax + 1)]
                                          : n_items;
      h.part_start[p] = first;
      P.part[p][0] = next - first;
      P.part[p][1] = 0;
    }
    // items of order bins k0 .. k1 (k = last - nm)
    auto range = [&](int nm_hi, int nm_lo, int *start) {
      const int k0 = max(0, last - nm_hi);
      const int k1 = last - nm_lo;  // exclusive
      if (k0 >= k1 || k0 > last) {
        *start = 0;
        return 0;
      }
      const int a = h.sol_cur[k0];
      const int b = k1 <= last ? h.sol_cur[k1] : n_items;
      *start = a;
      return b - a;
    };
    for (int b = 0; b < kPlanBuckets; ++b) {
      const int hi_b = plan_bucket_hi(b);
      const int lo_b = b + 1 < kPlanBuckets ? plan_bucket_hi(b + 1) : 0;
      int st;
      const int c = range(hi_b, lo_b, &st);
      h.bucket_start[b] = st;
      P.bucket[b][0] = 2 * c;
      P.bucket[b][1] = 0;
      P.fallback[b][0] = 0;
      P.fallback[b][1] = 0;
    }
    for (int b = 0; b < kPlanObsBuckets; ++b) {
      const int hi_b = plan_obs_hi(b, last);
      const int lo_b =
          b + 1 < kPlanObsBuckets ? plan_obs_hi(b + 1, last) : 0;
      int st;
      const int c = hi_b > lo_b ? range(hi_b, lo_b, &st) : (st = 0, 0);
      h.obs_start[b] = st;
      P.obs[b][0] = P.skip_x ? 0 : c;
      P.obs[b][1] = 0;
    }
  }
}

// One thread per item: its place in its part's assembly order (G
// descending), its two tasks in its solve bucket (order descending) and,
// for the observables, its place in its observable bucket; the tiles of
// the tiled parts (for the prefix sums) and the algorithmic work per
// launch.  Places are reserved per workgroup: the items of a workgroup
// count themselves per bin in LDS (their rank within the bin), one global
// atomic per bin reserves the workgroup's range, and each item takes base
// + rank; the work sums go the same way.
__global__ __launch_bounds__(256) void k_plan_scatter(const PlanLists L,
                                                      int mode, int phase,
                                                      int want_x,
                                                      int32_t *tiles) {
  constexpr int NA = 2 * kPlanParts + 2 * kPlanBuckets + kPlanObsBuckets;
  __shared__ double acc[NA];
  __shared__ int32_t lc[kHistLds];  // counts, then bases, per bin
  PlanPhase &P = L.phase[phase];
  const int n_items = P.n_items;
  const HistView h = hist_view(L);
  const int64_t cap = L.caps.items;
  const int nbp = h.nbins_part, nb = h.nbins_part + h.nbins_nm;
  const bool in_lds = nb <= kHistLds;
  const bool obs = want_x && mode == kPlanAllM;
  const int t = threadIdx.x;
  for (int i0 = blockIdx.x * blockDim.x; i0 < n_items;
       i0 += gridDim.x * blockDim.x) {
    if (t < NA) acc[t] = 0.;
    if (in_lds)
      for (int k = t; k < nb; k += blockDim.x) lc[k] = 0;
    __syncthreads();
    const int i = i0 + t;
    const bool live = i < n_items;
    Item it{};
    int nm = 0, part = 0, bp = 0, bn = 0, rp = 0, rn = 0;
    if (live) {
      it = L.items[i];
      nm = it.N + 1 - (it.m > 0 ? it.m : 1);
      part = block_part(nm, it.m == 0, L.micro_ratio);
      bp = part * (L.gmax + 1) + (L.gmax - it.G);
      bn = L.last_order - nm;
      if (in_lds) {
        rp = atomicAdd(&lc[bp], 1);
        rn = atomicAdd(&lc[nbp + bn], 1);
      }
    }
    __syncthreads();
    if (in_lds) {
      // reserve each bin's range (an item's place among the observable
      // lists is its place among the solve lists: the same order bins)
      for (int k = t; k < nb; k += blockDim.x) {
        const int c = lc[k];
        if (c == 0) continue;
        if (k < nbp) {
          lc[k] = atomicAdd(&h.part_cur[k], c);
        } else {
          lc[k] = atomicAdd(&h.sol_cur[k - nbp], c);
        }
      }
      __syncthreads();
    }
    if (live) {
      const bool m0 = it.m == 0;
      const double d = nm;
      const int pos = in_lds ? lc[bp] + rp : atomicAdd(&h.part_cur[bp], 1);
      const int lp = pos - h.part_start[part];
      L.ord[part * cap + lp] = i;
      if (part < 3) {
        const int tr = 16 * (3 - part);
        const int ntr = (nm + tr - 1) / tr;
        tiles[part * cap + lp] = ntr * (m0 ? (nm + 2 * tr - 1) / (2 * tr)
                                            : ntr);
      }
      atomicAdd(&acc[part], (m0 ? 32. : 96.) * it.G * d * d);
      atomicAdd(&acc[kPlanParts + part], 16. * d * it.G + 64. * d * d);
      const int posn = in_lds ? lc[nbp + bn] + rn : atomicAdd(&h.sol_cur[bn], 1);
      const int b = plan_bucket_of(nm);
      const int pos2 = posn - h.bucket_start[b];
      int32_t *sl = L.sol + (int64_t)b * 2 * cap;
      sl[2 * pos2] = 2 * i;
      sl[2 * pos2 + 1] = 2 * i + 1;
      atomicAdd(&acc[2 * kPlanParts + b],
                2. * (m0 ? 8. / 3. : 32. / 3.) * d * d * d);
      atomicAdd(&acc[2 * kPlanParts + kPlanBuckets + b],
                2. * (32. * d * d + (mode == kPlanAllM ? 16. * d * d : 0.)));
      if (obs) {
        // the observable cursors of a bin equal its solve cursors
        const int pos3 = (in_lds ? posn : atomicAdd(&h.obs_cur[bn], 1));
        const int ob = plan_obs_of(nm);
        L.obsl[ob * cap + pos3 - h.obs_start[ob]] = i;
        atomicAdd(&acc[2 * kPlanParts + 2 * kPlanBuckets + ob],
                  8. * (4. * d * d * L.n_in + 8. * d * L.n_pairs));
      }
    }
    __syncthreads();
    if (t < NA && acc[t] != 0.) {
      double *dst = t < kPlanParts ? &P.part_flops[t]
                    : t < 2 * kPlanParts ? &P.part_bytes[t - kPlanParts]
                    : t < 2 * kPlanParts + kPlanBuckets
                        ? &P.bucket_flops[t - 2 * kPlanParts]
                    : t < 2 * kPlanParts + 2 * kPlanBuckets
                        ? &P.bucket_bytes[t - 2 * kPlanParts - kPlanBuckets]
                        : &P.obs_flops[t - 2 * kPlanParts - 2 * kPlanBuckets];
      atomicAdd(dst, acc[t]);
    }
    __syncthreads();
  }
}

// Tile prefix sums of the three tiled parts, first pass: the exclusive
// prefix within each chunk of kScanChunk items and the chunk's total.
// Grid: (chunks, 3).
__global__ __launch_bounds__(kScanThreads) void k_plan_tscan1(
    const PlanLists L, int phase, const int32_t *tiles, int32_t *csum) {
  __shared__ int64_t sh[2 * kScanThreads / 64];
  const int p = blockIdx.y;
  const int n = L.phase[phase].part[p][0];
  const int64_t cap = L.caps.items;
  const int c = blockIdx.x;
  if ((int64_t)c * kScanChunk >= n) return;
  const int k = c * kScanChunk + threadIdx.x;
  const int v = k < n ? tiles[p * cap + k] : 0;
  int64_t total;
  const int64_t ex = block_scan_excl<kScanThreads>(v, sh, &total);
  if (k < n) L.pre[p * (cap + 1) + k] = (int32_t)ex;
  if (threadIdx.x == 0) csum[p * ((cap + kScanChunk - 1) / kScanChunk) + c] =
      (int32_t)total;
}

// Second pass: each chunk adds the sum of the chunks before it (scanned
// again per workgroup: at most caps.items / kScanChunk values); the last
// chunk writes the part's tile count.
__global__ __launch_bounds__(kScanThreads) void k_plan_tscan2(
    const PlanLists L, int phase, const int32_t *csum) {
  __shared__ int64_t sh[2 * kScanThreads / 64];
  __shared__ int64_t before_sh;
  PlanPhase &P = L.phase[phase];
  const int p = blockIdx.y;
  const int n = P.part[p][0];
  const int64_t cap = L.caps.items;
  const int64_t ncap = (cap + kScanChunk - 1) / kScanChunk;
  const int c = blockIdx.x;
  const int nch = (n + kScanChunk - 1) / kScanChunk;
  if (c >= nch && !(n == 0 && c == 0)) return;
  const int32_t *cs = csum + p * ncap;
  // sum of chunks 0 .. c - 1
  int64_t s = 0;
  for (int q = threadIdx.x; q < c; q += kScanThreads) s += cs[q];
  int64_t total;
  (void)block_scan_excl<kScanThreads>(s, sh, &total);
  if (threadIdx.x == 0) before_sh = total;
  __syncthreads();
  const int64_t before = before_sh;
  const int k = c * kScanChunk + threadIdx.x;
  int32_t *pre = L.pre + p * (cap + 1);
  if (k < n) pre[k] += (int32_t)before;
  if (c == (nch > 0 ? nch - 1 : 0) && threadIdx.x == 0) {
    const int64_t all = before + (n > 0 ? cs[c] : 0);
    pre[n] = (int32_t)all;
    P.part[p][1] = (int32_t)all;
  }
}

// One thread: the sizes of the phase's lists against the grids the host
// launches for them.  A list larger than its grid would leave units
// unprocessed, so the whole phase is emptied (its launches return at once)
// and the run flagged; the host enlarges the grids and runs again.
__global__ void k_plan_check(const PlanLists L, int phase) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  PlanPhase &P = L.phase[phase];
  int32_t c[kPlanLists];
  c[0] = P.sets_nodes[1];
  for (int p = 0; p < kPlanParts; ++p)
    c[plan_list_part(p)] = p < 3 ? P.part[p][1] : (P.part[p][0] + 3) / 4;
  for (int b = 0; b < kPlanBuckets; ++b) c[plan_list_bucket(b)] = P.bucket[b][0];
  for (int b = 0; b < kPlanObsBuckets; ++b) c[plan_list_obs(b)] = P.obs[b][0];
  const int32_t *cap = L.gcap + phase * kPlanLists;
  bool over = false;
  for (int k = 0; k < kPlanLists; ++k) {
    L.status->count[phase][k] = c[k];
    over = over || c[k] > cap[k];
  }
  if (!over) return;
  L.status->overflow |= kPlanOverflowGrid;
  P.skip = 1;
  P.skip_x = 1;
  P.n_items = 0;
  P.sets_nodes[0] = P.sets_nodes[1] = 0;
  for (int p = 0; p < kPlanParts; ++p) P.part[p][0] = P.part[p][1] = 0;
  for (int b = 0; b < kPlanBuckets; ++b) P.bucket[b][0] = 0;
  for (int b = 0; b < kPlanObsBuckets; ++b) P.obs[b][0] = 0;
}

// ---------------------------------------------------------------------------
// Convergence and the mixed path's fallback.

__global__ void k_plan_converge(const PlanLists L, int phase,
                                const Item *items, const double *qparts,
                                double tol) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= L.n || L.phase[phase].skip) return;
  ConvState st = L.state[g];
  if (st.done) return;
  const int q0 = L.gitem[g], q1 = L.gitem[g + 1];
  if (q0 >= q1) return;
  converge_items(st, q0, q1, nullptr, items, qparts, tol, L.last_order);
  L.state[g] = st;
}

// One workgroup per bucket: the flagged tasks of the bucket, in order.
__global__ __launch_bounds__(kScanThreads) void k_plan_fallback(
    const PlanLists L, int phase) {
  __shared__ int64_t sh[2 * kScanThreads / 64];
  __shared__ int64_t base_sh;
  PlanPhase &P = L.phase[phase];
  const int b = blockIdx.x;
  const int n = P.bucket[b][0];
  const int64_t cap2 = 2 * L.caps.items;
  const int32_t *fl = L.flags + b * cap2;
  const int32_t *sl = L.sol + b * cap2;
  int32_t *fb = L.fb + b * cap2;
  if (threadIdx.x == 0) base_sh = 0;
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += kScanThreads) {
    const int k = k0 + threadIdx.x;
    const int f = k < n && fl[k] ? 1 : 0;
    int64_t total;
    const int64_t ex = block_scan_excl<kScanThreads>(f, sh, &total);
    if (f) fb[base_sh + ex] = sl[k];
    __syncthreads();
    if (threadIdx.x == 0) base_sh += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    P.fallback[b][0] = (int32_t)base_sh;
    atomicAdd((unsigned long long *)&L.status->n_mixed,
              (unsigned long long)n);
    atomicAdd((unsigned long long *)&L.status->n_fallback,
              (unsigned long long)base_sh);
  }
}

__global__ void k_plan_iota(int32_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

}  // namespace

hipError_t launch_plan_first(const PlanLists &L, const GrainParams &gp,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_plan_first, dim3((L.n + 255) / 256), dim3(256), 0, s,
                     L, gp);
  return hipGetLastError();
}

hipError_t launch_plan_next(const PlanLists &L, bool final_round,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_plan_next, dim3((L.n + 255) / 256), dim3(256), 0, s,
                     L, final_round ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_plan_build(const PlanLists &L, int mode, int phase,
                             bool want_x, int grid, hipStream_t s) {
  // the tile counts per part and the chunk sums of their prefix sums sit
  // behind the fallback lists' room (unused while listing)
  int32_t *tiles = L.fb;
  const int64_t nchunk = (L.caps.items + kScanChunk - 1) / kScanChunk;
  int32_t *csum = L.fb + 3 * L.caps.items;
  hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(kScanThreads), 0, s, L, mode,
                     phase, want_x ? 1 : 0);
  hipLaunchKernelGGL(k_plan_items, dim3(grid), dim3(256), 0, s, L, mode,
                     phase);
  hipLaunchKernelGGL(k_plan_bins, dim3(1), dim3(kScanThreads), 0, s, L,
                     phase, want_x ? 1 : 0);
  hipLaunchKernelGGL(k_plan_scatter, dim3(grid), dim3(256), 0, s, L, mode,
                     phase, want_x ? 1 : 0, tiles);
  hipLaunchKernelGGL(k_plan_tscan1, dim3((unsigned)nchunk, 3),
                     dim3(kScanThreads), 0, s, L, phase, tiles, csum);
  hipLaunchKernelGGL(k_plan_tscan2, dim3((unsigned)nchunk, 3),
                     dim3(kScanThreads), 0, s, L, phase, csum);
  hipLaunchKernelGGL(k_plan_check, dim3(1), dim3(64), 0, s, L, phase);
  return hipGetLastError();
}

hipError_t launch_plan_converge(const PlanLists &L, int phase,
                                const double *qparts, double tol,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_plan_converge, dim3((L.n + 127) / 128), dim3(128), 0,
                     s, L, phase, L.items, qparts, tol);
  return hipGetLastError();
}

hipError_t launch_plan_fallback(const PlanLists &L, int phase, int n_buckets,
                                hipStream_t s) {
  if (n_buckets <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_plan_fallback, dim3(n_buckets), dim3(kScanThreads), 0,
                     s, L, phase);
  return hipGetLastError();
}

hipError_t launch_plan_iota(int32_t *p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_plan_iota, dim3(1024), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace ctm
