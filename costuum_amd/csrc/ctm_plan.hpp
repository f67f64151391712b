// ctm_plan.hpp -- device-built work lists of the batch plan: the
// convergence windows, all-m blocks and observables of a batch listed on
// the GPU (counting sorts and prefix sums), so a whole batch step runs
// without the host reading anything back between its phases.
//
// A phase (one window round, or the all-m blocks) is listed by
//   k_plan_scan    one workgroup: per-grain counts in closed form (table
//                  sets, items, nodes, table, Q and T sizes), their prefix
//                  sums, the capacity checks, histogram reset
//   k_plan_items   per item: the item (and its table set), histograms by
//                  (part, G) and by block order
//   k_plan_bins    one workgroup: histogram prefix sums, list counts
//   k_plan_scatter per item: assembly order per part (G descending), solve
//                  tasks per order bucket (order descending), observable
//                  lists, tile counts; algorithmic work per launch
//   k_plan_tscan1/2  tile prefix sums of the tiled parts (chunk scans, then
//                  the chunk offsets)
//   k_plan_check   one thread: list sizes against the launch grids
// and consumed by the kernels' device-count launches (dn pointers into
// PlanPhase).  Lists of one phase are ordered like the host's
// (ctm_api.hip build_tasks); only the order within equal keys differs,
// which changes no result.
#pragma once

#include "ctm_device.hpp"
#include "ctm_kernels.hpp"

namespace ctm {

// Rough order estimate from the largest size parameter of the spheroid
// (TMatrixCalculator.hpp:83-85 uses x + 4.05 x^(1/3) of the equal-volume
// sphere); only used to size the first window of orders.
__host__ __device__ inline int order_estimate(double rv, double d,
                                              double lam) {
  const double c = cbrt(d);
  const double rmax = rv * (c > 1. / (c * c) ? c : 1. / (c * c));
  const double x = 2. * M_PI * rmax / lam;
  return (int)(x + 4.05 * cbrt(x)) + 2;
}

// Last order of a grain's first window.  Small grains (estimate below 10)
// take the estimate + 3 + estimate / 8.  Larger ones take a model of the
// converged order in the size parameter x of the largest semi-axis, Re m
// and the shape: a least-squares fit of the converged orders of a seeded
// 4096-grain spheroid batch (tolerance 1e-4, 2 nodes per order; x up to 63,
// Re m 1.3-2, d 0.5-2).  Its residual (converged minus model, 5th-95th
// percentile) is -1.9 .. 1.5 orders on an independent 1024-grain batch and
// -3.5 .. 2.4 on a 61 440-grain size x wavelength x shape grid (x up to
// 126), against -6 .. 5 for the estimate alone; the window ends 3 + 4 %
// above the model.  Only the amount of work depends on it: the
// convergence decision is made order by order as the reference makes it.
// (scale, add: the margin above the model; the plan's autotuner may choose
// another.)
constexpr double kWindowScale = 1.04, kWindowAdd = 3.;
__host__ __device__ inline int first_window_end(double rv, double d,
                                                double lam, double mre,
                                                double scale = kWindowScale,
                                                double add = kWindowAdd) {
  const int est = order_estimate(rv, d, lam);
  if (est < 10) return est + 3 + est / 8;
  const double c = cbrt(d);
  const double rmax = rv * (c > 1. / (c * c) ? c : 1. / (c * c));
  const double x = 2. * M_PI * rmax / lam;
  const double cx = cbrt(x), ld = fabs(log(d));
  const double model = 0.2213 + 0.6409 * x + 2.4758 * cx +
                       0.4839 * (mre - 1.) * x + 0.7758 * ld * x -
                       0.1053 * ld * cx;
  return (int)ceil(scale * model + add);
}

// Window of the next convergence round of a grain that has not converged
// after window [lo, hi]: half its order more (a later round costs a set of
// small launches, so its extra orders are cheap), or every order up to the
// last one in the final round.
__host__ __device__ inline int next_window_end(int hi, int last_order,
                                               bool final_round) {
  if (final_round) return last_order;
  const int width = (hi + 1) / 2 > 8 ? (hi + 1) / 2 : 8;
  return hi + width < last_order ? hi + width : last_order;
}

// Assembly part of a block (micro-tile 3, 2, 1 -> 0, 1, 2; four per
// workgroup -> 3 for m > 0, 4 for m = 0).
__host__ __device__ inline int block_part(int nm, bool m0,
                                          double ratio = kMicroRatio) {
  const int micro = assemble_micro(nm, m0, ratio);
  return micro == 4 ? (m0 ? 4 : 3) : 3 - micro;
}

constexpr int kPlanParts = 5;
// Solve buckets of the plan (blocks up to kMaxLdsOrder): orders (70, 99],
// (64, 70], (48, 64], (32, 48], (16, 32], (0, 16] -- the host path's
// buckets below the global-scratch one.
constexpr int kPlanBuckets = 6;
__host__ __device__ inline int plan_bucket_hi(int b) {
  return b == 0 ? 99 : b == 1 ? 70 : b == 2 ? 64 : b == 3 ? 48 : b == 4 ? 32
                                                                       : 16;
}
__host__ __device__ inline int plan_bucket_of(int nm) {
  return nm > 70 ? 0 : nm > 64 ? 1 : nm > 48 ? 2 : nm > 32 ? 3 : nm > 16 ? 4
                                                                       : 5;
}
// Observable buckets: orders (80, last], (64, 80], (32, 64], (16, 32],
// (0, 16] (the host path's k_obs_x buckets).
constexpr int kPlanObsBuckets = 5;
__host__ __device__ inline int plan_obs_hi(int b, int last) {
  return b == 0 ? last : b == 1 ? 80 : b == 2 ? 64 : b == 3 ? 32 : 16;
}
__host__ __device__ inline int plan_obs_of(int nm) {
  return nm > 80 ? 0 : nm > 64 ? 1 : nm > 32 ? 2 : nm > 16 ? 3 : 4;
}

// Counts and work of one phase's lists.  The count pairs are the dn
// arguments of the consuming launches.
struct PlanPhase {
  int32_t sets_nodes[2];               // k_tables: [sets, nodes]
  int32_t part[kPlanParts][2];         // assembly: [items, tiles]
  int32_t bucket[kPlanBuckets][2];     // solve tasks
  int32_t fallback[kPlanBuckets][2];   // mixed path: tasks re-solved in FP64
  int32_t obs[kPlanObsBuckets][2];     // k_obs_x items
  int32_t n_items, skip;               // skip: a pool overflowed, no lists
  int32_t skip_x, pad;                 // skip_x: no observable lists
  int64_t tables, q, t;                // sizes (doubles, complex, complex)
  double part_flops[kPlanParts], part_bytes[kPlanParts];
  double bucket_flops[kPlanBuckets], bucket_bytes[kPlanBuckets];
  double obs_flops[kPlanObsBuckets];
  double table_bytes;
};

// The launches of a phase cover their lists with one workgroup per unit
// (one thread per node for the tables), up to a grid bound per (phase,
// list) that the host chooses: list 0 the table nodes, 1 + p the assembly
// part p (tiles for p < 3, groups of four blocks for p = 3, 4), 6 + b the
// solve bucket b (tasks; the mixed path's re-solves of a bucket are fewer),
// 12 + b the observable bucket b (items).
constexpr int kPlanPhases = 4;  // three window rounds, then all m
constexpr int kPlanLists = 1 + kPlanParts + kPlanBuckets + kPlanObsBuckets;
__host__ __device__ inline int plan_list_part(int p) { return 1 + p; }
__host__ __device__ inline int plan_list_bucket(int b) {
  return 1 + kPlanParts + b;
}
__host__ __device__ inline int plan_list_obs(int b) {
  return 1 + kPlanParts + kPlanBuckets + b;
}

// Status of a plan run: pools or grids that were too small (bits
// kPlanOverflow*), the sizes the run needed (largest over its phases) and
// the size of every list.
constexpr int kPlanOverflowTables = 1, kPlanOverflowQ = 2,
              kPlanOverflowT = 4, kPlanOverflowX = 8, kPlanOverflowNodes = 16,
              kPlanOverflowGrid = 32;
struct PlanStatus {
  int32_t overflow, pad;
  int64_t need_tables, need_q, need_t, need_x;
  int64_t n_mixed, n_fallback;
  int32_t count[kPlanPhases][kPlanLists];
};

// Capacities of the data pools (elements).
struct PlanCaps {
  int64_t items;   // items (and table sets) per phase: n (last_order + 1)
  int64_t tables;  // doubles
  int64_t q, t;    // complex values
  int64_t x;       // complex values of the observables' X
};

// Device arrays of the plan's lists.
struct PlanLists {
  TableSet *sets;
  Item *items;
  int32_t *ord;     // [kPlanParts][caps.items] assembly order per part
  int32_t *pre;     // [3][caps.items + 1] tile prefix sums
  int32_t *sol;     // [kPlanBuckets][2 caps.items] solve tasks
  int32_t *obsl;    // [kPlanObsBuckets][caps.items] k_obs_x items
  int32_t *flags;   // [kPlanBuckets][2 caps.items] mixed refinement flags
  int32_t *fb;      // [kPlanBuckets][2 caps.items] FP64 re-solves
  int32_t *hist;    // histograms and cursors (plan_hist_ints)
  int64_t *cnt;     // [n][6] per-grain counts
  int64_t *goff;    // [n + 1][6] per-grain offsets
  int32_t *gitem;   // [n + 1] first item of each grain
  int32_t *gobs;    // [n + 1] first item of each grain for the observables
  int32_t *lo, *hi; // per-grain window
  ConvState *state;
  const int64_t *gl_off;   // by N: Gauss-Legendre offset of G = glfac N
  const int32_t *wig_base; // by N: Wigner key of (N, m = 0); m adds m
  PlanPhase *phase;
  PlanStatus *status;
  const int32_t *gcap;     // [kPlanPhases][kPlanLists] grid bounds
  PlanCaps caps;
  int n, glfac, min_order, last_order;
  int gmax;  // glfac last_order
  int n_in, n_pairs;  // observables (0 without)
  // tunables: first-window margin, tile-shape cost ratio, and whether the
  // largest blocks' solves start before the smaller assembly parts
  double win_scale, win_add, micro_ratio;
  int big_first;
  int obs_lean;  // host-side knob: low-LDS k_obs_x for the largest blocks
};

// Ints of the histogram area: (part, G) bins with their cursors, order
// bins with solve and observable cursors, bucket starts.
__host__ __device__ inline int64_t plan_hist_ints(int gmax, int last) {
  return 2 * (int64_t)kPlanParts * (gmax + 1) + 3 * (int64_t)(last + 1) +
         kPlanParts + kPlanBuckets + kPlanObsBuckets + 16;
}

// Phase listing modes.
constexpr int kPlanWindow = 0, kPlanAllM = 1;

hipError_t launch_plan_first(const PlanLists &L, const GrainParams &gp,
                             hipStream_t s);
hipError_t launch_plan_next(const PlanLists &L, bool final_round,
                            hipStream_t s);
// Lists one phase into L (phase record L.phase[phase]) and checks its
// lists against the grid bounds L.gcap; want_x: the phase feeds the
// observables.
hipError_t launch_plan_build(const PlanLists &L, int mode, int phase,
                             bool want_x, int grid, hipStream_t s);
hipError_t launch_plan_converge(const PlanLists &L, int phase,
                                const double *qparts, double tol,
                                hipStream_t s);
// Mixed path: the tasks of bucket b whose refinement was flagged, into
// L.fb (count in phase.fallback[b]).
hipError_t launch_plan_fallback(const PlanLists &L, int phase, int n_buckets,
                                hipStream_t s);
// Identity list 0 .. n - 1.
hipError_t launch_plan_iota(int32_t *p, int64_t n, hipStream_t s);

}  // namespace ctm
