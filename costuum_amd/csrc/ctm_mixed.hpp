// ctm_mixed.hpp -- the mixed-precision solve of the all-m phase: FP32
// Gauss-Jordan elimination of the equilibrated Q blocks, then FP64
// iterative refinement of T = -RgQ Q^-1 against the FP64 Q and RgQ.
#pragma once

#include "ctm_device.hpp"

namespace ctm {

// Block orders the mixed solve takes: the smaller blocks stay on the
// register-resident FP64 elimination, the larger ones (global scratch) on
// the FP64 blocked one.
constexpr int kMixedMinOrder = 17;
constexpr int kMixedMaxOrder = 99;

// Refinement sweeps at most (each: P = T Q, R = -RgQ - P, T += R X), the
// predicted error of T (relative to max |T|) at which a task stops, and the
// correction size below which a stalled task is accepted (the FP64 noise
// floor of an ill-conditioned block).  1e-10 in T keeps Q_ext four orders
// inside the mixed path's 1e-6 bar.
constexpr int kRefineMaxIter = 6;
constexpr double kRefineTol = 1e-10;
constexpr double kRefineStall = 1e-9;

size_t invert32_lds_bytes(int64_t store_max, int nm_max);

// FP32 elimination of the tasks' Q_c (item << 1 | class): writes, per task,
// the FP32 inverse X of the equilibrated block D_r Q D_c in the class layout
// (xbuf, at the task's T offset) and the exponents (xexp, 2 nm ints at
// t_off + class 2 nm: the row scale of X by class row, then the column
// scale by class column), so that Q^-1(i, j) ~ 2^(xexp_r[i] + xexp_c[j])
// X(i, j).
hipError_t launch_invert32(const int32_t *tasks, int n, int nm_max,
                           int64_t store_max, int rows_max, const Item *items,
                           const cdbl *qbuf, cflt *xbuf, int32_t *xexp,
                           hipStream_t s, const int32_t *dn = nullptr,
                           int grid = 0);

// FP64 refinement: T = -RgQ X, then sweeps T += (-RgQ - T Q) X until the
// predicted error of T is below kRefineTol of its largest entry.  Writes T
// (tbuf), the per-task partial sums (qparts) and flags[i] = 0 when the
// refinement converged, 1 when the task must be solved in FP64 instead.
// rbuf holds the residuals (T layout).
hipError_t launch_refine(const int32_t *tasks, int n, int nm_max,
                         const Item *items, const cdbl *qbuf,
                         const cflt *xbuf, const int32_t *xexp, cdbl *tbuf,
                         cdbl *rbuf, double *qparts, int32_t *flags,
                         hipStream_t s, const int32_t *dn = nullptr,
                         int grid = 0);

}  // namespace ctm
