// ctm_average.hpp -- orientation-average kernels (ensemble T-matrix).
#pragma once

#include "ctm_device.hpp"

namespace ctm {

// Largest order the averaging kernels handle (LDS rows of 2 N + 1).
constexpr int kMaxAvgOrder = 400;

struct CgKey {
  int32_t n1, n2, N, pad;
  int64_t off;  // offset of the min(n1, n2) + 1 values in the CG table
};

struct AvgTask {
  int32_t grain, n1, n2, pad;
};

hipError_t launch_cg(const CgKey *keys, int n, double *cg, hipStream_t s);
hipError_t launch_average(const AvgTask *tasks, int n, const Item *items,
                          const int32_t *gi_off, const cdbl *tbuf,
                          const double *pN, int ncoef, const int32_t *dist,
                          const int32_t *cg_index, const double *cg,
                          int nmax_cg, cdbl *abuf, const int64_t *aoff,
                          hipStream_t s);

// a[0 .. n) = 0 (the ensemble entries of pairs no coefficient reaches)
hipError_t launch_zero(cdbl *a, int64_t n, hipStream_t s);

// out[o][k] = sum_s w[o][s] in[o][s][k] (divided by sum_s w[o][s] when
// normalise != 0), s in order: shape averaging and size integration.
hipError_t launch_weighted_reduce(int n_groups, int group, int64_t width,
                                  const double *in, const double *w,
                                  int normalise, double *out, hipStream_t s);

}  // namespace ctm
