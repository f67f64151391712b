// ctm_average.hpp -- orientation-average kernels (ensemble T-matrix).
#pragma once

#include "ctm_device.hpp"

namespace ctm {

// Largest order the averaging kernels handle (LDS rows of 2 N + 1).
constexpr int kMaxAvgOrder = 200;

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

}  // namespace ctm
