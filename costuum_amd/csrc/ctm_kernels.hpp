// ctm_kernels.hpp -- launch wrappers of the gfx950 T-matrix kernels.
#pragma once

#include "ctm_device.hpp"

namespace ctm {

// Per-grain inputs (device arrays).
struct GrainParams {
  const double *radius;
  const double *axis_ratio;
  const double *wavelength;
  const double *mr_re;
  const double *mr_im;
};

hipError_t launch_gauss_legendre(const GlKey *keys, int n, double *gl,
                                 hipStream_t s);
hipError_t launch_wigner(const WigKey *keys, int n, const double *gl,
                         double *wig, hipStream_t s);
// Device counts: every launcher of a per-item, per-tile or per-task kernel
// takes an optional `dn` (device pointer to the list's counts, as each
// kernel documents) and a grid size `grid` that bounds the list (the
// caller guarantees the count fits); the counts passed by value are then
// ignored and the workgroups (threads for k_tables) beyond the list return
// at once.  Lists built on the device run without the host knowing their
// sizes.
//
// sets: the launch's sets, whose node offsets start at node_base; n_nodes:
// the nodes of those sets (dn: [n_sets, n_nodes]).
hipError_t launch_tables(const TableSet *sets, int n, int node_base,
                         int n_nodes, const GrainParams &gp,
                         const double *gl, double *tables, hipStream_t s,
                         const int32_t *dn = nullptr, int grid = 0);
// micro = 1: 16 x 16 pair tiles (blocks up to order kSmallBlock);
// micro = 2: 32 x 32 tiles of 2 x 2 pairs per thread; micro = 3: 48 x 48
// tiles of 3 x 3 pairs per thread.
constexpr int kSmallBlock = 16;
// Blocks of order up to kPackRows, four per workgroup (order: their item
// indices, all with m = 0 when m0 is set, else all with m > 0).
constexpr int kPackRows = 24;
// smallest order assembled four per workgroup (orders 9-16 on the one-pair
// kernel's 16-row tiles measured 15.33 -> 15.24 ms per batch step against
// 9; 13, 15, 21 and 25 between or slower)
constexpr int kPackMin = 17;
hipError_t launch_assemble_pack4(const int32_t *order, int n_order, int m0,
                                 const Item *items, const TableSet *sets,
                                 const WigKey *wkeys, const GrainParams &gp,
                                 const double *tables, const double *wig,
                                 cdbl *qbuf, hipStream_t s,
                                 const int32_t *dn = nullptr, int grid = 0);
// Micro-tile of a block of order nm: 4 (four blocks per workgroup on
// kPackRows-row tiles) from kPackMin to kPackRows, 1 below, else 3 (48-row
// tiles) where their padded pair count, at `ratio` of the cost per pair of
// 32-row tiles (measured 0.5: the 3 x 3 micro-tile reads fewer LDS operands
// per pair), is smaller, else 2.  (m = 0 tiles are twice as wide.)  The
// count leaves out that edge tiles split their nodes: counting them as half
// moves orders 49-64 to 48-row tiles, measured slower (15.63 -> 16.56 ms
// per batch step).
constexpr double kMicroRatio = 0.5;
__host__ __device__ inline int assemble_micro(int nm, bool m0,
                                              double ratio = kMicroRatio) {
  if (nm >= kPackMin && nm <= kPackRows) return 4;
  if (nm <= kSmallBlock) return 1;
  const long r32 = (nm + 31) / 32, c32 = m0 ? (nm + 63) / 64 : r32;
  const long r48 = (nm + 47) / 48, c48 = m0 ? (nm + 95) / 96 : r48;
  return ratio * (double)(r48 * c48 * 2304) < (double)(r32 * c32 * 1024) ? 3
                                                                         : 2;
}
hipError_t launch_assemble(int micro, const int32_t *order,
                           const int32_t *tile_prefix, int n_order,
                           int n_tiles, const Item *items,
                           const TableSet *sets, const WigKey *wkeys,
                           const GrainParams &gp, const double *tables,
                           const double *wig, cdbl *qbuf, hipStream_t s,
                           const int32_t *dn = nullptr, int grid = 0);
size_t invert_solve_lds_bytes(int nm);
// Largest block order the LDS-resident inversion holds.
constexpr int kMaxLdsOrder = 99;
// Largest order the library computes (larger blocks are inverted in global
// scratch).
constexpr int kMaxOrder = 200;
// gscratch: n * nm_max^2 complex values, used when nm_max > kMaxLdsOrder;
// ainv_out (may be NULL): Q_c^-1 per task in the T layout.
hipError_t launch_invert_solve(const int32_t *tasks, int n, int nm_max,
                               int64_t store_max, int rows_max,
                               const Item *items, const cdbl *qbuf, cdbl *tbuf,
                               double *qparts, int diag_only, cdbl *gscratch,
                               cdbl *ainv_out, hipStream_t s,
                               const int32_t *dn = nullptr, int grid = 0);
hipError_t launch_converge(const int32_t *active, int n_active,
                           const int32_t *win_off, const int32_t *win_items,
                           const Item *items, const double *qparts, double tol,
                           int last_order, ConvState *state, hipStream_t s);
hipError_t launch_full_sums(int n_grains, const int32_t *off,
                            const int32_t *gi, const double *qparts,
                            double *qsca, double *qext, hipStream_t s);
hipError_t launch_write_orders(int n_grains, const ConvState *state,
                               uint32_t *nmax_out, int32_t *status_out,
                               hipStream_t s);

}  // namespace ctm
