// ctm_tangent.hpp -- forward-mode derivatives of the T-matrices with
// respect to the grain parameters (R_V, d, Re m, Im m).
#pragma once

#include "ctm_device.hpp"
#include "ctm_kernels.hpp"

namespace ctm {

// Parameter directions of the tangents.
constexpr int kTangentDirs = 4;  // R_V, d, Re m, Im m
// Tangent assembly tiles: 16 rows x kTangentTileCols columns of pairs for
// m > 0, twice as wide for m = 0.
constexpr int kTangentTileCols = 32;

// Tangent tables: direction dir of the set at base offset off starts at
// ttables[dir * 2 * table_size + 2 * off] (2 * kTangentDirs * table_size
// doubles in all).
hipError_t launch_tables_tangent(const TableSet *sets, int n_sets,
                                 const GrainParams &gp, const double *gl,
                                 const double *tables, int64_t table_size,
                                 double *ttables, hipStream_t s);
// dQ, dRgQ per direction: dq[dir * q_size + layout of Q].  order /
// tile_prefix: every item with 16 x kTangentTileCols pair tiles (twice as
// wide for m = 0).
hipError_t launch_assemble_tangent(const int32_t *order,
                                   const int32_t *tile_prefix, int n_order,
                                   int n_tiles, const Item *items,
                                   const TableSet *sets, const WigKey *wkeys,
                                   const GrainParams &gp, const double *tables,
                                   const double *ttables, int64_t table_size,
                                   const double *wig, cdbl *dq, int64_t q_size,
                                   hipStream_t s);
// M = dRgQ + T dQ (in place of dRgQ), then dT = -M Q^-1 and the per-block
// tangent sums tparts[(task * 4 + dir) * 2 + {0: d sum |T|^2, 1: d tr Re T}].
hipError_t launch_tangent_solve(int n_tasks, const Item *items,
                                const cdbl *tbuf, const cdbl *ainv, cdbl *dq,
                                int64_t q_size, cdbl *dt, int64_t t_size,
                                double *tparts, hipStream_t s);
// Directional derivatives of observables quadratic in T by the exact
// symmetric difference O(T + s dT) - O(T - s dT) = 2 s dO: per-grain scales
// s = max |T| / max |dT| (gi_off: the grains' item ranges), T +- s dT into
// `out` (T layout), and (plus - minus) / (2 s) per row of `width` values.
hipError_t launch_pm_scale(int n_grains, const int32_t *gi_off,
                           const Item *items, const cdbl *tbuf, const cdbl *dt,
                           double *scale, hipStream_t s);
hipError_t launch_pm_form(int n_items, const Item *items, const cdbl *tbuf,
                          const cdbl *dt, const double *scale, double sign,
                          cdbl *out, hipStream_t s);
hipError_t launch_pm_combine(int64_t n, int64_t width, const double *plus,
                             const double *minus, const double *scale,
                             double *out, hipStream_t s);
// Per-grain tangents of the full-T sums (f_m = 1 for m = 0, else 2).
hipError_t launch_tangent_sums(int n_grains, const int32_t *gi_off,
                               const double *tparts, double *dqsca,
                               double *dqext, hipStream_t s);

}  // namespace ctm
