// ctm_observables.hpp -- direction-pair layout of the observable kernels.
#pragma once

#include "ctm_device.hpp"
#include "ctm_kernels.hpp"

namespace ctm {

// Directions are rows of an angle table (cos, sin, 1/sin with the 9000
// sentinel at sin = 0, ExtinctionCoefficientTask.hpp:222-226).  Incoming
// and outgoing lists index that table; each direction pair (out, in) gets
// the four X_m^k values.  Pair blocks: forward pairs j (extinction, and the
// first term of absorption) at fwd_pair0 + j; absorption pairs (j, i) at
// abs_pair0 + j * ngauss + i; scattering-matrix pairs at sca_pair0 + s.
struct ObsSpec {
  int32_t n_ang, n_in, n_out, n_pairs;
  int32_t n_theta, ngauss, n_stheta, n_sphi;
  int32_t fwd_pair0, abs_pair0, sca_pair0, pad;
  const int32_t *in_angle;   // [n_in]
  const int32_t *out_angle;  // [n_out]
  const int32_t *pair_in;    // [n_pairs] index into the incoming list
  const int32_t *pair_out;   // [n_pairs] index into the outgoing list
  const int32_t *pair_by_in;  // [n_pairs] pairs ordered by incoming index
  const int32_t *in_pair_off; // [n_in + 1] start of each incoming direction
  const double *theta_w;     // [ngauss] GL weights of cos(theta_in)
  const double *phi_cos, *phi_sin, *phi_w;  // [ngauss] azimuths on [0, 2pi]
  const double *sphi_cos, *sphi_sin;        // [n_sphi] scattering azimuths
};

// An S/Z/K query for one grain in an arbitrary orientation (radians).
struct Query {
  int32_t grain, pad;
  double alpha, beta, theta_in, phi_in, theta_out, phi_out;
};
// Per query: S (4 complex, row-major 2x2), Z (16), K (16) doubles.
constexpr int kQueryOut = 8 + 16 + 16;
constexpr int kMaxQueryOrder = 256;

// nmax: the largest order among the queried grains
hipError_t launch_query(const Query *q, int n, int nmax, const Item *items,
                        const int32_t *gi_off, const GrainParams &gp,
                        const cdbl *tbuf, const cdbl *abuf,
                        const int64_t *aoff, double *out, hipStream_t s);

hipError_t launch_pitau(const double *ang, int n_ang, int nmax_all, double *PT,
                        hipStream_t s);
// ck: incoming directions per thread in the T b products (4 or 8); nt:
// threads per workgroup; gcap: most chunks side by side.  Rows of the Y
// buffer: G 2 nm for at most gcap groups that fit nt threads, and at least
// one group of 2 nm_cap rows.
__host__ __device__ inline int obs_x_rows_cap(int nm_cap, int nt, int gcap) {
  const int two = 2 * nm_cap;
  const int fit = gcap * two < nt ? gcap * two : nt;
  return fit > two ? fit : two;
}
// ptg: the outgoing pi/tau are read from the table rather than staged
size_t obs_x_lds_bytes(int nm_cap, int n_out, int ck, int nt, int gcap,
                       bool ptg);
constexpr size_t kObsXMaxLds = 159 * 1024;
// item_list: items to process (largest first); abuf/aoff (may be null):
// full-layout ensemble T-matrices per item.
hipError_t launch_obs_x(const ObsSpec &spec, const int32_t *item_list,
                        int n_items, int nm_cap, int ck, const Item *items,
                        const cdbl *tbuf, const cdbl *abuf,
                        const int64_t *aoff, const double *PT, int nmax_all,
                        cdbl *X, hipStream_t s, const int32_t *dn = nullptr,
                        int grid = 0, bool lean = false);
// nq_lo, nq_hi: only grains with that many m blocks (the rest untouched)
hipError_t launch_obs_q(const ObsSpec &spec, int n_grains,
                        const int32_t *gi_off, const GrainParams &gp,
                        const cdbl *X, double *qext, double *qextpol,
                        double *qextcpol, double *qabs, double *qabspol,
                        hipStream_t s, int nq_lo = 0,
                        int nq_hi = 0x7fffffff);
hipError_t launch_obs_z(const ObsSpec &spec, int n_grains,
                        const int32_t *gi_off, const GrainParams &gp,
                        const cdbl *X, double *Z, hipStream_t s);

}  // namespace ctm
