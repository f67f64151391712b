/*
 * costuum_amd.h -- C ABI of the MI355X (gfx950) T-matrix library.
 *
 * The library computes converged EBCM T-matrices of spheroids for whole
 * batches of grains on one GPU and the observables derived from them.  All
 * signatures use plain pointers and sizes; complex values are interleaved
 * (re, im) doubles.  Every call returns CTM_OK (0) or a negative status; the
 * message of the last failure is available from ctm_last_error().  Nothing
 * aborts the process.
 *
 * Pointers documented as "device" must point to GPU memory of the context's
 * device (e.g. a torch tensor's data_ptr()); "host" pointers to CPU memory.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * SKIRT/CosTuuM).
 */
#ifndef COSTUUM_AMD_H
#define COSTUUM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTM_OK 0
#define CTM_ERR_ARGUMENT -1
#define CTM_ERR_HIP -2
#define CTM_ERR_STATE -3
#define CTM_ERR_MEMORY -4

/* Convergence status per grain. */
#define CTM_CONVERGED 0
#define CTM_NOT_CONVERGED 1 /* last order kept, as the task path does */

/* Arithmetic of the batched path.  CTM_PRECISION_FP64: every step in FP64,
 * as the reference (Configuration.hpp.in:62).  CTM_PRECISION_MIXED: the
 * blocks of order 17..99 of the all-m phase are eliminated in FP32 (after a
 * power-of-two equilibration) and T = -RgQ Q^-1 is recovered by FP64
 * iterative refinement against the FP64 Q and RgQ; a block whose
 * refinement stalls is solved again in FP64.  The Q assembly, the Bessel
 * recurrences, the convergence windows (orders and their tests) and the
 * other blocks stay FP64. */
#define CTM_PRECISION_FP64 0
#define CTM_PRECISION_MIXED 1

typedef struct ctm_context ctm_context;

/* Creates a context bound to HIP device `device` with its own stream
 * (replaces the per-thread TMatrixAuxiliarySpaceManager scratch and the
 * QuickSched runner: TMatrixResource.hpp:281-434,
 * QuickSchedWrapper.hpp:302-318).  Inside a call, independent launches also
 * run on three side streams of the context; they are joined into the
 * context's stream before the call returns, so the context's stream orders
 * every result. */
int ctm_create(ctm_context **ctx, int device);
int ctm_destroy(ctm_context *ctx);
const char *ctm_last_error(const ctm_context *ctx);
/* The context's HIP stream as an opaque pointer (hipStream_t). */
void *ctm_stream(ctm_context *ctx);
int ctm_synchronize(ctm_context *ctx);

/* Convergence parameters of the task path (TaskManager.hpp:371-384,
 * CTMModule.cpp:160-175): orders min_order..max_order-1, ngauss = glfac *
 * order nodes per hemisphere, relative tolerance on Q_sca and Q_ext. */
typedef struct ctm_convergence {
  uint32_t min_order;
  uint32_t max_order;
  uint32_t glfac;
  uint32_t precision; /* CTM_PRECISION_* */
  double tolerance;
} ctm_convergence;

/*
 * Batched converged T-matrices: the task-path chain InteractionTask ->
 * TMatrixM0Task (per order, until converged) -> TMatrixMAllTask (all m)
 * (TaskManager.hpp:1146-1197,1314-1327; TMatrixResource.hpp:579-1120), for
 * n_grains independent grains.
 *
 * Inputs (device, n_grains doubles each): equal-volume radius, axis ratio
 * d = a/b (d < 1 prolate), wavelength (same length unit as the radius),
 * real and imaginary refractive index.
 * Outputs (device): converged order (uint32), status (int32), and the
 * unnormalised full-T sums Q_sca, Q_ext (TMatrix.hpp:2055-2110; may be NULL).
 * The T-matrices stay in the context until the next batch; fetch them with
 * ctm_tmatrix_get() or use them in the observable calls below.
 */
int ctm_tmatrix_batch(ctm_context *ctx, int n_grains, const double *radius,
                      const double *axis_ratio, const double *wavelength,
                      const double *mr_re, const double *mr_im,
                      const ctm_convergence *conv, uint32_t *nmax_out,
                      int32_t *status_out, double *qsca_out,
                      double *qext_out);

/*
 * The same chain at a forced order and ngauss per grain (host arrays), no
 * convergence loop: the parity entry point (the reference's
 * TMatrix(wavelength, m, R_V, d, nmax, ngauss) constructor plus
 * compute_additional_elements, TMatrix.hpp:532-992).
 */
int ctm_tmatrix_fixed(ctm_context *ctx, int n_grains, const double *radius,
                      const double *axis_ratio, const double *wavelength,
                      const double *mr_re, const double *mr_im,
                      const uint32_t *nmax_host, const uint32_t *ngauss_host,
                      uint32_t precision, double *qsca_out, double *qext_out);

/* Number of complex entries of grain g's T-matrix in the reference block
 * layout (sum over m of (2 nm)^2), or 0 when g is out of range. */
size_t ctm_tmatrix_size(ctm_context *ctx, int grain);
/* Copies grain g's T-matrix to host memory in the reference block layout
 * (TMatrix.hpp:1028-1046): for m = 0..nmax a row-major (2 nm) x (2 nm)
 * complex block, row/col index i*nm + n - max(m,1). */
int ctm_tmatrix_get(ctm_context *ctx, int grain, double *T_host,
                    size_t capacity);

/* Orders computed by the last batch (host copies of the per-grain results). */
int ctm_batch_orders(ctm_context *ctx, uint32_t *nmax_host,
                     uint32_t *ngauss_host, int32_t *status_host);

/* Observables of the context's T-matrices (last batch, or its orientation
 * average), per grain:
 *  - extinction at the n_theta incidence angles: Q_ext, Q_extpol, Q_extcpol
 *    (ExtinctionCoefficientTask.hpp:626-653);
 *  - absorption at the same angles: Q_abs, Q_abspol, minus the GL(ngauss) x
 *    GL(ngauss) quadrature of the scattered light over incoming directions
 *    when account_for_scattering is set (AbsorptionCoefficientTask.hpp:
 *    722-781);
 *  - the 4x4 scattering matrix for incidence theta = pi/2 on an
 *    n_scatter_theta x 2 n_scatter_theta (theta, phi) grid, GL or uniform
 *    (ScatteringMatrixTask.hpp:698-751, grid :107-277).
 * Q outputs are device arrays [n_grains][n_theta] normalised by pi R_V^2
 * (any may be NULL); Z is a device array
 * [n_grains][n_scatter_theta][2 n_scatter_theta][4][4] (may be NULL). */
typedef struct ctm_observable_spec {
  int32_t n_theta;
  const double *theta; /* host, radians */
  int32_t do_extinction;
  int32_t do_absorption;
  int32_t account_for_scattering;
  int32_t ngauss;
  int32_t n_scatter_theta;
  int32_t scatter_gl;
} ctm_observable_spec;

int ctm_observables(ctm_context *ctx, const ctm_observable_spec *spec,
                    double *qext, double *qextpol, double *qextcpol,
                    double *qabs, double *qabspol, double *Z);

/* ctm_observables with no allocation of any kind on the device: no
 * hipMalloc, hipFree or allocating library routine, and no profiler events.
 * No counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  Its scratch (direction tables, the pi/tau table, the
 * bilinear forms and the item order) lives in a fixed 64 MiB device array
 * of the library, one per GPU, held until the process ends; a call whose
 * scratch would exceed it returns CTM_ERR_ARGUMENT with the size needed
 * (X takes n_items * n_pairs * 64 bytes: n_items = sum over grains of
 * nmax + 1, n_pairs = n_theta (1 + ngauss) + n_scatter_theta).  The
 * outputs are the caller's, as for ctm_observables, and agree with it bit
 * for bit.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (several contexts or threads) race, and the caller
 * must serialise them. */
int ctm_observables_noalloc(ctm_context *ctx,
                            const ctm_observable_spec *spec, double *qext,
                            double *qextpol, double *qextcpol, double *qabs,
                            double *qabspol, double *Z);

/* Replaces, for the observable calls, the context's T-matrices by their
 * ensemble averages over an axially symmetric orientation distribution with
 * Legendre coefficients p_N (Mishchenko 1991 eq. 3.27;
 * AlignmentAverageTask.hpp:115-215, TMatrixCalculator.hpp:229-307).
 * pN_host: n_dist rows of ncoef >= 2 nmax + 1 coefficients; dist_of_grain
 * (host, may be NULL = all 0) selects a row per grain, as a size-based
 * alignment distribution does (SizeBasedAlignmentDistribution.hpp:68-146);
 * -1 marks a grain whose alignment is disabled, whose ensemble T-matrix is
 * its own (DisabledAlignmentOrientationDistribution.hpp:36-81,
 * AlignmentAverageTask.hpp:195-214).  Coefficients that are exactly 0 are
 * skipped. */
int ctm_orientation_average(ctm_context *ctx, int n_dist, int ncoef,
                            const double *pN_host,
                            const int32_t *dist_of_grain);
/* Weighted sums over groups of consecutive rows, on the context's stream:
 * out[o][k] = sum_s w[o][s] in[o][s][k], s = 0..group-1 in order, divided by
 * sum_s w[o][s] when normalise != 0.  in: n_groups x group x width, weights:
 * n_groups x group, out: n_groups x width, all device arrays.  With
 * normalise it is the shape average of ShapeAveragingTask::execute
 * (ShapeAveragingTask.hpp:127-166, ShapeDistribution.hpp:95-106); without,
 * the caller-side size integration sum_a w_a Q(a)
 * (scripts/BBOPResult_sim.py:9-71). */
int ctm_weighted_reduce(ctm_context *ctx, int n_groups, int group,
                        int64_t width, const double *in, const double *weights,
                        int normalise, double *out);

/* Grain g's ensemble T-matrix in the reference block layout. */
int ctm_average_get(ctm_context *ctx, int grain, double *T_host,
                    size_t capacity);

/* Replaces grain g's T-matrix, as seen by the observable and query calls,
 * by T_host (count = ctm_tmatrix_size complex values in the reference block
 * layout): TMatrix::set_element (TMatrix.hpp:1062-1080) applied to a whole
 * T-matrix.  The values go to the context's full-layout (ensemble) copy;
 * if the context holds no ensemble T-matrices yet, every grain's T-matrix
 * is first copied there unchanged, as for a disabled alignment. */
int ctm_tmatrix_set(ctm_context *ctx, int grain, const double *T_host,
                    size_t count);

/* S (2x2 complex), Z (4x4) and K (4x4) of grain g for a particle orientation
 * (alpha, beta) and laboratory incoming/outgoing directions, from the
 * context's T-matrices (ensemble ones after ctm_orientation_average):
 * TMatrix::get_forward_scattering_matrix / get_scattering_matrix /
 * get_extinction_matrix (TMatrix.hpp:1306-1482,1891-2010).  out_host
 * receives 40 doubles per query: S as 4 (re, im) pairs row-major, Z
 * row-major, K row-major (K uses the S of the incoming direction, i.e. call
 * with theta_out = theta_in, phi_out = phi_in for get_extinction_matrix). */
typedef struct ctm_query_spec {
  int32_t grain;
  int32_t pad;
  double alpha, beta, theta_in, phi_in, theta_out, phi_out;
} ctm_query_spec;

int ctm_query(ctm_context *ctx, int n, const ctm_query_spec *queries,
              double *out_host);

/* m = 0 convergence sums (sum (2n+1)|T_nn|^2, sum (2n+1) Re T_nn over the
 * m = 0 block, TMatrixCalculator.hpp:108-124) for n candidate
 * (grain, nmax, ngauss) triples; inputs are device arrays with one entry
 * per candidate, nmax/ngauss and outputs host arrays.  Used for the
 * two-phase convergence of TMatrixCalculator::calculate_TMatrix.
 * Invalidates the T-matrices held by the context. */
int ctm_m0_sums(ctm_context *ctx, int n, const double *radius,
                const double *axis_ratio, const double *wavelength,
                const double *mr_re, const double *mr_im,
                const uint32_t *nmax_host, const uint32_t *ngauss_host,
                double *qsca_host, double *qext_host);

/* Forward-mode derivatives of the last batch's T-matrices with respect to
 * the grain parameters p = (R_V, d, Re m, Im m), at the batch's orders
 * (frozen nmax and ngauss).  No reference counterpart: the reference has no
 * derivatives; this differentiates its assembly and T = -RgQ Q^-1
 * (TMatrixResource.hpp:623-730,993-1119) exactly:
 * dT = -(dRgQ + T dQ) Q^-1.  Outputs (host, may be NULL): [n_grains][4]
 * derivatives of the full-T sums Q_sca and Q_ext (TMatrix.hpp:2055-2110). */
int ctm_tmatrix_tangents(ctm_context *ctx, double *dqsca_host,
                         double *dqext_host);
/* dT/dp_dir of grain g in the reference block layout (dir 0..3). */
int ctm_tangent_get(ctm_context *ctx, int grain, int dir, double *dT_host,
                    size_t capacity);
/* Which T-matrices ctm_observables reads: -1 the batch's own, 0..3 the
 * tangent along p_dir.  With a tangent selected only the observables linear
 * in T are allowed (extinction; absorption without the scattered-light
 * correction; no Z, no orientation average): they return the derivative of
 * the cross section divided by pi R_V^2 (the caller adds -2 Q / R_V for
 * p = R_V). */
int ctm_observables_select(ctm_context *ctx, int which);
/* Derivatives along p_dir (after ctm_tmatrix_tangents) of every observable
 * of ctm_observables, including those quadratic in T: the absorption with
 * the scattered-light correction (AbsorptionCoefficientTask.hpp:742-773) and
 * the scattering matrix Z.  S is linear in T and each observable at most
 * quadratic in S, so O(T + s dT) - O(T - s dT) = 2 s dO exactly; the library
 * evaluates the observables at T +- s dT with a per-grain s = max |T| /
 * max |dT| and returns the difference / (2 s).  Same specification and
 * output layout as ctm_observables; Q derivatives are at fixed
 * normalisation (for p = R_V the caller adds -2 Q / R_V, Z is not
 * normalised).  No reference counterpart (the reference has no
 * derivatives).  The batch's own T-matrices only (no ensemble). */
int ctm_observables_derivative(ctm_context *ctx, int dir,
                               const ctm_observable_spec *spec, double *dqext,
                               double *dqextpol, double *dqextcpol,
                               double *dqabs, double *dqabspol, double *dZ);

/* Vector-Jacobian product of the full-T sums for per-grain cotangents
 * (host, may be NULL): grad[g][p] = cot_qsca[g] dQ_sca/dp +
 * cot_qext[g] dQ_ext/dp, after ctm_tmatrix_tangents. */
int ctm_vjp(ctm_context *ctx, const double *cot_qsca, const double *cot_qext,
            double *grad);

/* Host-only helpers (no GPU work).  ctm_real_fn is a real function of one
 * real variable with a caller argument, as the reference passes its Python
 * callbacks (PyOrientationDistribution.hpp:501-533). */
typedef double (*ctm_real_fn)(double x, void *arg);

/* Gauss-Legendre nodes (ascending) and weights of the given order on
 * [a, b]: SpecialFunctions::get_gauss_legendre_points_and_weights(_ab)
 * (SpecialFunctions.hpp:647-751). */
int ctm_gauss_legendre(uint32_t order, double a, double b, double *x,
                       double *w);

/* R_V / R_S of a spheroid with axis ratio d:
 * get_equal_volume_to_equal_surface_area_sphere_ratio
 * (SpecialFunctions.hpp:558-583). */
double ctm_rv_over_rs(double axis_ratio);

/* Adaptive Gauss-Legendre quadrature of f on [a, b], orders start..max-1
 * until two successive results agree to abs_err or rel_err:
 * SpecialFunctions::gauss_legendre_quadrature (SpecialFunctions.hpp:772-815).
 * Returns CTM_NOT_CONVERGED (with the last estimate) where the reference
 * warns. */
int ctm_adaptive_quadrature(ctm_real_fn f, void *arg, double a, double b,
                            uint32_t start_order, uint32_t max_order,
                            double abs_err, double rel_err, double *result);

/* Legendre coefficients p_0..p_order of an orientation distribution p(beta)
 * (NULL: the isotropic p = 1/2), normalised numerically:
 * OrientationDistribution::initialise (OrientationDistribution.hpp:220-243).
 * coefficients has order + 1 entries. */
int ctm_orientation_coefficients(ctm_real_fn p, void *arg, uint32_t order,
                                 double abs_err, double rel_err,
                                 double *coefficients);

/* Estimated device bytes one grain takes in the batched path (its
 * T-matrix blocks, Q/RgQ, ensemble copy, the Bessel tables of its
 * convergence orders and, for n_pairs observable direction pairs, their
 * bilinear forms), from an order estimate x + 4.05 x^(1/3) + 8 on the
 * largest semi-axis clamped to [min_order + 1, max_order - 1].  Callers
 * size batches with it to honour a memory budget, as the reference's
 * maximum_memory_usage does for its task resources
 * (TaskManager.hpp:101-348,718-760).  Host only. */
double ctm_grain_memory_estimate(double radius, double axis_ratio,
                                 double wavelength, uint32_t min_order,
                                 uint32_t max_order, int32_t n_pairs);

/* Kernel timing.  With profiling on, every launch of the library is
 * bracketed by two HIP events on the stream it runs on; ctm_profile_read
 * returns, per kernel, the number of launches, the summed event time and the
 * algorithmic FLOPs and compulsory HBM bytes those launches stand for.  No
 * reference counterpart: it replaces QuickSched's per-task tic/toc log
 * (qsched.c:572,620-622; QuickSchedWrapper.hpp:494-508). */
#define CTM_KERNEL_GAUSS_LEGENDRE 0
#define CTM_KERNEL_WIGNER 1
#define CTM_KERNEL_TABLES 2
#define CTM_KERNEL_ASSEMBLE 3
#define CTM_KERNEL_SOLVE 4
#define CTM_KERNEL_CONVERGE 5
#define CTM_KERNEL_SUMS 6
#define CTM_KERNEL_OBS_X 7
#define CTM_KERNEL_OBS_SUM 8
#define CTM_KERNEL_AVERAGE 9
#define CTM_KERNEL_ASSEMBLE_SMALL 10 /* blocks of order <= 16 */
#define CTM_KERNEL_ASSEMBLE_WIDE 11  /* blocks on 48-row tiles */
#define CTM_KERNEL_ASSEMBLE_PACK 12  /* blocks of order 9-24, four per group */
#define CTM_KERNEL_REFINE 13  /* FP64 refinement of the mixed solve */
/* 48-row tiles of the plan's convergence windows, which run beside the
 * other assembly parts (CTM_KERNEL_ASSEMBLE_WIDE: the launches that run
 * alone) */
#define CTM_KERNEL_ASSEMBLE_WINDOW 14
#define CTM_KERNEL_COUNT 15

typedef struct ctm_kernel_stats {
  uint64_t launches;
  double milliseconds;
  double flops;
  double bytes;
} ctm_kernel_stats;

/* Blocks of the last batch call that went through the FP32 elimination
 * (CTM_PRECISION_MIXED) and those re-solved in FP64 because their
 * refinement stalled.  No reference counterpart. */
int ctm_precision_stats(ctm_context *ctx, int64_t *blocks_mixed,
                        int64_t *blocks_fallback);

int ctm_profile_enable(ctm_context *ctx, int on);
int ctm_profile_reset(ctm_context *ctx);
int ctm_profile_read(ctm_context *ctx, int kernel, ctm_kernel_stats *out);

/*
 * Batch plan: the task path for a fixed batch (the whole of
 * TaskManager::generate_tasks + QuickSched::execute_tasks for converged
 * T-matrices and their averaged Q_ext / Q_abs, TaskManager.hpp:445-455,
 * 1146-1197, 1314-1327; QuickSchedWrapper.hpp:375-389) with the scheduler's
 * work queues built on the GPU: the convergence rounds, the all-m blocks
 * and the observables of the batch are listed by device kernels from the
 * device's convergence state, so a run is a fixed sequence of launches that
 * the host enqueues at once, with no read-back until the run has finished
 * (the host path reads the convergence state back after every round and
 * lists the next one on the CPU).  Results equal
 * ctm_tmatrix_batch + ctm_observables on the same inputs (same orders,
 * same sums).
 *
 * Inputs are device arrays (n_grains doubles each, as ctm_tmatrix_batch)
 * bound at creation; a run reads their current contents.  The last order
 * (conv->max_order - 1) is at most 99.  obs (may be NULL): Q_ext / Q_abs
 * at n_theta >= 1 angles, no scattering-matrix grid.  Outputs are device
 * arrays bound at creation (any may be NULL): nmax, status, qsca, qext per
 * grain; obs_* [n_grains][n_theta] as ctm_observables.  Destroy plans
 * before their context.  No reference counterpart as an entry point.
 */
typedef struct ctm_plan ctm_plan;
typedef struct ctm_plan_outputs {
  uint32_t *nmax;
  int32_t *status;
  double *qsca;
  double *qext;
  double *obs_qext;
  double *obs_qextpol;
  double *obs_qextcpol;
  double *obs_qabs;
  double *obs_qabspol;
} ctm_plan_outputs;
int ctm_plan_create(ctm_context *ctx, int n_grains, const double *radius,
                    const double *axis_ratio, const double *wavelength,
                    const double *mr_re, const double *mr_im,
                    const ctm_convergence *conv,
                    const ctm_observable_spec *obs,
                    const ctm_plan_outputs *outputs, ctm_plan **plan);
/* One run; returns when the outputs are complete.  A run whose device
 * lists outgrew the plan's memory pools is repeated with larger pools. */
int ctm_plan_run(ctm_plan *plan);
typedef struct ctm_plan_info {
  int64_t runs;            /* runs so far, repeats included */
  int64_t reruns;          /* repeats after a pool had to grow */
  int64_t mixed_blocks;    /* last run: blocks through the FP32 elimination */
  int64_t mixed_fallback;  /* last run: of those, re-solved in FP64 */
  int64_t pool_bytes;      /* device bytes of the data pools */
  int32_t allm_items;      /* last run: (grain, m) blocks */
  int32_t window_items[3]; /* last run: m = 0 orders of each window round */
} ctm_plan_info;
int ctm_plan_info_get(ctm_plan *plan, ctm_plan_info *info);
/* Autotuner (no reference counterpart; it replaces the fixed schedule of
 * QuickSched and the hand-set tile choice): times whole runs of the plan
 * (median of `runs` after one untimed run per candidate) under candidate
 * values of the first-window margin (end = ceil(scale * model + add)), of
 * the cost ratio that sends blocks to 48-row or 32-row assembly tiles and
 * of the launch order of the largest blocks' solves (big_first 1: right
 * after the part that assembles them; 2: alone, the other parts waiting
 * for them) and of the kernel variant that forms
 * the largest blocks' observable forms (obs_lean: the one with the smallest
 * LDS footprint, which shares CUs with the kernels beside it), by
 * coordinate descent, and keeps the fastest.  Orders and values do not
 * depend on them, except that the tile ratio can move a block to tiles of
 * the other height, whose edge tiles may sum the quadrature nodes in
 * another order: values then change in rounding only.  `out` (may be NULL)
 * receives the chosen values and times. */
typedef struct ctm_plan_tuning {
  double window_scale;
  double window_add;
  double micro_ratio;
  double ms_default;  /* per run, the values before tuning */
  double ms_chosen;   /* per run, the chosen values */
  int32_t trials;
  int32_t runs_per_trial;
  int32_t big_first;
  int32_t obs_lean;   /* largest blocks' X on the low-LDS kernel variant */
} ctm_plan_tuning;
int ctm_plan_tune(ctm_plan *plan, int runs, ctm_plan_tuning *out);
/* Applies tunable values (e.g. those ctm_plan_tune chose for a similar
 * batch earlier) without timing anything; the time fields are ignored.
 * No reference counterpart. */
int ctm_plan_tuning_set(ctm_plan *plan, const ctm_plan_tuning *tuning);
int ctm_plan_destroy(ctm_plan *plan);

#ifdef __cplusplus
}
#endif

#endif /* COSTUUM_AMD_H */
