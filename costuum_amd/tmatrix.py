"""Batched T-matrices on the GPU through the C ABI (include/costuum_amd.h).

Host-side mirror of the reference's T-matrix entry points:
  * ``tmatrix_batch`` -- the task-path chain InteractionTask ->
    TMatrixM0Task (orders until converged) -> TMatrixMAllTask for a whole
    batch of grains (TaskManager.hpp:1146-1197,1314-1327);
  * ``tmatrix_fixed`` -- all m blocks at a forced (nmax, ngauss), the
    TMatrix(wavelength, m, R_V, d, nmax, ngauss) constructor followed by
    compute_additional_elements (TMatrix.hpp:532-992).
Device memory and streams come from PyTorch; the arithmetic runs in the
HIP library.  Lengths (radius, wavelength) share one unit.
"""
import ctypes

import numpy as np
import torch

from . import _native

PRECISIONS = {"fp64": _native.CTM_PRECISION_FP64,
              "mixed": _native.CTM_PRECISION_MIXED}


class CtmError(RuntimeError):
    pass


class Context:
    """A library context on one HIP device (owns a stream and scratch)."""

    def __init__(self, device=0):
        self._lib = _native.lib()
        if not torch.cuda.is_available():
            raise CtmError("no GPU visible to torch")
        self.device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        rc = self._lib.ctm_create(ctypes.byref(h), device)
        if rc != 0:
            raise CtmError("ctm_create failed (%d)" % rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ctm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def check(self, rc, what):
        if rc != 0:
            msg = self._lib.ctm_last_error(self._h)
            raise CtmError("%s failed (%d): %s" % (
                what, rc, msg.decode() if msg else ""))

    def synchronize(self):
        self.check(self._lib.ctm_synchronize(self._h), "ctm_synchronize")

    def precision_stats(self):
        """(blocks through the FP32 elimination, blocks re-solved in FP64)
        of the last batch call."""
        a = ctypes.c_int64()
        b = ctypes.c_int64()
        self.check(self._lib.ctm_precision_stats(self._h, ctypes.byref(a),
                                                 ctypes.byref(b)),
                   "ctm_precision_stats")
        return a.value, b.value

    def stream_handle(self):
        return self._lib.ctm_stream(self._h)

    def profile(self, on=True):
        self.check(self._lib.ctm_profile_enable(self._h, int(on)),
                   "ctm_profile_enable")
        self.check(self._lib.ctm_profile_reset(self._h), "ctm_profile_reset")

    def profile_stats(self):
        """{kernel: (launches, ms, flops, bytes)} since the last reset."""
        out = {}
        for k, name in enumerate(_native.KERNELS):
            st = _native.CtmKernelStats()
            self.check(self._lib.ctm_profile_read(self._h, k,
                                                  ctypes.byref(st)),
                       "ctm_profile_read")
            out[name] = (st.launches, st.milliseconds, st.flops, st.bytes)
        return out


_default = {}


def default_context(device=0):
    ctx = _default.get(device)
    if ctx is None:
        ctx = Context(device)
        _default[device] = ctx
    return ctx


def _dev(x, device):
    t = torch.as_tensor(x, dtype=torch.float64)
    return t.to(device).contiguous()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class BatchResult:
    """Per-grain results of one batch; T-matrices stay on the device until
    the next batch on the same context."""

    def __init__(self, ctx, n, nmax, status, qsca, qext, ngauss=None):
        self.ctx = ctx
        self.n = n
        self.nmax = nmax
        self.status = status
        self.qsca = qsca
        self.qext = qext
        self.ngauss = ngauss

    def orders(self):
        L = self.ctx._lib
        nmax = np.zeros(self.n, dtype=np.uint32)
        ngauss = np.zeros(self.n, dtype=np.uint32)
        status = np.zeros(self.n, dtype=np.int32)
        self.ctx.check(L.ctm_batch_orders(
            self.ctx.handle, nmax.ctypes.data, ngauss.ctypes.data,
            status.ctypes.data), "ctm_batch_orders")
        return nmax, ngauss, status

    def average(self, pN, dist_of_grain=None):
        """Ensemble-average the batch's T-matrices over orientation
        distribution(s) with Legendre coefficients pN ([n_dist][ncoef]);
        later observables use the averages."""
        pN = np.ascontiguousarray(np.atleast_2d(pN), dtype=np.float64)
        d = None
        if dist_of_grain is not None:
            d = np.ascontiguousarray(dist_of_grain, dtype=np.int32)
        self.ctx.check(self.ctx._lib.ctm_orientation_average(
            self.ctx.handle, pN.shape[0], pN.shape[1], pN.ctypes.data,
            d.ctypes.data if d is not None else None),
            "ctm_orientation_average")

    def averaged_tmatrix(self, g):
        L = self.ctx._lib
        size = L.ctm_tmatrix_size(self.ctx.handle, g)
        out = np.zeros(size, dtype=np.complex128)
        self.ctx.check(L.ctm_average_get(self.ctx.handle, g, out.ctypes.data,
                                         size), "ctm_average_get")
        return out

    def tmatrix(self, g):
        """Grain g's T-matrix in the reference block layout (flat complex
        array; see the oracle's tmatrix_blocks to split it per m)."""
        L = self.ctx._lib
        size = L.ctm_tmatrix_size(self.ctx.handle, g)
        if size == 0:
            raise CtmError("no T-matrix for grain %d" % g)
        out = np.zeros(size, dtype=np.complex128)
        self.ctx.check(L.ctm_tmatrix_get(self.ctx.handle, g, out.ctypes.data,
                                         size), "ctm_tmatrix_get")
        return out


def observables(theta, ctx=None, do_extinction=True, do_absorption=True,
                account_for_scattering=True, ngauss=20, n_scatter_theta=0,
                scatter_gl=True, n_grains=None):
    """Extinction/absorption at the angles `theta` (radians) and optionally
    the scattering-matrix grid, for the T-matrices of the context's last
    batch.  Returns a dict of device tensors."""
    ctx = ctx or default_context()
    dev = ctx.device
    th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    n = n_grains
    if n is None:
        raise ValueError("n_grains is required")
    nt = len(th)
    out = {}
    for k in ("qext", "qextpol", "qextcpol", "qabs", "qabspol"):
        out[k] = torch.zeros((n, nt), dtype=torch.float64, device=dev)
    Z = None
    if n_scatter_theta > 0:
        Z = torch.zeros((n, n_scatter_theta, 2 * n_scatter_theta, 4, 4),
                        dtype=torch.float64, device=dev)
        out["Z"] = Z
    # the output buffers were zeroed on torch's stream; the library writes
    # them on its own stream
    torch.cuda.current_stream(dev).synchronize()
    spec = _native.CtmObservableSpec(nt, th.ctypes.data, int(do_extinction),
                                     int(do_absorption),
                                     int(account_for_scattering), ngauss,
                                     n_scatter_theta, int(scatter_gl))
    rc = ctx._lib.ctm_observables(
        ctx.handle, ctypes.byref(spec), _ptr(out["qext"]),
        _ptr(out["qextpol"]), _ptr(out["qextcpol"]), _ptr(out["qabs"]),
        _ptr(out["qabspol"]), _ptr(Z))
    ctx.check(rc, "ctm_observables")
    return out


def tmatrix_batch(radius, axis_ratio, wavelength, mr_re, mr_im, min_order=4,
                  max_order=100, glfac=2, tolerance=1e-4, precision="fp64",
                  ctx=None, sync=True):
    """Converged T-matrices for a batch of spheroids (task-path contract)."""
    ctx = ctx or default_context()
    dev = ctx.device
    arrs = [_dev(a, dev) for a in (radius, axis_ratio, wavelength, mr_re,
                                    mr_im)]
    n = arrs[0].numel()
    if any(a.numel() != n for a in arrs):
        raise ValueError("grain arrays differ in length")
    nmax = torch.zeros(n, dtype=torch.int32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    qsca = torch.zeros(n, dtype=torch.float64, device=dev)
    qext = torch.zeros(n, dtype=torch.float64, device=dev)
    conv = _native.CtmConvergence(min_order, max_order, glfac,
                                  PRECISIONS[precision], tolerance)
    torch.cuda.current_stream(dev).synchronize()
    rc = ctx._lib.ctm_tmatrix_batch(
        ctx.handle, n, *[_ptr(a) for a in arrs], ctypes.byref(conv),
        _ptr(nmax), _ptr(status), _ptr(qsca), _ptr(qext))
    ctx.check(rc, "ctm_tmatrix_batch")
    if sync:
        ctx.synchronize()
    return BatchResult(ctx, n, nmax, status, qsca, qext)


class BatchPlan:
    """A fixed batch whose whole step -- convergence rounds, all-m blocks,
    full-T sums and (optionally) Q_ext / Q_abs at the angles `theta` -- runs
    with its work lists built on the GPU (ctm_plan_*): no host round trip
    between the phases.

    The inputs are copied to device tensors owned by the plan
    (``self.inputs``: radius, axis_ratio, wavelength, mr_re, mr_im); a run
    reads their current contents, so a new batch of the same size is run by
    writing into them.  Outputs are device tensors: ``nmax``, ``status``,
    ``qsca``, ``qext`` and, with theta, ``obs`` (qext, qextpol, qextcpol,
    qabs, qabspol [n][n_theta])."""

    def __init__(self, radius, axis_ratio, wavelength, mr_re, mr_im,
                 min_order=4, max_order=100, glfac=2, tolerance=1e-4,
                 precision="fp64", theta=None, ngauss=20,
                 account_for_scattering=True, ctx=None):
        self.ctx = ctx or default_context()
        dev = self.ctx.device
        self.inputs = [_dev(a, dev) for a in (radius, axis_ratio, wavelength,
                                               mr_re, mr_im)]
        n = self.inputs[0].numel()
        if any(a.numel() != n for a in self.inputs):
            raise ValueError("grain arrays differ in length")
        self.n = n
        self.nmax = torch.zeros(n, dtype=torch.int32, device=dev)
        self.status = torch.zeros(n, dtype=torch.int32, device=dev)
        self.qsca = torch.zeros(n, dtype=torch.float64, device=dev)
        self.qext = torch.zeros(n, dtype=torch.float64, device=dev)
        self.obs = None
        spec = None
        if theta is not None:
            th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
            self._theta = th
            self.obs = {k: torch.zeros((n, len(th)), dtype=torch.float64,
                                       device=dev)
                        for k in ("qext", "qextpol", "qextcpol", "qabs",
                                  "qabspol")}
            spec = _native.CtmObservableSpec(len(th), th.ctypes.data, 1, 1,
                                             int(account_for_scattering),
                                             ngauss, 0, 0)
        o = self.obs or {}
        outs = _native.CtmPlanOutputs(
            self.nmax.data_ptr(), self.status.data_ptr(),
            self.qsca.data_ptr(), self.qext.data_ptr(),
            *[o[k].data_ptr() if k in o else None
              for k in ("qext", "qextpol", "qextcpol", "qabs", "qabspol")])
        conv = _native.CtmConvergence(min_order, max_order, glfac,
                                      PRECISIONS[precision], tolerance)
        torch.cuda.current_stream(dev).synchronize()
        h = ctypes.c_void_p()
        rc = self.ctx._lib.ctm_plan_create(
            self.ctx.handle, n, *[_ptr(a) for a in self.inputs],
            ctypes.byref(conv), ctypes.byref(spec) if spec else None,
            ctypes.byref(outs), ctypes.byref(h))
        self.ctx.check(rc, "ctm_plan_create")
        self._h = h

    def run(self):
        """One step; returns when the outputs are complete."""
        self.ctx.check(self.ctx._lib.ctm_plan_run(self._h), "ctm_plan_run")
        return self

    def tune(self, runs=3):
        """Autotunes the plan's first-window margin and tile-shape ratio by
        timing whole runs (ctm_plan_tune); returns the chosen values and
        the per-run times before and after."""
        t = _native.CtmPlanTuning()
        self.ctx.check(self.ctx._lib.ctm_plan_tune(self._h, runs,
                                                   ctypes.byref(t)),
                       "ctm_plan_tune")
        return {"window_scale": t.window_scale, "window_add": t.window_add,
                "micro_ratio": t.micro_ratio, "big_first": t.big_first,
                "obs_lean": t.obs_lean,
                "ms_default": t.ms_default,
                "ms_chosen": t.ms_chosen, "trials": t.trials,
                "runs_per_trial": t.runs_per_trial}

    def set_tuning(self, window_scale=1.04, window_add=3.0, micro_ratio=0.5,
                   big_first=0, obs_lean=1):
        """Applies tunable values without timing (ctm_plan_tuning_set),
        e.g. those tune() returned for a similar batch."""
        t = _native.CtmPlanTuning()
        t.window_scale = window_scale
        t.window_add = window_add
        t.micro_ratio = micro_ratio
        t.big_first = int(big_first)
        t.obs_lean = int(obs_lean)
        self.ctx.check(self.ctx._lib.ctm_plan_tuning_set(self._h,
                                                         ctypes.byref(t)),
                       "ctm_plan_tuning_set")

    def info(self):
        i = _native.CtmPlanInfo()
        self.ctx.check(self.ctx._lib.ctm_plan_info_get(self._h,
                                                       ctypes.byref(i)),
                       "ctm_plan_info_get")
        return {"runs": i.runs, "reruns": i.reruns,
                "mixed_blocks": i.mixed_blocks,
                "mixed_fallback": i.mixed_fallback,
                "pool_bytes": i.pool_bytes,
                "allm_items": i.allm_items,
                "window_items": list(i.window_items)}

    def close(self):
        if getattr(self, "_h", None):
            self.ctx._lib.ctm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tmatrix_fixed(radius, axis_ratio, wavelength, mr_re, mr_im, nmax, ngauss,
                  precision="fp64", ctx=None):
    """All m blocks at forced (nmax, ngauss) per grain."""
    ctx = ctx or default_context()
    dev = ctx.device
    arrs = [_dev(a, dev) for a in (radius, axis_ratio, wavelength, mr_re,
                                    mr_im)]
    n = arrs[0].numel()
    nm = np.ascontiguousarray(np.broadcast_to(nmax, (n,)), dtype=np.uint32)
    ng = np.ascontiguousarray(np.broadcast_to(ngauss, (n,)), dtype=np.uint32)
    qsca = torch.zeros(n, dtype=torch.float64, device=dev)
    qext = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.current_stream(dev).synchronize()
    rc = ctx._lib.ctm_tmatrix_fixed(
        ctx.handle, n, *[_ptr(a) for a in arrs], nm.ctypes.data,
        ng.ctypes.data, PRECISIONS[precision], _ptr(qsca), _ptr(qext))
    ctx.check(rc, "ctm_tmatrix_fixed")
    ctx.synchronize()
    return BatchResult(ctx, n, torch.as_tensor(nm.astype(np.int32)), None,
                       qsca, qext, ngauss=ng)


_QUERY_DTYPE = np.dtype([("grain", "<i4"), ("pad", "<i4"), ("alpha", "<f8"),
                         ("beta", "<f8"), ("theta_in", "<f8"),
                         ("phi_in", "<f8"), ("theta_out", "<f8"),
                         ("phi_out", "<f8")])
assert _QUERY_DTYPE.itemsize == ctypes.sizeof(_native.CtmQuery)


def query(ctx, grains, alpha, beta, theta_in, phi_in, theta_out, phi_out):
    """Amplitude (S, 2x2 complex), phase (Z, 4x4) and extinction (K, 4x4)
    matrices of the context's T-matrices (the averaged ones after
    ``BatchResult.average``) for particles at orientation (alpha, beta) and
    the given laboratory incident/scattered directions
    (TMatrix.hpp:1246-1482,1891-2010).  All arguments broadcast; returns
    (S [..., 2, 2], Z [..., 4, 4], K [..., 4, 4])."""
    b = np.broadcast_arrays(np.asarray(grains, dtype=np.int64), alpha, beta,
                            theta_in, phi_in, theta_out, phi_out)
    shape = b[0].shape
    n = b[0].size
    # packed like ctm_query_spec: int32 grain, int32 pad, six doubles
    qs = np.zeros(max(n, 1), dtype=_QUERY_DTYPE)
    qs["grain"][:n] = np.ravel(b[0])
    for name, x in zip(_QUERY_DTYPE.names[2:], b[1:]):
        qs[name][:n] = np.ravel(x)
    out = np.zeros((max(n, 1), _native.QUERY_OUT))
    ctx.check(ctx._lib.ctm_query(
        ctx.handle, n, qs.ctypes.data_as(ctypes.POINTER(_native.CtmQuery)),
        out.ctypes.data), "ctm_query")
    out = out[:n]
    S = (out[:, 0:8:2] + 1j * out[:, 1:8:2]).reshape(shape + (2, 2))
    Z = out[:, 8:24].reshape(shape + (4, 4))
    K = out[:, 24:40].reshape(shape + (4, 4))
    return S, Z, K


def m0_sums(radius, axis_ratio, wavelength, mr_re, mr_im, nmax, ngauss,
            ctx=None):
    """m = 0 convergence sums (sum (2n+1)|T_nn|^2, sum (2n+1) Re T_nn) of
    candidate (nmax, ngauss) pairs, the inner quantity of
    TMatrixCalculator::calculate_TMatrix (TMatrixCalculator.hpp:108-124).
    Replaces the T-matrices held by the context."""
    ctx = ctx or default_context()
    dev = ctx.device
    arrs = [_dev(a, dev) for a in (radius, axis_ratio, wavelength, mr_re,
                                    mr_im)]
    n = arrs[0].numel()
    nm = np.ascontiguousarray(np.broadcast_to(nmax, (n,)), dtype=np.uint32)
    ng = np.ascontiguousarray(np.broadcast_to(ngauss, (n,)), dtype=np.uint32)
    qs = np.zeros(n)
    qe = np.zeros(n)
    torch.cuda.current_stream(dev).synchronize()
    ctx.check(ctx._lib.ctm_m0_sums(
        ctx.handle, n, *[_ptr(a) for a in arrs], nm.ctypes.data,
        ng.ctypes.data, qs.ctypes.data, qe.ctypes.data), "ctm_m0_sums")
    return qs, qe


# ---------------------------------------------------------------------------
# Derivatives with respect to the grain parameters (forward mode in the
# library; no reference counterpart)

PARAMS = ("radius", "axis_ratio", "mr_re", "mr_im")


def tangents(res):
    """d(Q_sca, Q_ext full-T sums)/dp for p in PARAMS at the batch's frozen
    orders: two arrays [n_grains, 4].  The tangent T-matrices stay in the
    context for tangent_tmatrix / observables_tangent."""
    ctx = res.ctx
    dqs = np.zeros((res.n, 4))
    dqe = np.zeros((res.n, 4))
    ctx.check(ctx._lib.ctm_tmatrix_tangents(ctx.handle, dqs.ctypes.data,
                                            dqe.ctypes.data),
              "ctm_tmatrix_tangents")
    return dqs, dqe


def tangent_tmatrix(res, g, direction):
    """dT/dp of grain g in the reference block layout."""
    L = res.ctx._lib
    size = L.ctm_tmatrix_size(res.ctx.handle, g)
    out = np.zeros(size, dtype=np.complex128)
    res.ctx.check(L.ctm_tangent_get(res.ctx.handle, g, direction,
                                    out.ctypes.data, size), "ctm_tangent_get")
    return out


def observables_tangent(theta, direction, ctx, n_grains, radius=None,
                        do_extinction=True, do_absorption=True):
    """d/dp of the observables linear in T (extinction; absorption without
    the scattered-light correction) at the angles theta, for p =
    PARAMS[direction].  For p = R_V the normalisation 1 / (pi R_V^2) adds
    -2 Q / R_V; radius (per grain) is required then."""
    ctx.check(ctx._lib.ctm_observables_select(ctx.handle, int(direction)),
              "ctm_observables_select")
    try:
        d = observables(theta, ctx=ctx, do_extinction=do_extinction,
                        do_absorption=do_absorption,
                        account_for_scattering=False, n_grains=n_grains)
    finally:
        ctx.check(ctx._lib.ctm_observables_select(ctx.handle, -1),
                  "ctm_observables_select")
    if direction == 0:
        if radius is None:
            raise ValueError("radius is required for the R_V direction")
        q = observables(theta, ctx=ctx, do_extinction=do_extinction,
                        do_absorption=do_absorption,
                        account_for_scattering=False, n_grains=n_grains)
        rv = torch.as_tensor(np.asarray(radius, dtype=np.float64),
                             device=ctx.device).reshape(-1, 1)
        for k in d:
            d[k] = d[k] - 2.0 * q[k] / rv
    return d


def vjp(res, cot_qsca=None, cot_qext=None):
    """Vector-Jacobian product of the full-T sums: [n_grains, 4]."""
    ctx = res.ctx
    cs = None if cot_qsca is None else np.ascontiguousarray(cot_qsca,
                                                           dtype=np.float64)
    ce = None if cot_qext is None else np.ascontiguousarray(cot_qext,
                                                           dtype=np.float64)
    out = np.zeros((res.n, 4))
    ctx.check(ctx._lib.ctm_vjp(ctx.handle,
                               cs.ctypes.data if cs is not None else None,
                               ce.ctypes.data if ce is not None else None,
                               out.ctypes.data), "ctm_vjp")
    return out


def observables_derivative(theta, direction, ctx, n_grains, radius=None,
                           do_extinction=True, do_absorption=True,
                           account_for_scattering=True, ngauss=20,
                           n_scatter_theta=0, scatter_gl=True):
    """d/dp of every observable of ``observables`` (p = PARAMS[direction],
    after ``tangents``), including the absorption with the scattered-light
    correction and the scattering matrix Z (ctm_observables_derivative: the
    exact symmetric difference of the quadratic forms).  For p = R_V the Q
    values get the -2 Q / R_V of their 1 / (pi R_V^2) normalisation here
    (radius per grain is required then); Z is not normalised.  Returns a
    dict of device tensors like ``observables``."""
    dev = ctx.device
    th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    n, nt = n_grains, len(th)
    out = {k: torch.zeros((n, nt), dtype=torch.float64, device=dev)
           for k in ("qext", "qextpol", "qextcpol", "qabs", "qabspol")}
    Z = None
    if n_scatter_theta > 0:
        Z = torch.zeros((n, n_scatter_theta, 2 * n_scatter_theta, 4, 4),
                        dtype=torch.float64, device=dev)
        out["Z"] = Z
    torch.cuda.current_stream(dev).synchronize()
    spec = _native.CtmObservableSpec(nt, th.ctypes.data, int(do_extinction),
                                     int(do_absorption),
                                     int(account_for_scattering), ngauss,
                                     n_scatter_theta, int(scatter_gl))
    ctx.check(ctx._lib.ctm_observables_derivative(
        ctx.handle, int(direction), ctypes.byref(spec), _ptr(out["qext"]),
        _ptr(out["qextpol"]), _ptr(out["qextcpol"]), _ptr(out["qabs"]),
        _ptr(out["qabspol"]), _ptr(Z)), "ctm_observables_derivative")
    if direction == 0:
        if radius is None:
            raise ValueError("radius is required for the R_V direction")
        q = observables(th, ctx=ctx, do_extinction=do_extinction,
                        do_absorption=do_absorption,
                        account_for_scattering=account_for_scattering,
                        ngauss=ngauss, n_grains=n)
        rv = torch.as_tensor(np.asarray(radius, dtype=np.float64),
                             device=dev).reshape(-1, 1)
        for k in ("qext", "qextpol", "qextcpol", "qabs", "qabspol"):
            out[k] = out[k] - 2.0 * q[k] / rv
    return out


class _Observables(torch.autograd.Function):
    """(radius, axis_ratio, mr_re, mr_im) -> (Q_ext, Q_abs, Q_sca) at the
    incidence angles theta (Q_abs with the ngauss x ngauss scattered-light
    correction, Q_sca = Q_ext - Q_abs) at frozen orders.  The forward pass
    also forms the per-grain Jacobian [n, n_theta, 2, 4] with the library's
    forward-mode tangents (four directions), from which backward forms any
    vector-Jacobian product."""

    @staticmethod
    def forward(fctx, radius, axis_ratio, mr_re, mr_im, wavelength, nmax,
                ngauss, theta, ngauss_abs, ctx):
        res = tmatrix_fixed(radius.detach(), axis_ratio.detach(),
                            wavelength, mr_re.detach(), mr_im.detach(), nmax,
                            ngauss, ctx=ctx)
        n = res.n
        q = observables(theta, ctx=ctx, ngauss=ngauss_abs, n_grains=n)
        tangents(res)
        rv = radius.detach().cpu().numpy()
        cols = []
        for j in range(4):
            d = observables_derivative(theta, j, ctx, n, radius=rv,
                                       ngauss=ngauss_abs)
            cols.append(torch.stack([d["qext"], d["qabs"]], dim=-1))
        fctx.save_for_backward(torch.stack(cols, dim=-1))
        return q["qext"], q["qabs"], q["qext"] - q["qabs"]

    @staticmethod
    def backward(fctx, g_ext, g_abs, g_sca):
        J, = fctx.saved_tensors
        ge = (g_ext + g_sca).unsqueeze(-1)
        ga = (g_abs - g_sca).unsqueeze(-1)
        grad = (ge * J[:, :, 0, :] + ga * J[:, :, 1, :]).sum(dim=1)
        return (grad[:, 0], grad[:, 1], grad[:, 2], grad[:, 3], None, None,
                None, None, None, None)


def observables_autograd(radius, axis_ratio, wavelength, mr_re, mr_im, nmax,
                         ngauss, theta, ngauss_abs=20, ctx=None):
    """Differentiable (Q_ext, Q_abs, Q_sca) [n, n_theta] at the incidence
    angles theta (torch autograd over radius, axis_ratio, mr_re, mr_im) at
    forced orders."""
    ctx = ctx or default_context()
    return _Observables.apply(radius, axis_ratio, mr_re, mr_im, wavelength,
                              nmax, ngauss, theta, ngauss_abs, ctx)


class _TMatrixSums(torch.autograd.Function):
    """(radius, axis_ratio, mr_re, mr_im) -> (Q_sca, Q_ext) full-T sums at
    frozen orders, differentiable through the library's tangents."""

    @staticmethod
    def forward(fctx, radius, axis_ratio, mr_re, mr_im, wavelength, nmax,
                ngauss, ctx):
        res = tmatrix_fixed(radius.detach(), axis_ratio.detach(),
                            wavelength, mr_re.detach(), mr_im.detach(), nmax,
                            ngauss, ctx=ctx)
        dqs, dqe = tangents(res)
        fctx.save_for_backward(torch.as_tensor(dqs, device=radius.device),
                               torch.as_tensor(dqe, device=radius.device))
        return res.qsca.clone(), res.qext.clone()

    @staticmethod
    def backward(fctx, gs, ge):
        dqs, dqe = fctx.saved_tensors
        g = gs.reshape(-1, 1) * dqs + ge.reshape(-1, 1) * dqe
        return g[:, 0], g[:, 1], g[:, 2], g[:, 3], None, None, None, None


def tmatrix_sums(radius, axis_ratio, wavelength, mr_re, mr_im, nmax, ngauss,
                 ctx=None):
    """Differentiable (Q_sca, Q_ext) full-T sums per grain (torch autograd
    over radius, axis_ratio, mr_re, mr_im) at forced orders."""
    ctx = ctx or default_context()
    return _TMatrixSums.apply(radius, axis_ratio, mr_re, mr_im, wavelength,
                              nmax, ngauss, ctx)


def weighted_sums(values, weights, group, normalise=False, ctx=None):
    """Weighted sums over consecutive groups of rows of a device tensor
    values [n, ...] with weights [n], on the GPU: out[o] = sum_s w x (divided
    by sum_s w when normalise).  normalise=True is the shape average of the
    reference (ShapeAveragingTask.hpp:127-166); normalise=False the size
    integration sum_a w_a Q(a) its scripts do (scripts/BBOPResult_sim.py:
    9-71)."""
    ctx = ctx or default_context()
    x = torch.as_tensor(values, dtype=torch.float64, device=ctx.device)
    n = x.shape[0]
    if group < 1 or n % group:
        raise ValueError("rows must split into groups of %d" % group)
    width = int(np.prod(x.shape[1:])) if x.dim() > 1 else 1
    xin = x.reshape(n, width).contiguous()
    w = torch.as_tensor(weights, dtype=torch.float64,
                        device=ctx.device).contiguous()
    if w.numel() != n:
        raise ValueError("one weight per row")
    out = torch.zeros((n // group, width), dtype=torch.float64,
                      device=ctx.device)
    torch.cuda.current_stream(ctx.device).synchronize()
    ctx.check(ctx._lib.ctm_weighted_reduce(
        ctx.handle, n // group, group, width, _ptr(xin), _ptr(w),
        int(bool(normalise)), _ptr(out)), "ctm_weighted_reduce")
    ctx.synchronize()
    return out.reshape((n // group,) + tuple(x.shape[1:]))
