"""The CosTuuM Python module surface on the MI355X library.

Mirrors the reference's CPython extension (CTMModule.cpp:481-541): the same
names, keyword arguments, defaults, output shapes and squeeze semantics, so
that ``from costuum_amd import CosTuuM`` can stand in for ``import CosTuuM``.
T-matrices, ensemble averages and every observable are computed by the HIP
library through its C ABI (include/costuum_amd.h); this module only builds
the batches (distributions, dust callbacks, sorted axes) and reshapes the
results.

Error convention: where the reference warns and returns NULL (a Python
exception) this raises ``ValueError``; where it aborts on a numerical
failure (``ctm_error``) this raises ``CtmError``.

Lengths are in the caller's unit (the reference uses SI metres); radius and
wavelength must share it.
"""
import ctypes
import math
import sys
import warnings

import numpy as np
import torch

from . import _native
from .tmatrix import (Context, CtmError, m0_sums, observables, query,
                      tmatrix_batch, tmatrix_fixed)

# CTMModule.cpp:515-524 and DustProperties.hpp:36-47
DAVIS_GREENSTEIN_ALIGNMENT = 0
MISHCHENKO_ALIGNMENT = 1
DISABLE_ALIGNMENT = 2
CUSTOM_ALIGNMENT = 3
CARBON_PARALLEL = 0
CARBON_PERPENDICULAR = 1
SILICON = 2

# grains per library batch in get_table (bounds the T-matrix storage of one
# batch; whole shape groups are kept together)
_GRAINS_PER_BATCH = 4096


def _lib():
    return _native.lib()


class _Callback:
    """A Python function wrapped as a ctm_real_fn; an exception raised by the
    function is kept and re-raised after the library call returns."""

    def __init__(self, fn):
        self.error = None

        def call(x, _arg):
            if self.error is not None:
                return 0.0
            try:
                return float(fn(x))
            except BaseException as e:  # re-raised by check()
                self.error = e
                return 0.0
        self.c = _native.REAL_FN(call)

    def check(self):
        if self.error is not None:
            raise self.error


def gauss_legendre(order, a=-1.0, b=1.0):
    """Gauss-Legendre nodes (ascending) and weights of the given order on
    [a, b] (SpecialFunctions.hpp:647-751)."""
    x = np.zeros(order)
    w = np.zeros(order)
    rc = _lib().ctm_gauss_legendre(int(order), float(a), float(b),
                                   x.ctypes.data, w.ctypes.data)
    if rc != 0:
        raise ValueError("invalid Gauss-Legendre order or interval")
    return x, w


def adaptive_quadrature(function, a, b, start_order=10, maximum_order=1000,
                        absolute_error=1e-10, relative_error=1e-10):
    """SpecialFunctions::gauss_legendre_quadrature (SpecialFunctions.hpp:
    772-815) of a Python function of one variable."""
    cb = _Callback(function)
    out = ctypes.c_double()
    rc = _lib().ctm_adaptive_quadrature(cb.c, None, float(a), float(b),
                                        int(start_order), int(maximum_order),
                                        float(absolute_error),
                                        float(relative_error),
                                        ctypes.byref(out))
    cb.check()
    if rc == _native.CTM_NOT_CONVERGED:
        warnings.warn("Tolerance for quadrature could not be reached")
    elif rc != 0:
        raise ValueError("invalid quadrature arguments")
    return out.value


def get_equal_volume_radius(radius, axis_ratio):
    """Equal-volume radius of a spheroid given its equal-surface-area radius
    (CTMModule.cpp:66-95, SpecialFunctions.hpp:558-583)."""
    return _lib().ctm_rv_over_rs(float(axis_ratio)) * float(radius)


def _wigner_d00(cosx, order):
    """d^n_00 for n = 1..order at each cosx (SpecialFunctions.hpp:317-335);
    returns an array [order, len(cosx)]."""
    cosx = np.asarray(cosx, dtype=np.float64)
    out = np.zeros((order,) + cosx.shape)
    if order == 0:
        return out
    prev = np.ones_like(cosx)
    cur = cosx.copy()
    out[0] = cur
    for k in range(1, order):
        nxt = ((2.0 * k + 1.0) * cosx * cur - k * prev) / (k + 1.0)
        prev, cur = cur, nxt
        out[k] = cur
    return out


# ---------------------------------------------------------------------------
# orientation distributions (PyOrientationDistribution.hpp:57-670)
# ---------------------------------------------------------------------------

class OrientationDistribution:
    """Isotropic orientation distribution p(beta) = 1/2 with Legendre
    coefficients p_0..p_order computed numerically
    (OrientationDistribution.hpp:38-271).  Exported as
    RandomOrientationDistribution."""

    compute_alignment = True

    def __init__(self, order=100):
        self._coefficients = self._numeric(None, int(order), 1e-10, 1e-5)

    @staticmethod
    def _numeric(function, order, abs_err, rel_err):
        coef = np.zeros(order + 1)
        cb = _Callback(function) if function is not None else None
        fn = cb.c if cb is not None else ctypes.cast(None, _native.REAL_FN)
        rc = _lib().ctm_orientation_coefficients(fn, None, order,
                                                 float(abs_err),
                                                 float(rel_err),
                                                 coef.ctypes.data)
        if cb is not None:
            cb.check()
        if rc == _native.CTM_NOT_CONVERGED:
            warnings.warn("Tolerance for quadrature could not be reached")
        elif rc != 0:
            raise ValueError("invalid orientation distribution")
        return coef

    def get_coefficient(self, n):
        return float(self._coefficients[n])

    def get_coefficients(self):
        return self._coefficients.copy()

    def get_maximum_order(self):
        return len(self._coefficients) - 1

    def evaluate(self, theta):
        """p(theta) from the expansion: sum over n < order of
        (2n + 1)/2 p_n d^n_00(theta) (PyOrientationDistribution.hpp:
        153-199)."""
        th = np.atleast_1d(np.asarray(theta, dtype=np.float64))
        order = self.get_maximum_order()
        d = np.ones((max(order, 1),) + th.shape)
        if order > 1:
            d[1:] = _wigner_d00(np.cos(th), order - 1)
        value = np.zeros_like(th)
        for n in range(order):
            value += 0.5 * (2.0 * n + 1.0) * self._coefficients[n] * d[n]
        return np.squeeze(value)


RandomOrientationDistribution = OrientationDistribution


class MishchenkoOrientationDistribution(OrientationDistribution):
    """p(beta) = 1/2 + 5/4 p_2 (3 cos^2 beta - 1) with p_2 = 3/2 <cos^2 beta>
    - 1/2 (MishchenkoOrientationDistribution.hpp:65-122)."""

    def __init__(self, cos2beta, order=100):
        if not 0.2 <= cos2beta <= 0.6:
            raise ValueError("cos2beta must lie in [0.2, 0.6]")
        self._coefficients = np.zeros(int(order) + 1)
        self._coefficients[0] = 1.0
        if order >= 2:
            self._coefficients[2] = 1.5 * cos2beta - 0.5


class DavisGreensteinOrientationDistribution(OrientationDistribution):
    """Perfect Davis-Greenstein alignment: oblate grains (d > 1) along the
    field (p_N = 1), prolate grains perpendicular (p_2i = (-1)^i (2i-1)!!/
    (2i)!!) (DavisGreensteinOrientationDistribution.hpp:66-105)."""

    def __init__(self, axis_ratio, order=100):
        order = int(order)
        c = np.zeros(order + 1)
        if axis_ratio < 1.0:
            num, den = 1.0, 2.0
            c[0] = 1.0
            i = 1
            while 2 * i < order + 1:
                c[2 * i] = (1 if i % 2 == 0 else -1) * num / den
                num *= 2.0 * i + 1.0
                den *= 2.0 * (i + 1.0)
                i += 1
        else:
            c[:] = 1.0
        self._coefficients = c


class CustomOrientationDistribution(OrientationDistribution):
    """A user density function(beta[, additional_arguments]), normalised and
    expanded numerically with tolerances 1e-4 (PyOrientationDistribution.hpp:
    462-533,609-632)."""

    def __init__(self, function, order=100, additional_arguments=None):
        if additional_arguments is not None:
            fn = lambda b: function(b, additional_arguments)  # noqa: E731
        else:
            fn = function
        self._coefficients = self._numeric(fn, int(order), 1e-4, 1e-4)


class _DisabledAlignment(OrientationDistribution):
    """Alignment switched off: the ensemble T-matrix is the grain's own
    (DisabledAlignmentOrientationDistribution.hpp:36-81)."""

    compute_alignment = False

    def __init__(self):
        self._coefficients = np.zeros(1)


# ---------------------------------------------------------------------------
# shape distributions (PyShapeDistribution.hpp:124-300)
# ---------------------------------------------------------------------------

class ShapeDistribution:
    """Base: GL points on [d_min, d_max] weighted by the density
    (ShapeDistribution.hpp:36-140)."""

    def __call__(self, axis_ratio):
        if 0.5 <= axis_ratio <= 1.5:
            x = axis_ratio - 1.0
            return math.exp(-0.5 * x * x / 0.027777778)
        return 0.0

    def get_minimum_axis_ratio(self):
        return 0.5

    def get_maximum_axis_ratio(self):
        return 1.5

    def evaluate(self, npoints):
        x, w = gauss_legendre(npoints, self.get_minimum_axis_ratio(),
                              self.get_maximum_axis_ratio())
        self._shapes = x
        self._weights = w * np.array([self(float(d)) for d in x])

    def get_number_of_points(self):
        return len(self._shapes)

    def get_shape(self, index):
        return float(self._shapes[index])

    def get_weight(self, index):
        return float(self._weights[index])

    def get_limits(self):
        return (self.get_minimum_axis_ratio(), self.get_maximum_axis_ratio())


class SingleShapeShapeDistribution(ShapeDistribution):
    """One axis ratio with weight 1 (SingleShapeShapeDistribution.hpp:
    34-76)."""

    def __init__(self, axis_ratio):
        self._shapes = np.array([float(axis_ratio)])
        self._weights = np.array([1.0])

    def __call__(self, axis_ratio):
        return 1.0 if axis_ratio == self._shapes[0] else 0.0

    def get_minimum_axis_ratio(self):
        return float(self._shapes[0])

    def get_maximum_axis_ratio(self):
        return float(self._shapes[0])


def _dh_shape_factor(d):
    """Depolarisation factor L of a spheroid with axis ratio d
    (DraineHensleyShapeDistribution.hpp:67-84)."""
    if d < 1.0:
        e2 = 1.0 - d * d
        e = math.sqrt(e2)
        return (1.0 - e2) * (0.5 * math.log((1.0 + e) / (1.0 - e)) / e - 1.0) \
            / e2
    if d > 1.0:
        e2 = 1.0 - 1.0 / (d * d)
        e = math.sqrt(e2)
        return (1.0 - math.sqrt(1.0 - e2) * math.asin(e) / e) / e2
    return 1.0 / 3.0


def _dh_jacobian(d):
    """|dL/dd| (DraineHensleyShapeDistribution.hpp:93-121)."""
    if d < 1.0:
        e2 = 1.0 - d * d
        if e2 < 1e-5:
            return 4.0 / 15.0 + 2.0 * e2 / 21.0 + 3.0 * e2 * e2 / 70.0
        e = math.sqrt(e2)
        return 0.5 * math.sqrt(1.0 - e2) * (
            (3.0 - e2) * math.log((1.0 + e) / (1.0 - e)) - 6.0 * e) / (
            e2 * e2 * e)
    if d > 1.0:
        e2 = 1.0 - 1.0 / (d * d)
        if e2 < 1e-5:
            return 4.0 / 15.0 - 2.0 * e2 / 21.0 - 11.0 * e2 * e2 / 210.0
        e = math.sqrt(e2)
        return (1.0 - e2) * ((3.0 - 2.0 * e2) * math.asin(e) -
                             3.0 * e * math.sqrt(1.0 - e2)) / (e2 * e2 * e)
    return 4.0 / 15.0


def _dh_cde2(d):
    """CDE2 density 12 L (1 - L)^2 in L, times the Jacobian
    (DraineHensleyShapeDistribution.hpp:128-147)."""
    L = _dh_shape_factor(d)
    return 12.0 * L * (L - 1.0) ** 2 * _dh_jacobian(d)


class DraineHensleyShapeDistribution(ShapeDistribution):
    """Draine & Hensley (2017) CDE2 shape distribution, limited where the
    density drops below `cutoff`, or to the central `fraction` of the
    prolate and oblate mass (DraineHensleyShapeDistribution.hpp:35-417)."""

    def __init__(self, npoints=20, cutoff=0.15, fraction=-1.0):
        self._dmin = 1.0
        self._dmax = 1.0
        if fraction > 0.0:
            if fraction >= 1.0:
                raise ValueError("fraction must be below 1")
            self._limits_from_fraction(fraction)
        else:
            if not 0.0 < cutoff < 0.5:
                raise ValueError("cutoff must lie in (0, 0.5)")
            self._limits_from_cutoff(cutoff)
        self.evaluate(int(npoints))

    def __call__(self, axis_ratio):
        if self._dmin <= axis_ratio <= self._dmax:
            return _dh_cde2(axis_ratio)
        return 0.0

    def get_minimum_axis_ratio(self):
        return self._dmin

    def get_maximum_axis_ratio(self):
        return self._dmax

    def _limits_from_fraction(self, fraction):
        def q(a, b):
            return adaptive_quadrature(_dh_cde2, a, b, 10, 1000, 1e-10, 1e-10)
        prolate = q(0.0, 1.0)
        target = fraction * prolate
        lo, hi, half = 0.0, 1.0, 0.5
        integral = q(half, 1.0)
        while abs(integral - target) > 0.005 * abs(integral + target):
            if half <= 0.0:
                raise CtmError("shape limit search failed")
            if integral > target:
                lo = half
            else:
                hi = half
            half = 0.5 * (lo + hi)
            integral = q(half, 1.0)
        self._dmin = half
        target = fraction * (1.0 - prolate)
        lo, hi = 1.0, 2.0
        integral = q(1.0, hi)
        while integral < target:
            hi *= 2.0
            integral = q(1.0, hi)
        half = 0.5 * (lo + hi)
        integral = q(1.0, half)
        while abs(integral - target) > 1e-5 * abs(integral + target):
            if half <= 0.0:
                raise CtmError("shape limit search failed")
            if integral > target:
                hi = half
            else:
                lo = half
            half = 0.5 * (lo + hi)
            integral = q(1.0, half)
        self._dmax = half

    def _limits_from_cutoff(self, cutoff):
        # the density is evaluated inside [dmin, dmax] throughout, as the
        # reference's operator() is during its search
        val = _dh_cde2(self._dmin) - cutoff
        while val > 0.0:
            self._dmin *= 0.5
            val = _dh_cde2(self._dmin) - cutoff
        upper = 1.0
        if not _dh_cde2(upper) - cutoff > 0.0:
            raise CtmError("cutoff above the density at d = 1")
        while abs(self._dmin - upper) > 0.005 * abs(self._dmin + upper):
            mid = 0.5 * (self._dmin + upper)
            if _dh_cde2(mid) - cutoff < 0.0:
                self._dmin = mid
            else:
                upper = mid
        val = _dh_cde2(self._dmax) - cutoff
        while val > 0.0:
            self._dmax += 1.0
            val = _dh_cde2(self._dmax) - cutoff
        lower = 1.0
        while abs(self._dmax - lower) > 0.005 * abs(self._dmax + lower):
            mid = 0.5 * (lower + self._dmax)
            if _dh_cde2(mid) - cutoff > 0.0:
                lower = mid
            else:
                self._dmax = mid


# ---------------------------------------------------------------------------
# alignment (PyAlignmentDistribution.hpp:94-212)
# ---------------------------------------------------------------------------

class SizeBasedAlignmentDistribution:
    """Grains larger than `minimum_size` (equal-volume radius) are aligned
    with the oblate (d > 1) or prolate distribution of the chosen type;
    smaller ones are isotropic (SizeBasedAlignmentDistribution.hpp:
    39-147)."""

    def __init__(self, minimum_size, aligned_orientation_distribution_type,
                 maximum_order=100, oblate_orientation_distribution=None,
                 prolate_orientation_distribution=None):
        self.minimum_size = float(minimum_size)
        n = 2 * int(maximum_order)
        kind = int(aligned_orientation_distribution_type)
        self._non_aligned = OrientationDistribution(n)
        if kind == DAVIS_GREENSTEIN_ALIGNMENT:
            self._oblate = DavisGreensteinOrientationDistribution(2.0, n)
            self._prolate = DavisGreensteinOrientationDistribution(0.5, n)
        elif kind == MISHCHENKO_ALIGNMENT:
            self._oblate = MishchenkoOrientationDistribution(0.6, n)
            self._prolate = MishchenkoOrientationDistribution(0.2, n)
        elif kind == DISABLE_ALIGNMENT:
            self._oblate = _DisabledAlignment()
            self._prolate = _DisabledAlignment()
        elif kind == CUSTOM_ALIGNMENT:
            if (oblate_orientation_distribution is None or
                    prolate_orientation_distribution is None):
                raise ValueError("custom alignment needs both oblate and "
                                 "prolate orientation distributions")
            self._oblate = oblate_orientation_distribution
            self._prolate = prolate_orientation_distribution
        else:
            raise ValueError("Unknown orientation distribution!")

    def get_distribution(self, equal_volume_radius, axis_ratio):
        if equal_volume_radius > self.minimum_size:
            return self._oblate if axis_ratio > 1.0 else self._prolate
        return self._non_aligned


# ---------------------------------------------------------------------------
# dust properties (PyDustProperties.hpp:160-430)
# ---------------------------------------------------------------------------

class CustomDustProperties:
    """Refractive index from function(wavelength, grain_size, grain_type[,
    additional_arguments]) returning a complex number
    (PyDustProperties.hpp:242-357)."""

    def __init__(self, function, additional_arguments=None):
        self._function = function
        self._args = additional_arguments

    def _index(self, wavelength, grain_size, grain_type):
        if self._args is not None:
            r = self._function(float(wavelength), float(grain_size),
                               int(grain_type), self._args)
        else:
            r = self._function(float(wavelength), float(grain_size),
                               int(grain_type))
        if isinstance(r, np.ndarray):
            if r.size != 1:
                raise ValueError("Function returns an array!")
            r = r.reshape(()).item()
        if not isinstance(r, complex) and not np.iscomplexobj(r):
            raise ValueError("Function does not return a complex number!")
        return complex(r)

    def get_refractive_index(self, type, size, wavelength):
        return self._index(wavelength, size, type)


def _eps_to_mr(eps):
    """Refractive index from the dielectric function
    (DraineDustProperties.hpp:135-142)."""
    a = abs(eps)
    return complex(math.sqrt(0.5 * (a + eps.real)),
                   math.sqrt(0.5 * (a - eps.real)))


def _carbon_free_electrons(T, wavelength, grain_size, parallel):
    """Free-electron (Drude) contribution to the dielectric function of
    graphite with the size-limited collision time
    (DraineDustProperties.hpp:84-127); SI units."""
    if parallel:
        omega_p, tau_bulk, TF, vF = 1.53e14, 1.4e-14, 255.0, 3.7e4
    else:
        T2 = T * T
        omega_p = 4.33e14 * math.sqrt(1.0 - 6.24e-3 * T + 3.66e-5 * T2)
        tau_bulk = 4.2e-11 / (1.0 + 0.322 * T + 0.0013 * T2)
        TF, vF = 255.0, 4.5e5
    veff = vF * math.sqrt(1.0 + T / TF)
    tau = 1.0 / (1.0 / tau_bulk + veff / grain_size)
    omega = 2.0 * math.pi * 299792458.0 / wavelength
    wt = omega * tau
    wpt = omega_p * tau
    return complex(-wpt * wpt, 0.0) / complex(wt * wt, wt)


class _LinearTable:
    """Columns of a whitespace text table with linear interpolation in the
    first column, ascending or descending (Table.hpp:179-260)."""

    def __init__(self, path):
        rows = []
        with open(path) as f:
            for line in f:
                t = line.split()
                if not t or t[0].startswith("#"):
                    continue
                rows.append([float(v) for v in t[:5]])
        self.rows = np.array(rows)

    def row(self, x):
        col = self.rows[:, 0]
        lo, hi = 0, len(col) - 1
        xlo, xhi = col[lo], col[hi]
        if xlo == xhi:
            raise CtmError("Table endpoints are the same!")
        asc = xlo < xhi
        if not (min(xlo, xhi) <= x <= max(xlo, xhi)):
            raise ValueError("wavelength %g outside the table" % x)
        while lo + 1 != hi:
            mid = (lo + hi) >> 1
            if (col[mid] <= x) if asc else (col[mid] >= x):
                lo, xlo = mid, col[mid]
            else:
                hi, xhi = mid, col[mid]
        f = (x - xlo) / (xhi - xlo)
        return (1.0 - f) * self.rows[lo] + f * self.rows[hi]


class DraineDustProperties:
    """Draine's (2003) astrosilicate and graphite dielectric tables with the
    graphite free-electron correction for the grain temperature and size
    (DraineDustProperties.hpp:70-239).  The callindex.out_* tables are not
    shipped with this package: data_directory (or the environment variable
    COSTUUM_DRAINE_DATA) names the folder that holds callindex.out_silD03,
    callindex.out_CpaD03_0.10 and callindex.out_CpeD03_0.10.  Wavelengths
    and sizes in metres."""

    def __init__(self, dust_temperature=20.0, data_directory=None):
        import os
        d = data_directory or os.environ.get("COSTUUM_DRAINE_DATA")
        if not d:
            raise NotImplementedError(
                "DraineDustProperties needs the callindex.out tables: pass "
                "data_directory or set COSTUUM_DRAINE_DATA")
        self.dust_temperature = float(dust_temperature)
        self._sil = self._load(os.path.join(d, "callindex.out_silD03"))
        self._cpa = self._load(os.path.join(d, "callindex.out_CpaD03_0.10"),
                               True)
        self._cpe = self._load(os.path.join(d, "callindex.out_CpeD03_0.10"),
                               False)

    @staticmethod
    def _load(path, parallel=None):
        t = _LinearTable(path)
        t.rows[:, 0] *= 1e-6
        t.rows[:, 1] += 1.0
        t.rows[:, 3] += 1.0
        if parallel is not None:
            # remove the free-electron term the table holds (20 K, 0.1 um)
            for i in range(len(t.rows)):
                c = _carbon_free_electrons(20.0, t.rows[i, 0], 1e-7, parallel)
                t.rows[i, 1] -= c.real
                t.rows[i, 2] -= c.imag
        return t

    def _index(self, wavelength, grain_size, grain_type):
        if grain_type in (CARBON_PARALLEL, CARBON_PERPENDICULAR):
            par = grain_type == CARBON_PARALLEL
            row = (self._cpa if par else self._cpe).row(wavelength)
            eps = complex(row[1], row[2]) + _carbon_free_electrons(
                self.dust_temperature, wavelength, grain_size, par)
            return _eps_to_mr(eps)
        if grain_type == SILICON:
            row = self._sil.row(wavelength)
            return complex(row[3], row[4])
        raise ValueError("Unknown dust grain type: %d!" % grain_type)

    def get_refractive_index(self, type, size, wavelength):
        return self._index(wavelength, size, type)


# ---------------------------------------------------------------------------
# get_table (CTMModule.cpp:146-480, TaskManager.hpp:445-1330)
# ---------------------------------------------------------------------------

def _as_1d(a, dtype):
    return np.atleast_1d(np.asarray(a, dtype=dtype)).ravel()


def _reduce_groups(ctx, x, group, weights):
    """Weighted mean over consecutive groups of `group` rows of a device
    tensor x [n, width] (ShapeAveragingTask.hpp:127-166)."""
    n = x.shape[0]
    width = int(np.prod(x.shape[1:])) if x.dim() > 1 else 1
    out = torch.zeros((n // group, width), dtype=torch.float64,
                      device=x.device)
    w = torch.as_tensor(weights, dtype=torch.float64, device=x.device)
    xin = x.reshape(n, width).contiguous()
    torch.cuda.current_stream(x.device).synchronize()
    ctx.check(ctx._lib.ctm_weighted_reduce(
        ctx.handle, n // group, group, width, ctypes.c_void_p(xin.data_ptr()),
        ctypes.c_void_p(w.data_ptr()), 1, ctypes.c_void_p(out.data_ptr())),
        "ctm_weighted_reduce")
    ctx.synchronize()
    return out.reshape((n // group,) + tuple(x.shape[1:]))


def _memory_batches(inter, shapes, min_order, max_order, n_pairs, budget,
                    device):
    """Consecutive ranges of interactions (each with all its shapes) whose
    estimated device memory (ctm_grain_memory_estimate per grain) stays
    within min(budget, 80 % of the free device memory), at most
    _GRAINS_PER_BATCH grains each: the maximum_memory_size semantics of the
    reference's task resources (TaskManager.hpp:101-348,718-760) applied to
    the batches the GPU holds at once."""
    free, _ = torch.cuda.mem_get_info(device)
    cap = min(budget, 0.8 * free)
    est = _lib().ctm_grain_memory_estimate
    per = max(1, _GRAINS_PER_BATCH // max(len(shapes), 1))
    out, i0 = [], 0
    while i0 < len(inter):
        i1, used = i0, 0.0
        while i1 < len(inter) and i1 - i0 < per:
            a, lam, _ = inter[i1]
            b = sum(est(float(a), float(d), float(lam), min_order, max_order,
                        n_pairs) for d in shapes)
            if i1 > i0 and used + b > cap:
                break
            used += b
            i1 += 1
        out.append((i0, i1))
        i0 = i1
    return out


def get_table(types, sizes, wavelengths, thetas, shape_distribution,
              alignment_distribution, dust_properties, minimum_order=10,
              maximum_order=100, gauss_legendre_factor=2, tolerance=1e-4,
              number_of_quadrature_angles=20, maximum_memory_size=10000000000,
              number_of_threads=4, quicksched_graph_log=None,
              quicksched_task_log=None, quicksched_task_type_log=None,
              memory_log=None, do_absorption=True, do_extinction=False,
              do_scattering=False, account_for_scattering=False,
              verbose=False):
    """Shape- and orientation-averaged absorption/extinction coefficients
    and scattering matrices on a (composition, size, wavelength, theta)
    grid (CTMModule.cpp:146-480).

    Returns squeezed float64 arrays: [ncomp, nsize, nlambda, ntheta,
    (Qabs, Qabspol)? + (Qext, Qextpol, Qextcpol)?] and/or [ncomp, nsize,
    nlambda, ntheta, 2 ntheta, 4, 4] (both, as a tuple, when both are asked
    for).  Sizes and wavelengths are sorted ascending first
    (TaskManager.hpp:457-459).  maximum_memory_size bounds the device memory
    of each batch of grains (with 80 % of the free device memory); the grid
    streams through the GPU in as many batches as that takes.
    number_of_threads has no GPU meaning and is accepted for compatibility;
    the QuickSched logs are replaced by a per-kernel timing log written to
    quicksched_task_log when it is given."""
    comps = _as_1d(types, np.int64)
    size_arr = np.sort(_as_1d(sizes, np.float64))
    lam_arr = np.sort(_as_1d(wavelengths, np.float64))
    th = _as_1d(thetas, np.float64)
    if not (do_absorption or do_extinction or do_scattering):
        return None
    for t in comps:
        if not 0 <= t < 3:
            raise ValueError("unknown grain type %d" % t)
    nshape = shape_distribution.get_number_of_points()
    shapes = [shape_distribution.get_shape(i) for i in range(nshape)]
    weights = [shape_distribution.get_weight(i) for i in range(nshape)]
    # one interaction per (composition, size, wavelength), in the
    # reference's loop order, with the dust callback called once each
    inter = []
    for t in comps:
        for a in size_arr:
            for lam in lam_arr:
                mr = dust_properties._index(lam, a, int(t))
                inter.append((a, lam, mr))
    ninter = len(inter)
    nth = len(th)
    ncol = 2 * int(bool(do_absorption)) + 3 * int(bool(do_extinction))
    q_out = np.zeros((ninter, nth, ncol)) if ncol else None
    z_out = np.zeros((ninter, nth, 2 * nth, 4, 4)) if do_scattering else None
    ctx = Context(0)
    if quicksched_task_log:
        ctx.profile(True)
    chunks = _memory_batches(inter, shapes, int(minimum_order),
                             int(maximum_order),
                             nth * (1 + (int(number_of_quadrature_angles)
                                         if account_for_scattering else 0)) +
                             (nth if do_scattering else 0),
                             float(maximum_memory_size), ctx.device)
    try:
        for i0, i1 in chunks:
            chunk = inter[i0:i1]
            rv = np.repeat([c[0] for c in chunk], nshape)
            lam = np.repeat([c[1] for c in chunk], nshape)
            mre = np.repeat([c[2].real for c in chunk], nshape)
            mim = np.repeat([c[2].imag for c in chunk], nshape)
            dd = np.tile(shapes, len(chunk))
            n = len(rv)
            if verbose:
                print("get_table: grains %d..%d of %d" % (
                    i0 * nshape, i0 * nshape + n, ninter * nshape),
                    file=sys.stderr, flush=True)
            res = tmatrix_batch(rv, dd, lam, mre, mim,
                                min_order=int(minimum_order),
                                max_order=int(maximum_order),
                                glfac=int(gauss_legendre_factor),
                                tolerance=float(tolerance), ctx=ctx)
            nmax = res.nmax.cpu().numpy()
            _apply_alignment(res, alignment_distribution, rv, dd, nmax)
            obs = observables(th, ctx=ctx, do_extinction=bool(do_extinction),
                              do_absorption=bool(do_absorption),
                              account_for_scattering=bool(
                                  account_for_scattering),
                              ngauss=int(number_of_quadrature_angles),
                              n_scatter_theta=nth if do_scattering else 0,
                              scatter_gl=True, n_grains=n)
            cols = []
            if do_absorption:
                cols += [obs["qabs"], obs["qabspol"]]
            if do_extinction:
                cols += [obs["qext"], obs["qextpol"], obs["qextcpol"]]
            wts = np.tile(weights, len(chunk))
            if cols:
                q = torch.stack(cols, dim=2)
                q_out[i0:i0 + len(chunk)] = _reduce_groups(
                    ctx, q, nshape, wts).cpu().numpy()
            if do_scattering:
                z_out[i0:i0 + len(chunk)] = _reduce_groups(
                    ctx, obs["Z"], nshape, wts).cpu().numpy()
        if quicksched_task_log:
            with open(quicksched_task_log, "w") as f:
                f.write("# kernel\tlaunches\tmilliseconds\n")
                for k, v in ctx.profile_stats().items():
                    f.write("%s\t%d\t%g\n" % (k, v[0], v[1]))
    finally:
        ctx.close()
    dims = (len(comps), len(size_arr), len(lam_arr))
    first = second = None
    if q_out is not None:
        first = np.squeeze(q_out.reshape(dims + (nth, ncol)))
    if z_out is not None:
        second = np.squeeze(z_out.reshape(dims + (nth, 2 * nth, 4, 4)))
    if first is not None and second is not None:
        return first, second
    return first if first is not None else second


def _apply_alignment(res, alignment, rv, dd, nmax):
    """Ensemble-average each grain over the orientation distribution the
    alignment distribution assigns to its (size, shape)
    (TaskManager.hpp:1112-1120, AlignmentAverageTask.hpp:115-215)."""
    rows, index, dist = [], {}, np.zeros(len(rv), dtype=np.int32)
    for g in range(len(rv)):
        od = alignment.get_distribution(float(rv[g]), float(dd[g]))
        if not od.compute_alignment:
            dist[g] = -1
            continue
        key = id(od)
        if key not in index:
            index[key] = len(rows)
            rows.append(od.get_coefficients())
        if len(rows[index[key]]) < 2 * int(nmax[g]) + 1:
            raise CtmError("orientation distribution of order %d is too "
                           "small for nmax %d" % (len(rows[index[key]]) - 1,
                                                  int(nmax[g])))
        dist[g] = index[key]
    if not rows:
        return  # alignment disabled for every grain: T is used as it is
    ncoef = max(max(len(r) for r in rows), 2 * int(nmax.max()) + 1)
    pN = np.zeros((len(rows), ncoef))
    for i, r in enumerate(rows):
        pN[i, :len(r)] = r
    res.average(pN, dist)


# ---------------------------------------------------------------------------
# TMatrix (PyTMatrix.hpp:118-748)
# ---------------------------------------------------------------------------

_WINDOW = 8  # candidate orders evaluated per library call


def _calculate_tmatrix(ctx, ratio_of_radii, axis_ratio, axi, wavelength,
                       maximum_order, tolerance, ndgs, mr, maximum_ngauss):
    """TMatrixCalculator::calculate_TMatrix (TMatrixCalculator.hpp:64-213):
    nmax from max(4, x + 4.05 x^(1/3)) upward at ngauss = ndgs nmax until
    both m = 0 sums change by at most the tolerance, then ngauss upward
    from ndgs nmax + 1; then every m block.  Candidates are evaluated in
    windows on the GPU and walked in order on the host."""
    if abs(ratio_of_radii - 1.0) > 1e-8:
        ratio_of_radii = _lib().ctm_rv_over_rs(float(axis_ratio))
    rv = ratio_of_radii * axi
    xev = 2.0 * math.pi * rv / wavelength
    nmax = int(max(4.0, xev + 4.05 * np.cbrt(xev)))
    args = (rv, axis_ratio, wavelength, mr.real, mr.imag)
    old_s = old_e = 0.0
    dext = dsca = 1.0

    def going(n, limit):
        return n < limit and (dext > tolerance or dsca > tolerance)

    while going(nmax, maximum_order):
        cand = list(range(nmax, min(nmax + _WINDOW, maximum_order)))
        qs, qe = m0_sums(*[[a] * len(cand) for a in args], cand,
                         [ndgs * c for c in cand], ctx=ctx)
        for s, e in zip(qs, qe):
            dsca = abs((old_s - s) / s)
            dext = abs((old_e - e) / e)
            old_s, old_e = s, e
            nmax += 1
            if not going(nmax, maximum_order):
                break
    if nmax == maximum_order:
        raise CtmError("Unable to converge!")
    nmax -= 1
    ngauss = nmax * ndgs + 1
    dext = dsca = tolerance + 1.0
    while going(ngauss, maximum_ngauss):
        cand = list(range(ngauss, min(ngauss + _WINDOW, maximum_ngauss)))
        qs, qe = m0_sums(*[[a] * len(cand) for a in args],
                         [nmax] * len(cand), cand, ctx=ctx)
        for s, e in zip(qs, qe):
            dsca = abs((old_s - s) / s)
            dext = abs((old_e - e) / e)
            old_s, old_e = s, e
            ngauss += 1
            if not going(ngauss, maximum_ngauss):
                break
    if ngauss == maximum_ngauss:
        raise CtmError("Unable to converge!")
    ngauss -= 1
    res = tmatrix_fixed([rv], [axis_ratio], [wavelength], [mr.real],
                        [mr.imag], nmax, ngauss, ctx=ctx)
    return res, rv, nmax, ngauss


def _angles(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim > 1:
        raise ValueError("Wrong shape for input array!")
    return np.atleast_1d(a)


class TMatrix:
    """A single converged (and optionally orientation-averaged) T-matrix
    (PyTMatrix.hpp:118-748).

    cos2beta < 0 keeps the fixed-orientation T-matrix; 0 or 1 selects
    Davis-Greenstein alignment, 1/3 the isotropic distribution and any other
    value a Mishchenko distribution with that <cos^2 beta>."""

    def __init__(self, particle_radius, axis_ratio, wavelength,
                 refractive_index, cos2beta=1.0 / 3.0, tolerance=1e-4,
                 maximum_order=200, gauss_legendre_factor=2,
                 maximum_ngauss=500, is_equal_volume_radius=True):
        self._ctx = Context(0)
        mr = complex(refractive_index)
        ratio = 1.0 if is_equal_volume_radius else 0.1
        self._res, self._rv, self._nmax, self._ngauss = _calculate_tmatrix(
            self._ctx, ratio, float(axis_ratio), float(particle_radius),
            float(wavelength), int(maximum_order), float(tolerance),
            int(gauss_legendre_factor), mr, int(maximum_ngauss))
        self._wavelength = float(wavelength)
        self._is_averaged = cos2beta >= 0.0
        if cos2beta >= 0.0:
            order = 2 * self._nmax
            if cos2beta == 0.0 or cos2beta == 1.0:
                od = DavisGreensteinOrientationDistribution(axis_ratio, order)
            elif cos2beta == 1.0 / 3.0:
                od = OrientationDistribution(order)
            else:
                od = MishchenkoOrientationDistribution(cos2beta, order)
            self._res.average(od.get_coefficients()[None, :])

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None

    def get_nmax(self):
        return self._nmax

    def get_ngauss(self):
        return self._ngauss

    def get_tmatrix(self):
        """The (ensemble) T-matrix in the reference block layout."""
        if self._is_averaged:
            return self._res.averaged_tmatrix(0)
        return self._res.tmatrix(0)

    def _cabs(self, theta, ngauss):
        obs = observables(theta, ctx=self._ctx, do_extinction=False,
                          do_absorption=True, account_for_scattering=True,
                          ngauss=int(ngauss), n_grains=1)
        norm = math.pi * self._rv * self._rv
        return (obs["qabs"][0].cpu().numpy() * norm,
                obs["qabspol"][0].cpu().numpy() * norm)

    def get_absorption_cross_section(self, theta, ngauss=100):
        """C_abs(theta) with the ngauss x ngauss scattered-light quadrature
        (TMatrix.hpp:2229-2251)."""
        th = _angles(theta)
        c, _ = self._cabs(th, ngauss)
        return c

    def get_absorption_cross_sections(self, theta, ngauss=100):
        """[C_abs, C_abspol] per theta (TMatrix.hpp:2194-2222), squeezed."""
        th = _angles(theta)
        c, cp = self._cabs(th, ngauss)
        return np.squeeze(np.stack([c, cp], axis=1))

    def get_average_absorption_cross_section(self):
        """Half the GL(20)-weighted sum of C_abs over incidence angles
        (TMatrix.hpp:2261-2287)."""
        x, w = gauss_legendre(20)
        c, _ = self._cabs(np.arccos(x), 20)
        return float(0.5 * np.dot(c, w))

    def get_extinction_matrix(self, theta, phi, alpha=0.0, beta=0.0):
        """K [ntheta, nphi, 4, 4] for a particle at (alpha, beta)
        (TMatrix.hpp:1972-2010), squeezed."""
        th, ph = _angles(theta), _angles(phi)
        T, P = np.meshgrid(th, ph, indexing="ij")
        _, _, K = query(self._ctx, 0, float(alpha), float(beta), T, P, T, P)
        return np.squeeze(K)

    def get_scattering_matrix(self, theta_in, phi_in, theta_out, phi_out,
                              alpha=0.0, beta=0.0):
        """Z [ntheta_in, nphi_in, ntheta_out, nphi_out, 4, 4] for a particle
        at (alpha, beta) (TMatrix.hpp:1891-1952), squeezed."""
        a = [_angles(x) for x in (theta_in, phi_in, theta_out, phi_out)]
        g = np.meshgrid(*a, indexing="ij")
        _, Z, _ = query(self._ctx, 0, float(alpha), float(beta), *g)
        return np.squeeze(Z)
