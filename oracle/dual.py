"""Forward-mode (dual-number) restatement of the oracle's fixed-order
T-matrix chain: T of every m block and its exact derivative with respect to
one of (R_V, d, Re m, Im m), for the derivative tests.

Test infrastructure only, like the rest of oracle/: nothing in the
costuum_amd package imports it.

Every quantity that depends on the parameters is carried as a pair
(value, tangent) through the same steps as oracle/ctm_oracle.c, which
restates the reference:
  * geometry r^2(theta), (1/r) dr/dtheta (SpecialFunctions.hpp:873-893);
  * j_n(kr), y_n(kr) and j_n(m kr) with [z f_n(z)]'/z by the downward ratio
    recurrence from n = 800 and the forward recurrences
    (SpecialFunctions.hpp:107-140, 202-240);
  * the Q/RgQ assembly sums (TMatrixResource.hpp:601-709, 927-1100);
  * T = -RgQ Q^-1 (TMatrixResource.hpp:712-730, 1102-1119), whose tangent
    is taken from the exact identity d(Q^-1) = -Q^-1 dQ Q^-1.
Gauss-Legendre nodes and Wigner d values do not depend on the parameters
and come from the C oracle.  The value part therefore equals the oracle's
T to rounding (the sums over nodes and the inverse are evaluated in another
order), and the tangent part is the derivative of the same function to
rounding, with no step-size error.  Both are checked in
tests/test_oracle_dual.py against the C oracle: the values directly, the
tangents against Richardson-extrapolated central differences.
"""
import numpy as np

from . import gauss_legendre, tmatrix_offset, tmatrix_size, wigner_dn_0m

BESSEL_NSTART = 800  # start of the downward ratio recurrence (ctm_oracle.c)

# A planted mistake for the tests that show the checks have teeth: the
# named tangent is dropped.  None in normal use.
_PLANT = None


class Dual:
    """A value and its tangent along one direction (numpy arrays or
    scalars, real or complex)."""
    __slots__ = ("v", "d")
    # numpy arrays on the left of an operator defer to the Dual methods
    __array_ufunc__ = None

    def __init__(self, v, d=None):
        self.v = v
        self.d = (v * 0) if d is None else d

    def __add__(self, o):
        o = _lift(o)
        return Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = _lift(o)
        return Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return _lift(o) - self

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        o = _lift(o)
        return Dual(self.v * o.v, self.d * o.v + self.v * o.d)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = _lift(o)
        return Dual(self.v / o.v, (self.d * o.v - self.v * o.d) / (o.v * o.v))

    def __rtruediv__(self, o):
        return _lift(o) / self

    def __getitem__(self, k):
        return Dual(self.v[k], self.d[k])


def _lift(x):
    return x if isinstance(x, Dual) else Dual(x, x * 0)


def dsin(x):
    x = _lift(x)
    return Dual(np.sin(x.v), np.cos(x.v) * x.d)


def dcos(x):
    x = _lift(x)
    return Dual(np.cos(x.v), -np.sin(x.v) * x.d)


def dsqrt(x):
    x = _lift(x)
    s = np.sqrt(x.v)
    return Dual(s, x.d / (2. * s))


def dcbrt(x):
    x = _lift(x)
    c = np.cbrt(x.v)
    return Dual(c, x.d / (3. * c * c))


def sph_j(nmax, z):
    """j_n(z) and [z j_n(z)]'/z for n = 1..nmax (lists of Duals over the
    nodes): downward ratio recurrence from n = 800, then forward products
    (ctm_oracle.c ctmo_sph_j_real / sph_j_cplx; SpecialFunctions.hpp:
    202-240)."""
    zinv = 1. / z
    rho = [None] * BESSEL_NSTART
    rho[BESSEL_NSTART - 1] = z / (2. * BESSEL_NSTART + 1.)
    for idx in range(BESSEL_NSTART - 1, 0, -1):
        rho[idx - 1] = 1. / ((2. * idx + 1.) * zinv - rho[idx])
    j0 = dsin(z) * zinv
    j = [rho[0] * j0]
    dj = [j0 - j[0] * zinv]
    for i in range(1, nmax):
        j.append(rho[i] * j[i - 1])
        dj.append(j[i - 1] - (i + 1.) * zinv * j[i])
    return j, dj


def sph_y(nmax, z):
    """y_n(x) and [x y_n(x)]'/x for n = 1..nmax by the forward recurrence
    from the closed forms of y_1, y_2 (ctm_oracle.c ctmo_sph_y_real;
    SpecialFunctions.hpp:107-140)."""
    c = dcos(z)
    s = dsin(z)
    zi = 1. / z
    zi2 = zi * zi
    zi3 = zi2 * zi
    y0 = -c * zi2 - s * zi
    y1 = (-3. * zi3 + zi) * c - 3. * zi2 * s
    y = [y0]
    dy = [-zi * (c + y0)]
    if nmax > 1:
        y.append(y1)
        dy.append(y0 - 2. * zi * y1)
    for i in range(2, nmax):
        y.append((2. * i + 1.) * zi * y[i - 1] - y[i - 2])
        dy.append(y[i - 1] - (i + 1.) * zi * y[i])
    return y, dy


def _stack(lst):
    """List over n of Duals over nodes -> Dual [n][node]."""
    return Dual(np.stack([e.v for e in lst]), np.stack([e.d for e in lst]))


def _ann(n):
    return np.sqrt((2. * n + 1.) / (n * (n + 1.)))


def tmatrix_all_dual(rv, d, lam, mr, nmax, ngauss, direction):
    """T of every m block at (nmax, ngauss) in the oracle's block layout and
    its derivative along direction 0..3 = (R_V, d, Re m, Im m)
    (ctm_oracle.c ctmo_tmatrix_all).  Returns (T, dT), complex arrays of
    tmatrix_size(nmax) elements."""
    rv = Dual(float(rv), 1. if direction == 0 else 0.)
    dd = Dual(float(d), 1. if direction == 1 else 0.)
    mr = complex(mr)
    mrd = Dual(mr, 1. + 0j if direction == 2 else (1j if direction == 3
                                                   else 0j))
    k = 2. * np.pi / lam
    k2 = k * k
    k2mr = mrd * k2
    kmr = mrd * k
    # quadrature and geometry on the first ngauss nodes of the 2 ngauss
    # rule (ParticleGeometryResource.hpp:135-149)
    x, w = gauss_legendre(2 * ngauss)
    x, w = x[:ngauss], w[:ngauss]
    a = dcbrt(dd)
    a2 = a * a
    d2 = dd * dd
    c2 = x * x
    s2 = (1. - x) * (1. + x)
    r2a2 = 1. / (s2 + d2 * c2)
    r2 = a2 * r2a2
    dr = r2a2 * (x * np.sqrt(s2)) * (d2 - 1.)
    if _PLANT == "dr":
        dr = Dual(dr.v)
    s2i = 1. / s2
    sinv = np.sqrt(s2i)
    r = dsqrt(r2)
    kr = k * r * rv
    krmr = kmr * r * rv
    krinv = 1. / kr
    krmrinv = 1. / krmr
    jr, djr = sph_j(nmax, kr)
    yr, dyr = sph_y(nmax, kr)
    jm, djm = sph_j(nmax, krmr)
    jr, djr, yr, dyr = _stack(jr), _stack(djr), _stack(yr), _stack(dyr)
    jm, djm = _stack(jm), _stack(djm)
    h = Dual(jr.v + 1j * yr.v, jr.d + 1j * yr.d)
    dh = Dual(djr.v + 1j * dyr.v, djr.d + 1j * dyr.d)
    wr2 = r2 * w
    size = tmatrix_size(nmax)
    T = np.zeros(size, dtype=np.complex128)
    dT = np.zeros(size, dtype=np.complex128)
    for m in range(nmax + 1):
        nmin = 1 if m == 0 else m
        nm = nmax if m == 0 else nmax + 1 - m
        n = np.arange(nmin, nmax + 1, dtype=np.float64)
        nn1 = n * (n + 1.)
        wd = np.zeros((nmax, ngauss))
        dwd = np.zeros((nmax, ngauss))
        for ig in range(ngauss):
            wd[:, ig], dwd[:, ig] = wigner_dn_0m(x[ig], nmax, m)
        sel = slice(nmin - 1, nmax)
        # [n1, n2, node] broadcasts: n1 on axis 0, n2 on axis 1
        w1, dw1 = wd[sel][:, None, :], dwd[sel][:, None, :]
        w2, dw2 = wd[sel][None, :, :], dwd[sel][None, :, :]
        jk, djk = jr[sel][:, None, :], djr[sel][:, None, :]
        h1, dh1 = h[sel][:, None, :], dh[sel][:, None, :]
        jmn2, djmn2 = jm[sel][None, :, :], djm[sel][None, :, :]
        a12 = (_ann(n)[:, None] * _ann(n)[None, :])[:, :, None]
        n1n1p1 = nn1[:, None, None]
        n2n2p1 = nn1[None, :, None]
        even = ((np.arange(nm)[:, None] + np.arange(nm)[None, :]) % 2) == 0
        c1 = jmn2 * jk
        b1 = jmn2 * h1
        c2_ = jmn2 * djk
        b2 = jmn2 * dh1
        c4 = djmn2 * jk
        b4 = djmn2 * h1
        Qv = np.zeros((2 * nm, 2 * nm), dtype=np.complex128)
        Qd = np.zeros_like(Qv)
        Rv = np.zeros_like(Qv)
        Rd = np.zeros_like(Qv)
        i1 = np.arange(nm)[:, None]
        i2 = np.arange(nm)[None, :]

        def put(M_v, M_d, rows, cols, val, mask):
            s = Dual(np.sum(val.v, axis=2), np.sum(val.d, axis=2))
            M_v[rows, cols] += np.where(mask, s.v, 0.)
            M_d[rows, cols] += np.where(mask, s.d, 0.)

        if m > 0:
            # off-diagonal blocks J11/J22 (TMatrixResource.hpp:993-1044)
            dsi = wr2 * (sinv * float(m))
            c6 = djmn2 * djk
            b6 = djmn2 * dh1
            c7 = c4 * krinv
            b7 = b4 * krinv
            c8 = c2_ * krmrinv
            b8 = b2 * krmrinv
            e1 = dsi * (w1 * dw2 + dw1 * w2)
            f = dsi * dr * (w1 * w2)
            e2 = f * n1n1p1
            e3 = f * n2n2p1
            J11 = 1j * a12 * (e1 * b1)
            RgJ11 = 1j * a12 * (e1 * c1)
            J22 = 1j * a12 * (e1 * b6 + e2 * b7 + e3 * b8)
            RgJ22 = 1j * a12 * (e1 * c6 + e2 * c7 + e3 * c8)
            odd = ~even
            put(Qv, Qd, i1, i2 + nm, k2mr * J11 + k2 * J22, odd)
            put(Rv, Rd, i1, i2 + nm, k2mr * RgJ11 + k2 * RgJ22, odd)
            put(Qv, Qd, i1 + nm, i2, k2mr * J22 + k2 * J11, odd)
            put(Rv, Rd, i1 + nm, i2, k2mr * RgJ22 + k2 * RgJ11, odd)
        # diagonal blocks J12/J21 (TMatrixResource.hpp:623-705, 1045-1097)
        c3 = krinv * c1
        b3 = krinv * b1
        c5 = c1 * krmrinv
        b5 = b1 * krmrinv
        dssi = s2i * float(m * m)
        f1 = wr2 * ((w1 * w2) * dssi + dw1 * dw2)
        f2 = wr2 * dr * n1n1p1 * (w1 * dw2)
        f3 = wr2 * dr * n2n2p1 * (dw1 * w2)
        J12 = -a12 * (f1 * b2 + f2 * b3)
        RgJ12 = -a12 * (f1 * c2_ + f2 * c3)
        J21 = a12 * (f1 * b4 + f3 * b5)
        RgJ21 = a12 * (f1 * c4 + f3 * c5)
        put(Qv, Qd, i1, i2, k2mr * J21 + k2 * J12, even)
        put(Rv, Rd, i1, i2, k2mr * RgJ21 + k2 * RgJ12, even)
        put(Qv, Qd, i1 + nm, i2 + nm, k2mr * J12 + k2 * J21, even)
        put(Rv, Rd, i1 + nm, i2 + nm, k2mr * RgJ12 + k2 * RgJ21, even)
        # T = -RgQ Q^-1 and its tangent
        Qi = np.linalg.inv(Qv)
        dQi = -Qi @ Qd @ Qi
        Tm = -(Rv @ Qi)
        dTm = -(Rd @ Qi + Rv @ dQi)
        off = tmatrix_offset(nmax, m)
        T[off:off + 4 * nm * nm] = Tm.ravel()
        dT[off:off + 4 * nm * nm] = dTm.ravel()
    return T, dT
