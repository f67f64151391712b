"""ctypes bindings of the CPU FP64 oracle (liboracle.so).

Test infrastructure only: the parity checker for the HIP path and the CPU
baseline of bench.py.  Nothing in the costuum_amd package imports this.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


_D = ctypes.c_double
_U = ctypes.c_uint
_I = ctypes.c_int
_PD = ctypes.POINTER(ctypes.c_double)
_PU = ctypes.POINTER(ctypes.c_uint)
_PI = ctypes.POINTER(ctypes.c_int)


def _declare(L):
    L.ctmo_gauss_legendre.argtypes = [_U, _PD, _PD]
    L.ctmo_gauss_legendre_ab.argtypes = [_U, _D, _D, _PD, _PD]
    L.ctmo_rv_over_rs.argtypes = [_D]
    L.ctmo_rv_over_rs.restype = _D
    L.ctmo_r_dr_spheroid.argtypes = [_PD, _U, _D, _PD, _PD]
    L.ctmo_sph_j_real.argtypes = [_U, _D, _PD, _PD]
    L.ctmo_sph_j_cplx.argtypes = [_U, _PD, _PD, _PD]
    L.ctmo_sph_y_real.argtypes = [_U, _D, _PD, _PD]
    L.ctmo_wigner_dn_0m.argtypes = [_D, _U, _U, _PD, _PD]
    L.ctmo_wigner_dn_0m_sinx.argtypes = [_D, _D, _D, _U, _U, _PD, _PD]
    L.ctmo_plu_inverse.argtypes = [_U, _PD]
    L.ctmo_clebsch_gordan.argtypes = [_I] * 6
    L.ctmo_clebsch_gordan.restype = _D
    L.ctmo_tmatrix_size.argtypes = [_U]
    L.ctmo_tmatrix_size.restype = ctypes.c_size_t
    L.ctmo_tmatrix_offset.argtypes = [_U, _U]
    L.ctmo_tmatrix_offset.restype = ctypes.c_size_t
    L.ctmo_tmatrix_m.argtypes = [_D, _D, _D, _D, _D, _U, _U, _U, _PD]
    L.ctmo_tmatrix_all.argtypes = [_D, _D, _D, _D, _D, _U, _U, _PD]
    L.ctmo_m0_qsums.argtypes = [_U, _PD, _PD, _PD]
    L.ctmo_qcoeffs.argtypes = [_U, _PD, _PD, _PD]
    L.ctmo_converge_task.argtypes = [_D, _D, _D, _D, _D, _U, _U, _U, _D,
                                     _PU, _PU, _PD, _PD]
    L.ctmo_converge_task.restype = _I
    L.ctmo_converge_calculator.argtypes = [_D, _D, _D, _D, _U, _D, _U, _D, _D,
                                           _U, _PU, _PU, _PD]
    L.ctmo_converge_calculator.restype = _I
    L.ctmo_batch_task.argtypes = [_I, _PD, _PD, _PD, _PD, _PD, _U, _U, _U, _D,
                                  _I, _PU, _PI, _PD, _PD]


def _p(a):
    return a.ctypes.data_as(_PD)


def gauss_legendre(order, a=None, b=None):
    x = np.zeros(order)
    w = np.zeros(order)
    if a is None:
        lib().ctmo_gauss_legendre(order, _p(x), _p(w))
    else:
        lib().ctmo_gauss_legendre_ab(order, a, b, _p(x), _p(w))
    return x, w


def rv_over_rs(d):
    return lib().ctmo_rv_over_rs(d)


def r_dr_spheroid(costheta, d):
    costheta = np.ascontiguousarray(costheta, dtype=np.float64)
    r2 = np.zeros_like(costheta)
    dr = np.zeros_like(costheta)
    lib().ctmo_r_dr_spheroid(_p(costheta), len(costheta), d, _p(r2), _p(dr))
    return r2, dr


def sph_j(nmax, z):
    if isinstance(z, complex):
        zz = np.array([z.real, z.imag])
        j = np.zeros(nmax, dtype=np.complex128)
        dj = np.zeros(nmax, dtype=np.complex128)
        lib().ctmo_sph_j_cplx(nmax, _p(zz), _p(j), _p(dj))
        return j, dj
    j = np.zeros(nmax)
    dj = np.zeros(nmax)
    lib().ctmo_sph_j_real(nmax, float(z), _p(j), _p(dj))
    return j, dj


def sph_y(nmax, z):
    y = np.zeros(nmax)
    dy = np.zeros(nmax)
    lib().ctmo_sph_y_real(nmax, float(z), _p(y), _p(dy))
    return y, dy


def wigner_dn_0m(cosx, nmax, m):
    y = np.zeros(nmax)
    dy = np.zeros(nmax)
    lib().ctmo_wigner_dn_0m(cosx, nmax, m, _p(y), _p(dy))
    return y, dy


def wigner_dn_0m_sinx(cosx, sinx, sinx_inv, nmax, m):
    y = np.zeros(nmax)
    dy = np.zeros(nmax)
    lib().ctmo_wigner_dn_0m_sinx(cosx, sinx, sinx_inv, nmax, m, _p(y), _p(dy))
    return y, dy


def plu_inverse(A):
    A = np.ascontiguousarray(A, dtype=np.complex128).copy()
    lib().ctmo_plu_inverse(A.shape[0], _p(A))
    return A


def clebsch_gordan(n1, m1, n2, m2, N, M):
    return lib().ctmo_clebsch_gordan(n1, m1, n2, m2, N, M)


def tmatrix_size(nmax):
    return lib().ctmo_tmatrix_size(nmax)


def tmatrix_offset(nmax, m):
    return lib().ctmo_tmatrix_offset(nmax, m)


def tmatrix_m(rv, d, lam, mr, nmax, ngauss, m):
    nm = nmax if m == 0 else nmax + 1 - m
    T = np.zeros((2 * nm, 2 * nm), dtype=np.complex128)
    lib().ctmo_tmatrix_m(rv, d, lam, mr.real, mr.imag, nmax, ngauss, m, _p(T))
    return T


def tmatrix_all(rv, d, lam, mr, nmax, ngauss):
    T = np.zeros(tmatrix_size(nmax), dtype=np.complex128)
    lib().ctmo_tmatrix_all(rv, d, lam, mr.real, mr.imag, nmax, ngauss, _p(T))
    return T


def tmatrix_blocks(T, nmax):
    """Split a flat full T-matrix into its list of m blocks."""
    out = []
    for m in range(nmax + 1):
        nm = nmax if m == 0 else nmax + 1 - m
        o = tmatrix_offset(nmax, m)
        out.append(T[o:o + 4 * nm * nm].reshape(2 * nm, 2 * nm))
    return out


def m0_qsums(nmax, T0):
    T0 = np.ascontiguousarray(T0, dtype=np.complex128)
    s = ctypes.c_double()
    e = ctypes.c_double()
    lib().ctmo_m0_qsums(nmax, _p(T0), ctypes.byref(s), ctypes.byref(e))
    return s.value, e.value


def qcoeffs(nmax, T):
    T = np.ascontiguousarray(T, dtype=np.complex128)
    s = ctypes.c_double()
    e = ctypes.c_double()
    lib().ctmo_qcoeffs(nmax, _p(T), ctypes.byref(s), ctypes.byref(e))
    return s.value, e.value


def converge_task(rv, d, lam, mr, min_order=4, max_order=100, glfac=2,
                  tol=1e-4):
    n = ctypes.c_uint()
    g = ctypes.c_uint()
    dqs = ctypes.c_double()
    dqe = ctypes.c_double()
    st = lib().ctmo_converge_task(rv, d, lam, mr.real, mr.imag, min_order,
                                  max_order, glfac, tol, ctypes.byref(n),
                                  ctypes.byref(g), ctypes.byref(dqs),
                                  ctypes.byref(dqe))
    return st, n.value, g.value, dqs.value, dqe.value


def converge_calculator(ratio_of_radii, d, axi, lam, mr, maximum_order=200,
                        tol=1e-4, ndgs=2, maximum_ngauss=500):
    n = ctypes.c_uint()
    g = ctypes.c_uint()
    rv = ctypes.c_double()
    st = lib().ctmo_converge_calculator(ratio_of_radii, d, axi, lam,
                                        maximum_order, tol, ndgs, mr.real,
                                        mr.imag, maximum_ngauss,
                                        ctypes.byref(n), ctypes.byref(g),
                                        ctypes.byref(rv))
    return st, n.value, g.value, rv.value


def batch_task(rv, d, lam, mr_re, mr_im, min_order=4, max_order=100,
               glfac=2, tol=1e-4, nthreads=0):
    rv, d, lam, mr_re, mr_im = [np.ascontiguousarray(a, dtype=np.float64)
                                for a in (rv, d, lam, mr_re, mr_im)]
    n = len(rv)
    nmax = np.zeros(n, dtype=np.uint32)
    status = np.zeros(n, dtype=np.int32)
    qs = np.zeros(n)
    qe = np.zeros(n)
    lib().ctmo_batch_task(n, _p(rv), _p(d), _p(lam), _p(mr_re), _p(mr_im),
                          min_order, max_order, glfac, tol, nthreads,
                          nmax.ctypes.data_as(_PU),
                          status.ctypes.data_as(_PI), _p(qs), _p(qe))
    return nmax, status, qs, qe


def _declare_obs(L):
    L.ctmo_extinction.argtypes = [_U, _PD, _D, _D, _U, _PD, _PD, _PD, _PD]
    L.ctmo_absorption.argtypes = [_U, _PD, _D, _D, _U, _PD, _U, _I, _PD, _PD]
    L.ctmo_scattering_grid.argtypes = [_U, _PD, _D, _U, _I, _PD]
    L.ctmo_orientation_coefficients.argtypes = [_I, _D, _U, _PD]
    L.ctmo_orientation_average.argtypes = [_U, _PD, _PD, _PD]


_obs_declared = False


def _olib():
    global _obs_declared
    L = lib()
    if not _obs_declared:
        _declare_obs(L)
        _obs_declared = True
    return L


def extinction(nmax, T, wavelength, rv, theta):
    T = np.ascontiguousarray(T, dtype=np.complex128)
    th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    q = [np.zeros(len(th)) for _ in range(3)]
    _olib().ctmo_extinction(nmax, _p(T), wavelength, rv, len(th), _p(th),
                            *[_p(a) for a in q])
    return q


def absorption(nmax, T, wavelength, rv, theta, ngauss=20,
               account_for_scattering=True):
    T = np.ascontiguousarray(T, dtype=np.complex128)
    th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    q = [np.zeros(len(th)) for _ in range(2)]
    _olib().ctmo_absorption(nmax, _p(T), wavelength, rv, len(th), _p(th),
                            ngauss, int(account_for_scattering),
                            *[_p(a) for a in q])
    return q


def scattering_grid(nmax, T, wavelength, ntheta, use_gl=True):
    T = np.ascontiguousarray(T, dtype=np.complex128)
    Z = np.zeros((ntheta, 2 * ntheta, 4, 4))
    _olib().ctmo_scattering_grid(nmax, _p(T), wavelength, ntheta, int(use_gl),
                                 _p(Z))
    return Z


ORIENT_ISOTROPIC, ORIENT_MISHCHENKO, ORIENT_DAVIS_GREENSTEIN = 0, 1, 2


def orientation_coefficients(kind, param, nmax):
    p = np.zeros(nmax + 1)
    _olib().ctmo_orientation_coefficients(kind, param, nmax, _p(p))
    return p


def orientation_average(nmax, T, pN):
    T = np.ascontiguousarray(T, dtype=np.complex128)
    pN = np.ascontiguousarray(pN, dtype=np.float64)
    assert len(pN) >= 2 * nmax + 1
    out = np.zeros_like(T)
    _olib().ctmo_orientation_average(nmax, _p(T), _p(pN), _p(out))
    return out


def batch_task_obs(rv, d, lam, mr_re, mr_im, theta, wtheta, ngauss=20,
                   min_order=4, max_order=100, glfac=2, tol=1e-4, nthreads=0):
    L = _olib()
    if not getattr(L, "_bto", False):
        L.ctmo_batch_task_obs.argtypes = [_I, _PD, _PD, _PD, _PD, _PD, _U, _U,
                                          _U, _D, _I, _U, _PD, _PD, _U, _PU,
                                          _PI, _PD, _PD]
        L._bto = True
    rv, d, lam, mr_re, mr_im, theta, wtheta = [
        np.ascontiguousarray(a, dtype=np.float64)
        for a in (rv, d, lam, mr_re, mr_im, theta, wtheta)]
    n = len(rv)
    nmax = np.zeros(n, dtype=np.uint32)
    status = np.zeros(n, dtype=np.int32)
    qe = np.zeros(n)
    qa = np.zeros(n)
    L.ctmo_batch_task_obs(n, _p(rv), _p(d), _p(lam), _p(mr_re), _p(mr_im),
                          min_order, max_order, glfac, tol, nthreads,
                          len(theta), _p(theta), _p(wtheta), ngauss,
                          nmax.ctypes.data_as(_PU), status.ctypes.data_as(_PI),
                          _p(qe), _p(qa))
    return nmax, status, qe, qa


def scattering_query(nmax, T, wavelength, alpha, beta, th_in, ph_in, th_out,
                     ph_out):
    """Returns (S 2x2 complex, Z 4x4, K 4x4)."""
    L = _olib()
    if not getattr(L, "_sq", False):
        L.ctmo_scattering_query.argtypes = [_U, _PD, _D, _D, _D, _D, _D, _D,
                                            _D, _PD]
        L._sq = True
    T = np.ascontiguousarray(T, dtype=np.complex128)
    out = np.zeros(40)
    L.ctmo_scattering_query(nmax, _p(T), wavelength, alpha, beta, th_in,
                            ph_in, th_out, ph_out, _p(out))
    S = (out[0:8:2] + 1j * out[1:8:2]).reshape(2, 2)
    return S, out[8:24].reshape(4, 4), out[24:40].reshape(4, 4)


def batch_task_obs_ext(rv, d, lam, mr_re, mr_im, theta, wtheta, ngauss=20,
                       min_order=4, max_order=100, glfac=2, tol=1e-4,
                       nthreads=0, nmax_forced=None, per_angle=False):
    """batch_task_obs plus the full-T sums and, with per_angle, Q_ext and
    Q_abs at every incidence angle.  With nmax_forced the convergence loop
    is skipped and every grain is computed at its given order.  Returns a
    dict of numpy arrays."""
    L = _olib()
    if not getattr(L, "_btox", False):
        L.ctmo_batch_task_obs_ext.argtypes = [
            _I, _PD, _PD, _PD, _PD, _PD, _U, _U, _U, _D, _I, _U, _PD, _PD, _U,
            _PU, _PU, _PI, _PD, _PD, _PD, _PD, _PD, _PD]
        L._btox = True
    rv, d, lam, mr_re, mr_im, theta, wtheta = [
        np.ascontiguousarray(a, dtype=np.float64)
        for a in (rv, d, lam, mr_re, mr_im, theta, wtheta)]
    n = len(rv)
    nt = len(theta)
    out = dict(nmax=np.zeros(n, dtype=np.uint32),
               status=np.zeros(n, dtype=np.int32),
               qext_avg=np.zeros(n), qabs_avg=np.zeros(n),
               qsca_sum=np.zeros(n), qext_sum=np.zeros(n))
    if per_angle:
        out["qext_theta"] = np.zeros((n, nt))
        out["qabs_theta"] = np.zeros((n, nt))
    forced = None
    if nmax_forced is not None:
        nf = np.ascontiguousarray(nmax_forced, dtype=np.uint32)
        forced = nf.ctypes.data_as(_PU)
    pa = (lambda k: _p(out[k]) if per_angle else None)
    L.ctmo_batch_task_obs_ext(
        n, _p(rv), _p(d), _p(lam), _p(mr_re), _p(mr_im), min_order,
        max_order, glfac, tol, nthreads, nt, _p(theta), _p(wtheta), ngauss,
        forced, out["nmax"].ctypes.data_as(_PU),
        out["status"].ctypes.data_as(_PI), _p(out["qext_avg"]),
        _p(out["qabs_avg"]), _p(out["qsca_sum"]), _p(out["qext_sum"]),
        pa("qext_theta"), pa("qabs_theta"))
    return out
