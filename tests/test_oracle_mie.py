"""Spheres: the T-matrix reduces to Mie coefficients (T11_n = -b_n,
T22_n = -a_n, diagonal and m-independent), computed here independently
from scipy's spherical Bessel functions."""
import numpy as np
from scipy.special import spherical_jn, spherical_yn

import oracle as O


def mie_ab(x, m, nmax):
    n = np.arange(1, nmax + 1)
    mx = m * x
    jx, yx = spherical_jn(n, x), spherical_yn(n, x)
    djx, dyx = spherical_jn(n, x, True), spherical_yn(n, x, True)
    jmx, djmx = spherical_jn(n, mx), spherical_jn(n, mx, True)
    psi, dpsi = x * jx, jx + x * djx
    xi, dxi = x * (jx + 1j * yx), (jx + 1j * yx) + x * (djx + 1j * dyx)
    psim, dpsim = mx * jmx, jmx + mx * djmx
    a = (m * psim * dpsi - psi * dpsim) / (m * psim * dxi - xi * dpsim)
    b = (psim * dpsi - m * psi * dpsim) / (psim * dxi - m * xi * dpsim)
    return a, b


def test_sphere_equals_mie():
    for rv, lam, mr in ((0.3, 0.63, 1.7 + 0.1j), (0.1, 0.5, 1.33 + 1e-3j),
                        (1.0, 0.9, 1.5 + 0.02j)):
        x = 2 * np.pi * rv / lam
        nmax = int(x + 4.05 * x ** (1 / 3)) + 6
        a, b = mie_ab(x, mr, nmax)
        T = O.tmatrix_all(rv, 1.0, lam, mr, nmax, 2 * nmax)
        blocks = O.tmatrix_blocks(T, nmax)
        B0 = blocks[0]
        assert np.allclose(np.diag(B0)[:nmax], -b, rtol=1e-9, atol=1e-13)
        assert np.allclose(np.diag(B0)[nmax:], -a, rtol=1e-9, atol=1e-13)
        for m in range(1, nmax + 1):
            nm = nmax + 1 - m
            assert np.allclose(np.diag(blocks[m])[:nm], -b[m - 1:],
                               rtol=1e-9, atol=1e-13)
        # Q_ext, Q_sca of the full T against the Mie series
        qs, qe = O.qcoeffs(nmax, T)
        n = np.arange(1, nmax + 1)
        assert abs(qe + np.sum((2 * n + 1) * (a + b).real)) < 1e-9 * abs(qe)
        assert abs(qs - np.sum((2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2))) \
            < 1e-9 * qs


def test_nonspherical_not_diagonal():
    T = O.tmatrix_m(0.3, 0.6, 0.63, 1.5 + 0.01j, 10, 20, 1)
    assert np.abs(T - np.diag(np.diag(T))).max() > 1e-6
