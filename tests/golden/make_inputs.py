"""Seeded synthetic grain batches for the T-matrix workloads.

Draws are made once with numpy's PCG64 (default_rng(seed)), field by field in
the order radius, axis ratio, wavelength, Re(m), Im(m), and committed as
grains_seed{seed}_n{n}.npz so the oracle, the GPU path and any later run read
identical bits.

  4096-grain batch, seed 0: radius log-uniform in [0.01, 1] um, axis ratio
  uniform in [0.5, 2], wavelength log-uniform in [0.1, 10] um, Re(m) uniform
  in [1.3, 2.0], Im(m) log-uniform in [1e-4, 0.5].
  1024-grain batch, seed 1: same distributions.

The (size, wavelength, shape) grid: 64 wavelengths log-spaced in
[0.1, 20] um, 48 equal-volume radii log-spaced in [0.005, 2] um, 20
Gauss-Legendre axis ratios on [0.5, 2] weighted by the Draine-Hensley CDE2
density, and a refractive index m(lambda) = sqrt(eps(lambda)) from a seeded
three-oscillator Lorentz dielectric function (61 440 T-matrices).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def draw(seed, n):
    rng = np.random.default_rng(seed)
    radius = np.exp(rng.uniform(np.log(0.01), np.log(1.0), n))
    axis_ratio = rng.uniform(0.5, 2.0, n)
    wavelength = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n))
    mr_re = rng.uniform(1.3, 2.0, n)
    mr_im = np.exp(rng.uniform(np.log(1e-4), np.log(0.5), n))
    return dict(radius=radius, axis_ratio=axis_ratio, wavelength=wavelength,
                mr_re=mr_re, mr_im=mr_im)


def lorentz_index(wavelength, seed=5):
    """m = sqrt(eps) (principal branch, Im m >= 0) of
    eps = 1 + sum_k f_k w_k^2 / (w_k^2 - w^2 - i g_k w), w = 1 / lambda,
    with three oscillators drawn from default_rng(seed): two ultraviolet
    resonances log-uniform in [0.03, 0.09] um and one infrared resonance
    log-uniform in [9, 20] um (a silicate-like band), strengths uniform in
    [0.2, 1.0], damping g_k / w_k uniform in [0.1, 0.3]."""
    rng = np.random.default_rng(seed)
    lam_k = np.concatenate([np.exp(rng.uniform(np.log(0.03), np.log(0.09),
                                               2)),
                            np.exp(rng.uniform(np.log(9.0), np.log(20.0),
                                               1))])
    f_k = rng.uniform(0.2, 1.0, 3)
    gr_k = rng.uniform(0.1, 0.3, 3)
    w = 1.0 / np.asarray(wavelength, dtype=np.float64)
    eps = np.ones_like(w, dtype=np.complex128)
    for lk, fk, gk in zip(lam_k, f_k, gr_k):
        wk = 1.0 / lk
        eps += fk * wk * wk / (wk * wk - w * w - 1j * gk * wk * w)
    return np.sqrt(eps)


def grid_axes():
    """(wavelengths, radii, axis ratios, shape weights) of the grid."""
    from numpy.polynomial.legendre import leggauss
    lam = np.logspace(np.log10(0.1), np.log10(20.0), 64)
    a = np.logspace(np.log10(0.005), np.log10(2.0), 48)
    x, w = leggauss(20)
    d = 1.25 + 0.75 * x
    w = 0.75 * w * np.array([_cde2(v) for v in d])
    return lam, a, d, w


def _cde2(d):
    """Draine-Hensley CDE2 density in the axis ratio
    (DraineHensleyShapeDistribution.hpp:67-147)."""
    if d < 1.0:
        e2 = 1.0 - d * d
        e = np.sqrt(e2)
        L = (1.0 - e2) * (0.5 * np.log((1.0 + e) / (1.0 - e)) / e - 1.0) / e2
        jac = 0.5 * np.sqrt(1.0 - e2) * ((3.0 - e2) * np.log((1.0 + e) /
                                                               (1.0 - e)) -
                                         6.0 * e) / (e2 * e2 * e)
    elif d > 1.0:
        e2 = 1.0 - 1.0 / (d * d)
        e = np.sqrt(e2)
        L = (1.0 - np.sqrt(1.0 - e2) * np.arcsin(e) / e) / e2
        jac = (1.0 - e2) * ((3.0 - 2.0 * e2) * np.arcsin(e) -
                            3.0 * e * np.sqrt(1.0 - e2)) / (e2 * e2 * e)
    else:
        L, jac = 1.0 / 3.0, 4.0 / 15.0
    return 12.0 * L * (L - 1.0) ** 2 * jac


def grid():
    """The grid as grain arrays, (lambda, radius, shape) order with the
    shape fastest, plus the shape weights."""
    lam, a, d, w = grid_axes()
    L, A, D = np.meshgrid(lam, a, d, indexing="ij")
    m = lorentz_index(L.ravel())
    return dict(radius=A.ravel(), axis_ratio=D.ravel(), wavelength=L.ravel(),
                mr_re=m.real.copy(), mr_im=m.imag.copy()), w


def path(seed, n):
    return os.path.join(HERE, "grains_seed%d_n%d.npz" % (seed, n))


def load(seed, n):
    p = path(seed, n)
    if os.path.exists(p):
        with np.load(p) as f:
            return {k: f[k].copy() for k in f.files}
    return draw(seed, n)


if __name__ == "__main__":
    for seed, n in ((0, 4096), (1, 1024)):
        np.savez(path(seed, n), **draw(seed, n))
        print("wrote", path(seed, n))
