"""Seeded synthetic grain batches for the T-matrix workloads.

Draws are made once with numpy's PCG64 (default_rng(seed)), field by field in
the order radius, axis ratio, wavelength, Re(m), Im(m), and committed as
grains_seed{seed}_n{n}.npz so the oracle, the GPU path and any later run read
identical bits.

  4096-grain batch, seed 0: radius log-uniform in [0.01, 1] um, axis ratio
  uniform in [0.5, 2], wavelength log-uniform in [0.1, 10] um, Re(m) uniform
  in [1.3, 2.0], Im(m) log-uniform in [1e-4, 0.5].
  1024-grain batch, seed 1: same distributions.
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
