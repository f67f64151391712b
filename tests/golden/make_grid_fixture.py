"""Oracle results for two cells of the (wavelength, radius, shape) grid
(make_inputs.grid), computed once on the CPU here.

Each cell holds all 20 Draine-Hensley-weighted axis ratios of one
(wavelength, radius) pair with m(lambda) from the seeded Lorentz dielectric
function.  Per shape the oracle runs the task-path convergence (orders
4..159, ngauss = 2 order, tolerance 1e-4), the orientation average with the
Mishchenko alignment (<cos^2 beta> = 0.6 for oblate, 0.2 for prolate
grains), Q_ext and Q_abs (20 x 20 scattered light) at the 8 Gauss-Legendre
incidence angles theta = arccos(x_i), and Z at 91 scattering angles 0, 2,
..., 180 deg for incidence along the symmetry axis (alpha = beta = 0,
theta_in = phi_in = phi_out = 0).  Stored per cell: the converged orders and
status of the shapes, their Q values and the same at R_V moved up by one
ulp (the oracle's own sensitivity), each shape's Z at 19 of the angles (0,
10, ..., 180 deg) and the same at R_V + 1 ulp, the shape weights, and the
shape-averaged Z.  Some shapes of the first cell reach order 159 without
converging; their results change by orders of magnitude under one ulp of
R_V (the reference's EBCM shares this), which the stored ulp values show.

    cell (lambda index 5, radius index 47): x = 82.5, orders above 100
    cell (lambda index 10, radius index 40): x = 22.2, orders about 35

Writes oracle_grid_cells.npz beside this file.  Each finished shape is
checkpointed in oracle_grid_cells.partial.npz, so an interrupted run
resumes; the partial file is removed at the end.

    python tests/golden/make_grid_fixture.py [--processes 8]
"""
import argparse
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import make_inputs  # noqa: E402

CELLS = ((5, 47), (10, 40))
OUT = os.path.join(HERE, "oracle_grid_cells.npz")
PARTIAL = os.path.join(HERE, "oracle_grid_cells.partial.npz")
FIELDS = ("n", "st", "qe", "qa", "Z", "qe2", "qa2", "Z2")


def one_shape(args):
    import oracle as O
    idx, rv, d, lam, mr = args
    theta = np.arccos(np.polynomial.legendre.leggauss(8)[0])
    Theta = np.radians(np.arange(0.0, 181.0, 2.0))
    st, n, ng, _, _ = O.converge_task(rv, d, lam, mr, 4, 160, 2, 1e-4)
    T = O.tmatrix_all(rv, d, lam, mr, n, ng)
    pN = np.zeros(2 * n + 1)
    pN[0] = 1.0
    pN[2] = 1.5 * (0.6 if d > 1.0 else 0.2) - 0.5
    Ta = O.orientation_average(n, T, pN)
    qe, _, _ = O.extinction(n, Ta, lam, rv, theta)
    qa, _ = O.absorption(n, Ta, lam, rv, theta, ngauss=20)
    Z = np.array([O.scattering_query(n, Ta, lam, 0., 0., 0., 0., th, 0.)[1]
                  for th in Theta])
    # the oracle's own change under one ulp of R_V, at the same order
    rv2 = np.nextafter(rv, np.inf)
    Ta2 = O.orientation_average(n, O.tmatrix_all(rv2, d, lam, mr, n, ng), pN)
    qe2, _, _ = O.extinction(n, Ta2, lam, rv2, theta)
    qa2, _ = O.absorption(n, Ta2, lam, rv2, theta, ngauss=20)
    Z2 = np.array([O.scattering_query(n, Ta2, lam, 0., 0., 0., 0., th, 0.)[1]
                   for th in Theta[::5]])
    return idx, (n, st, qe, qa, Z, qe2, qa2, Z2)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--processes", type=int, default=8)
    a = ap.parse_args()
    lam, rad, d, w = make_inputs.grid_axes()
    out = {"cells": np.array(CELLS), "shape_weights": w,
           "axis_ratios": d}
    jobs = []
    for il, ia in CELLS:
        mr = complex(make_inputs.lorentz_index(lam[il]))
        jobs += [(rad[ia], dk, lam[il], mr) for dk in d]
    done = {}
    if os.path.exists(PARTIAL):
        with np.load(PARTIAL) as f:
            for key in f.files:
                i, fld = key.split("_", 1)
                done.setdefault(int(i[1:]), {})[fld] = f[key].copy()
    todo = [(i,) + jobs[i] for i in range(len(jobs)) if i not in done]
    # largest size parameters first, so the long shapes start early
    todo.sort(key=lambda j: -j[1] / j[3])
    with Pool(a.processes) as pool:
        for i, r in pool.imap_unordered(one_shape, todo, chunksize=1):
            done[i] = dict(zip(FIELDS, [np.asarray(v) for v in r]))
            np.savez(PARTIAL, **{"i%d_%s" % (k, f): v for k, rr in
                                 done.items() for f, v in rr.items()})
            print("shape %d done (%d/%d), order %d" % (
                i, len(done), len(jobs), int(done[i]["n"])), flush=True)
    res = [tuple(done[i][f] for f in FIELDS) for i in range(len(jobs))]
    ns = len(d)
    for c in range(len(CELLS)):
        r = res[c * ns:(c + 1) * ns]
        out["nmax_%d" % c] = np.array([x[0] for x in r], dtype=np.int32)
        out["status_%d" % c] = np.array([x[1] for x in r], dtype=np.int32)
        out["qext_%d" % c] = np.array([x[2] for x in r])
        out["qabs_%d" % c] = np.array([x[3] for x in r])
        out["qext_ulp_%d" % c] = np.array([x[5] for x in r])
        out["qabs_ulp_%d" % c] = np.array([x[6] for x in r])
        Z = np.array([x[4] for x in r])
        out["z_avg_%d" % c] = np.tensordot(w, Z, axes=1) / w.sum()
        out["z_shape_%d" % c] = Z[:, ::5]
        out["z_shape_ulp_%d" % c] = np.array([x[7] for x in r])
        print("cell", CELLS[c], "orders", out["nmax_%d" % c].min(), "..",
              out["nmax_%d" % c].max(), "status", out["status_%d" % c].max(),
              flush=True)
    np.savez_compressed(OUT, **out)
    os.remove(PARTIAL)
    print("wrote", OUT, flush=True)


if __name__ == "__main__":
    main()
