"""Oracle results on the whole 4096-grain seed-0 batch (CPU, run once here).

For every grain of grains_seed0_n4096.npz the CPU oracle runs the task-path
convergence (orders 4..99, ngauss = 2 order, tolerance 1e-4) and records:

  nmax, status            converged order and convergence status
  qsca_sum, qext_sum      unnormalised full-T sums (TMatrix.hpp:2055-2110)
  qext_avg, qabs_avg      Q_ext and Q_abs (scattered light subtracted with
                          the 20 x 20 quadrature) at the 32 Gauss-Legendre
                          incidence angles theta = arccos(x_i), averaged
                          with weights w_i / 2
  *_ulp                   the same four values recomputed at the same order
                          with the radius moved up by one ulp: the oracle's
                          own sensitivity to one rounding of its inputs,
                          the parity bar for ill-conditioned grains
  nmax_ulp                the converged order with the radius moved up by
                          one ulp
  dqs, dqe                the relative changes of the m = 0 sums at the
                          converged order (TMatrixResource.hpp:743-775)

A grain's converged order is fixed by FP64 arithmetic only when nmax_ulp
equals nmax and dqs is above rounding noise; otherwise a change of rounding
(another summation order) can move it.

Writes oracle_batch_seed0_n4096.npz beside this file.  Chunks are
checkpointed in oracle_batch_seed0_n4096.partial.npz so an interrupted run
resumes; the partial file is removed at the end.

    python tests/golden/make_batch_fixture.py [--threads 8] [--orders-only]

--orders-only adds nmax_ulp, dqs and dqe to an existing fixture.
"""
import argparse
import os
import sys
import time
from multiprocessing.pool import ThreadPool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import make_inputs  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(HERE, "oracle_batch_seed0_n4096.npz")
PARTIAL = os.path.join(HERE, "oracle_batch_seed0_n4096.partial.npz")
KEYS = ("radius", "axis_ratio", "wavelength", "mr_re", "mr_im")
FIELDS = ("nmax", "status", "qsca_sum", "qext_sum", "qext_avg", "qabs_avg",
          "qsca_sum_ulp", "qext_sum_ulp", "qext_avg_ulp", "qabs_avg_ulp")


def add_orders(g, out, threads):
    """nmax_ulp, dqs and dqe of every grain (convergence loop only)."""
    keys = list(zip(*[g[k] for k in KEYS]))

    def one(args):
        rv, d, lam, mre, mim = args
        mr = complex(mre, mim)
        st, n, _, dqs, dqe = oracle.converge_task(rv, d, lam, mr, 4, 100, 2,
                                                  1e-4)
        _, nu, _, _, _ = oracle.converge_task(np.nextafter(rv, np.inf), d,
                                              lam, mr, 4, 100, 2, 1e-4)
        return n, nu, dqs, dqe

    with ThreadPool(threads) as pool:
        r = pool.map(one, keys, chunksize=1)
    n = np.array([x[0] for x in r])
    assert np.array_equal(n, out["nmax"]), "convergence differs from batch"
    out["nmax_ulp"] = np.array([x[1] for x in r], dtype=np.uint16)
    out["dqs"] = np.array([x[2] for x in r])
    out["dqe"] = np.array([x[3] for x in r])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--chunk", type=int, default=128)
    ap.add_argument("--orders-only", action="store_true")
    a = ap.parse_args()
    g = make_inputs.load(0, 4096)
    if a.orders_only:
        with np.load(OUT) as f:
            out = {k: f[k].copy() for k in f.files}
        add_orders(g, out, a.threads)
        np.savez_compressed(OUT, **out)
        print("added orders to", OUT, flush=True)
        return
    n = len(g["radius"])
    xg, wg = np.polynomial.legendre.leggauss(32)
    theta = np.arccos(xg)
    wt = 0.5 * wg
    res = {k: np.zeros(n, dtype=np.float64) for k in FIELDS}
    done = np.zeros(n, dtype=bool)
    if os.path.exists(PARTIAL):
        with np.load(PARTIAL) as f:
            for k in FIELDS:
                res[k] = f[k].copy()
            done = f["done"].copy()
    # chunks of grains in size-parameter order mixed by a stride, so every
    # chunk holds a few large grains and the threads stay busy
    x = 2 * np.pi * g["radius"] / g["wavelength"]
    order = np.argsort(x)
    nchunk = (n + a.chunk - 1) // a.chunk
    chunks = [order[c::nchunk] for c in range(nchunk)]
    t0 = time.time()
    for ci, sel in enumerate(chunks):
        if done[sel].all():
            continue
        args = [g[k][sel] for k in KEYS]
        o = oracle.batch_task_obs_ext(*args, theta, wt, ngauss=20,
                                      min_order=4, max_order=100, glfac=2,
                                      tol=1e-4, nthreads=a.threads)
        rv_up = np.nextafter(args[0], np.inf)
        p = oracle.batch_task_obs_ext(rv_up, *args[1:], theta, wt, ngauss=20,
                                      min_order=4, max_order=100, glfac=2,
                                      tol=1e-4, nthreads=a.threads,
                                      nmax_forced=o["nmax"])
        for k in ("nmax", "status", "qsca_sum", "qext_sum", "qext_avg",
                  "qabs_avg"):
            res[k][sel] = o[k]
        for k in ("qsca_sum", "qext_sum", "qext_avg", "qabs_avg"):
            res[k + "_ulp"][sel] = p[k]
        done[sel] = True
        np.savez(PARTIAL, done=done, **res)
        print("chunk %d/%d: %d grains, largest order %d, %.0f s" %
              (ci + 1, nchunk, len(sel), int(o["nmax"].max()),
               time.time() - t0), flush=True)
    out = {k: res[k] for k in FIELDS}
    out["nmax"] = res["nmax"].astype(np.uint16)
    out["status"] = res["status"].astype(np.int8)
    add_orders(g, out, a.threads)
    np.savez_compressed(OUT, **out)
    os.remove(PARTIAL)
    print("wrote", OUT, flush=True)


if __name__ == "__main__":
    main()
