#!/usr/bin/env python
"""Converged T-matrices/s on one MI355X for the 4096-spheroid batch.

Workload: 4096 spheroids drawn from seed 0 (radius log-uniform in
[0.01, 1] um, axis ratio uniform in [0.5, 2], wavelength log-uniform in
[0.1, 10] um, Re(m) uniform in [1.3, 2], Im(m) log-uniform in [1e-4, 0.5]);
task-path convergence with minimum order 4, maximum order 100, 2 Gauss
nodes per order and tolerance 1e-4.  One step = one pass of the hot path
over the whole batch with the inputs resident in HBM: every order of the
convergence loop, all m blocks at the converged order, the per-grain
Q_sca/Q_ext sums, and Q_ext and Q_abs (minus the scattered light, 20 x 20
quadrature) at 32 Gauss-Legendre incidence angles, averaged over them.

Both arithmetic paths are timed, each over its own steps: plain FP64 and
mixed (FP32 elimination with FP64 refinement).  A grain counts as converged
only when its convergence test passed, its order equals the oracle's and
its averaged Q_ext is within 1e-6 of the oracle's (the all-grain oracle
fixture tests/golden/oracle_batch_seed0_n4096.npz); ``value`` is the faster
path's count per second, and the line carries both paths, their error
distributions against the oracle, the mixed/plain ratio and the plain FP64
path run one grain per call.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: 20 timed steps after 3 warm-up steps for the batch workload
    # (about half a second timed), 1 + 1 for the grid and gradient workloads
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--precision", default="both",
                    choices=["both", "fp64", "mixed"],
                    help="arithmetic path(s) timed on the batch workload")
    ap.add_argument("--driver", default="plan", choices=["plan", "host"],
                    help="batch workload: plan = work lists built on the "
                         "GPU, no host round trip inside a step "
                         "(ctm_plan_run); host = lists built on the CPU "
                         "after each convergence round (ctm_tmatrix_batch "
                         "+ ctm_observables).  The other driver's FP64 step "
                         "time is reported beside it")
    ap.add_argument("--tune", type=int, default=1,
                    help="plan driver: autotune the plan on the batch "
                         "before the warm-up (ctm_plan_tune; results do "
                         "not depend on it)")
    ap.add_argument("--secondary", type=int, default=1,
                    help="batch workload: also run the single-grain, "
                         "gradient and grid workloads (one short "
                         "measurement each) and attach their results")
    ap.add_argument("--compare-driver", type=int, default=1,
                    help="also time the other driver's FP64 step")
    ap.add_argument("--per-grain", type=int, default=64,
                    help="grains of the per-grain (one call each) FP64 "
                         "comparator, every k-th in size order")
    ap.add_argument("--grains", type=int, default=4096)
    ap.add_argument("--cpu-baseline", type=int, default=1,
                    help="time the CPU oracle on a sample (rank 0, N=1)")
    ap.add_argument("--cpu-stride", type=int, default=64,
                    help="CPU sample = every k-th grain in size order")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "grid", "gradients", "single"],
                    help="batch: the 4096-grain seed-0 batch (default); "
                         "grid: the 61 440 T-matrix size x wavelength x "
                         "shape grid at orders up to 160; gradients: the "
                         "1024-grain seed-1 batch with d/d(R_V, d, Re m, "
                         "Im m) of Q_ext, Q_abs, Q_sca; single: one "
                         "grain through the calculator path (prolate d = "
                         "0.5 and oblate d = 2), isotropic Q at 30 deg, Z at "
                         "91 angles")
    ap.add_argument("--grid-cells", type=int, default=204,
                    help="grid workload: (size, wavelength) cells per batch "
                         "(each cell holds its 20 shapes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 512 MiB five times "
                         "(reference kernel for the HBM byte counters)")
    a = ap.parse_args()
    batch = a.workload == "batch"
    if a.steps is None:
        a.steps = 20 if batch else (10 if a.workload == "single" else 1)
    if a.warmup is None:
        a.warmup = 3 if batch else 1
    return a


def copy_calibration():
    """A plain device-to-device copy of known size: 512 MiB read and 512 MiB
    written per launch.  rocprofv3's FETCH_SIZE / WRITE_SIZE of this kernel
    give the counter-to-byte scale used for the traffic figure."""
    import torch
    n = 512 * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": 8 * n, "launches": 5}),
          flush=True)


# the assembly kernels that can dominate the batch step: (symbol, profiler
# name, traffic summary under profiles/)
ASSEMBLY_KERNELS = (("k_assemble<3>", "assemble_wide",
                     "k_assemble3_batch4096_traffic.md"),
                    ("k_assemble<2>", "assemble",
                     "k_assemble2_batch4096_traffic.md"))


def dominant_kernel(stats):
    """The assembly kernel with the most algorithmic work in the timed
    steps.  (By work, not by event time: the smaller assembly parts run
    beside each other and beside solves, so their event durations include
    time they shared the GPU; the 3 x 3 part of the all-m phase runs alone.
    With the plan driver its launches in the convergence windows run beside
    the other parts and are counted apart, as "assemble_window".)"""
    return max(ASSEMBLY_KERNELS, key=lambda k: stats[k[1]][2])


def measured_traffic(summary):
    """HBM bytes per launch of an assembly kernel on this workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/ (scaled
    by the same counters on a copy of known size); None when the summary is
    absent."""
    import re
    p = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+-]+)",
                  open(p).read())
    return float(m.group(1)) if m else None


def cpu_baseline(g, stride, threads, theta, wtheta):
    """The CPU oracle (FP64 restatement of the reference algorithm) on a
    size-stratified sample of the batch, doing the same work per grain as
    the GPU step (T-matrix, then Q_ext and Q_abs at the same angles)."""
    import numpy as np
    import oracle
    x = 2 * np.pi * g["radius"] / g["wavelength"]
    order = np.argsort(x)
    sel = order[stride // 2::stride]
    if threads <= 0:
        threads = len(os.sched_getaffinity(0)) or os.cpu_count() or 1
    args = [g[k][sel] for k in ("radius", "axis_ratio", "wavelength",
                                 "mr_re", "mr_im")]
    t0 = time.perf_counter()
    nmax, status, qe, qa = oracle.batch_task_obs(*args, theta, wtheta,
                                                 ngauss=20, min_order=4,
                                                 max_order=100, glfac=2,
                                                 tol=1e-4, nthreads=threads)
    dt = time.perf_counter() - t0
    return ({"value": float((status == 0).sum() / dt),
             "unit": "converged T-matrices/s", "cores": int(threads),
             "kind": "port", "host": host_info(),
             "sample": "%d of 4096 grains (every %dth in size-parameter "
                       "order, largest N %d), T-matrix plus Q_ext/Q_abs at "
                       "32 angles, %.2f s wall on %d threads"
                       % (len(sel), stride, int(nmax.max()), dt, threads)},
            sel, nmax, qe, qa)


def host_info():
    """nproc, the CPUs this process may use, and the CPU model (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("Socket(s):"):
                info["sockets"] = line.split(":", 1)[1].strip()
    except Exception:  # lscpu missing: the counts above still stand
        pass
    return info


def flops_per_tmatrix(N, min_order=4, glfac=2):
    """Algorithmic FLOPs of one converged T-matrix of order N (complex
    multiply-add = 8 flops): the all-m assembly and parity-split solves plus
    the m = 0 passes of the convergence orders min_order..N."""
    import numpy as np
    N = int(N)
    G = glfac * N
    nm = np.arange(N, 0, -1, dtype=np.float64)  # m = 1..N
    f = np.sum(96. * G * nm * nm + 64. / 3. * nm ** 3)
    n = np.arange(min_order, N + 1, dtype=np.float64)
    f += np.sum(32. * glfac * n * n * n + 16. / 3. * n ** 3)
    return float(f)


def load_fixture():
    """The oracle on all 4096 grains (tests/golden/make_batch_fixture.py),
    or None when it has not been generated."""
    import numpy as np
    p = os.path.join(ROOT, "tests", "golden", "oracle_batch_seed0_n4096.npz")
    if not os.path.exists(p):
        return None
    with np.load(p) as f:
        return {k: f[k].copy() for k in f.files}


def verify(fx, nmax, status, qext, qabs):
    """Per-grain agreement with the oracle fixture: same order, converged,
    and the relative errors of the averaged Q_ext and Q_abs; the parity bar
    of a grain is max(1e-8, 50 x the oracle's change under one ulp of R_V)."""
    import numpy as np
    same = (status == 0) & (fx["status"] == 0) & (nmax == fx["nmax"])
    ee = np.abs(qext - fx["qext_avg"]) / np.abs(fx["qext_avg"])
    ea = np.abs(qabs - fx["qabs_avg"]) / np.abs(fx["qabs_avg"])
    se = np.abs(fx["qext_avg_ulp"] - fx["qext_avg"]) / np.abs(fx["qext_avg"])
    sa = np.abs(fx["qabs_avg_ulp"] - fx["qabs_avg"]) / np.abs(fx["qabs_avg"])
    bar = np.maximum(1e-8, 50. * np.maximum(se, sa))
    ok6 = same & (ee <= 1e-6)
    ok8 = same & (ee <= bar) & (ea <= bar)
    sel = same
    return {"grains_same_order_converged": int(same.sum()),
            "within_1e-6_qext": int(ok6.sum()),
            "within_parity_bar": int(ok8.sum()),
            "max_rel_err_qext": float(ee[sel].max()),
            "p99_rel_err_qext": float(np.percentile(ee[sel], 99)),
            "max_rel_err_qabs": float(ea[sel].max()),
            "p99_rel_err_qabs": float(np.percentile(ea[sel], 99)),
            "other_order": [int(i) for i in np.flatnonzero(
                (status == 0) & (nmax != fx["nmax"]))],
            # grains whose order the oracle's FP64 arithmetic does not fix:
            # it moves under one ulp of R_V, or the relative change at it is
            # rounding noise (not counted in the value when they differ)
            "oracle_order_rounding_sensitive": int(
                ((fx["status"] == 0) & ((fx["nmax_ulp"] != fx["nmax"]) |
                                        (fx["dqs"] < 1e-12))).sum())
            if "nmax_ulp" in fx else None}, ok6


def run_grid(a, emit=True):
    """The (size, wavelength, shape) grid: 64 wavelengths x 48 radii x 20
    Draine-Hensley-weighted axis ratios in [0.5, 2], m(lambda) from a
    seeded Lorentz dielectric function, task-path convergence up to order
    159, Mishchenko alignment (<cos^2 beta> = 0.6 oblate, 0.2 prolate),
    Q_ext and Q_abs (with the 20 x 20 scattered-light quadrature) at 8
    Gauss-Legendre incidence angles, Z at 91 scattering angles 0..180 deg
    (incidence along the alignment axis), all shape-averaged.  One step is
    the whole grid in chunks of 204 (size, wavelength) cells."""
    import numpy as np
    import torch
    import make_inputs
    from costuum_amd.tmatrix import (Context, observables, query,
                                     tmatrix_batch, weighted_sums)
    g, wshape = make_inputs.grid()
    nshape = len(wshape)
    ncell = len(g["radius"]) // nshape
    ctx = Context(0)
    dev = ctx.device
    keys = ("radius", "axis_ratio", "wavelength", "mr_re", "mr_im")
    dg = [torch.as_tensor(g[k], dtype=torch.float64, device=dev) for k in keys]
    xg, _ = np.polynomial.legendre.leggauss(8)
    theta = np.arccos(xg)
    Theta = np.radians(np.arange(0.0, 181.0, 2.0))
    cells = a.grid_cells
    pN = np.zeros((2, 321))
    pN[:, 0] = 1.0
    pN[0, 2] = 1.5 * 0.6 - 0.5   # oblate (d > 1)
    pN[1, 2] = 1.5 * 0.2 - 0.5   # prolate
    wt = torch.as_tensor(np.tile(wshape, cells), device=dev)

    # parity sample: grains of moderate order spread over the grid
    x = 2 * np.pi * g["radius"] / g["wavelength"]
    cand = np.flatnonzero((x > 0.5) & (x < 12.0))
    sample = cand[np.linspace(0, len(cand) - 1, 6).astype(int)]
    per_grain = {}

    def step():
        nconv = 0
        q_all, z_all, st_all = [], [], []
        for c0 in range(0, ncell, cells):
            c1 = min(ncell, c0 + cells)
            s0, s1 = c0 * nshape, c1 * nshape
            n = s1 - s0
            r = tmatrix_batch(*[t[s0:s1] for t in dg], ctx=ctx, min_order=4,
                              max_order=160, glfac=2, tolerance=1e-4,
                              sync=False)
            st_all.append(r.status)
            dist = np.where(g["axis_ratio"][s0:s1] > 1.0, 0, 1).astype(
                np.int32)
            r.average(pN, dist)
            obs = observables(theta, ctx=ctx, ngauss=20, n_grains=n)
            q = torch.stack([obs["qext"], obs["qabs"]], dim=2)
            for i in sample[(sample >= s0) & (sample < s1)]:
                per_grain[int(i)] = (q[i - s0], r.nmax[i - s0])
            q_all.append(weighted_sums(q, wt[:n], nshape, normalise=True,
                                       ctx=ctx))
            _, Z, _ = query(ctx, np.arange(n)[:, None], 0.0, 0.0, 0.0, 0.0,
                            Theta[None, :], 0.0)
            z_all.append(weighted_sums(torch.as_tensor(Z, device=dev),
                                       wt[:n], nshape, normalise=True,
                                       ctx=ctx))
        st = torch.cat(st_all)
        return int((st == 0).sum().item()), torch.cat(q_all), \
            torch.cat(z_all), st

    for _ in range(a.warmup):
        step()
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        nconv, q, z, st = step()
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = ctx.profile_stats()
    ms = (t1 - t0) * 1e3 / max(a.steps, 1)
    symbol, pname, _ = dominant_kernel(stats)
    launches, kms, flops, _ = stats[pname]
    achieved = flops / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    qn = q.cpu().numpy()
    # cells whose shapes all converged (the others keep the last order, as
    # in the reference, and carry meaningless values)
    good = (st.cpu().numpy().reshape(-1, nshape) == 0).all(axis=1)
    parity = None
    cpu = None
    if a.cpu_baseline:
        import oracle
        errs = []
        t_cpu = time.perf_counter()
        for i, (qg, nm) in sorted(per_grain.items()):
            rv, d, lam = g["radius"][i], g["axis_ratio"][i], \
                g["wavelength"][i]
            mr = complex(g["mr_re"][i], g["mr_im"][i])
            stc, n, ng, _, _ = oracle.converge_task(rv, d, lam, mr, 4, 160, 2,
                                                    1e-4)
            T = oracle.tmatrix_all(rv, d, lam, mr, n, ng)
            Ta = oracle.orientation_average(n, T, pN[0 if d > 1.0 else 1])
            qe, _, _ = oracle.extinction(n, Ta, lam, rv, theta)
            qa, _ = oracle.absorption(n, Ta, lam, rv, theta, ngauss=20)
            qq = qg.cpu().numpy()
            errs.append(max(np.abs(qq[:, 0] - qe).max() / np.abs(qe).max(),
                            np.abs(qq[:, 1] - qa).max() / np.abs(qa).max()))
            assert int(nm) == n, (i, int(nm), n)
        dt = time.perf_counter() - t_cpu
        parity = {"grains": len(errs), "same_nmax": True,
                  "max_rel_err_q": float(max(errs))}
        cpu = {"value": round(len(errs) / dt, 3),
               "unit": "converged T-matrices/s", "cores": 1, "kind": "port",
               "host": host_info(),
               "sample": "the %d parity grains (x 0.5-12): convergence, "
                         "T-matrix, Mishchenko average, Q_ext and Q_abs at "
                         "8 angles (no Z), one thread, %.2f s"
                         % (len(errs), dt)}
    out = {
        "metric": "converged T-matrices/s on 1 GPU, size x wavelength x "
                  "shape grid",
        "value": round(nconv / (ms / 1e3), 2),
        "unit": "converged T-matrices/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "grid: 64 lambda x 48 radii x 20 DH shapes, "
                               "Lorentz m(lambda), orders 4..159, "
                               "Mishchenko alignment, Q at 8 angles, Z at "
                               "91 angles",
                   "t_matrices": int(len(g["radius"])), "converged": nconv,
                   "cells_all_converged": int(good.sum()),
                   "mean_qext_converged_cells": float(qn[good][..., 0].mean()),
                   "mean_qabs_converged_cells": float(
                       qn[good][..., 1].mean())},
        "roofline": {"kernel": symbol, "bound": "mfma",
                     "achieved": round(achieved, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "kernels": {k: {"launches": int(v[0]), "ms": round(v[1], 3)}
                    for k, v in stats.items()},
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def gradients_cpu_baseline(g, theta, k, threads):
    """The CPU oracle on k grains stratified over the size parameter (the
    smallest 95 %: one grain of order 99 alone would take minutes): the
    task-path convergence, then central differences (relative step 1e-6)
    of Q_ext and Q_abs (20 x 20 scattered light) at the angles along
    (R_V, d, Re m, Im m) at the converged order -- eight forced-order
    evaluations per grain, grains in parallel on the host cores."""
    import numpy as np
    import oracle
    x = 2 * np.pi * g["radius"] / g["wavelength"]
    sel = np.argsort(x)[np.linspace(0, int(0.95 * (len(x) - 1)),
                                    k).astype(int)]
    keys = ("radius", "axis_ratio", "wavelength", "mr_re", "mr_im")
    base = [g[key][sel].copy() for key in keys]
    w = np.full(len(theta), 1.0 / len(theta))
    t0 = time.perf_counter()
    conv = oracle.batch_task_obs_ext(*base, theta, w, ngauss=20,
                                     nthreads=threads)
    nmax = conv["nmax"]
    for j in (0, 1, 3, 4):  # R_V, d, Re m, Im m
        for sgn in (1.0, -1.0):
            args = [b.copy() for b in base]
            args[j] = args[j] * (1.0 + sgn * 1e-6)
            oracle.batch_task_obs_ext(*args, theta, w, ngauss=20,
                                      nthreads=threads, nmax_forced=nmax,
                                      per_angle=True)
    dt = time.perf_counter() - t0
    return {"value": round(k / dt, 3),
            "unit": "converged T-matrices with d/d(R_V, d, Re m, Im m) per s",
            "cores": int(threads), "kind": "port", "host": host_info(),
            "sample": "%d grains (stratified over the smallest 95 %% in x, "
                      "largest order %d): "
                      "convergence plus 8 central-difference evaluations of "
                      "Q_ext, Q_abs at %d angles, %.2f s on %d threads"
                      % (k, int(nmax.max()), len(theta), dt, threads)}


def run_gradients(a, emit=True):
    """The 1024-grain seed-1 batch: converged T-matrices, Q_ext and Q_abs
    (with the 20 x 20 scattered-light correction; Q_sca = Q_ext - Q_abs) at
    16 Gauss-Legendre incidence angles, the full-T Q_sca/Q_ext sums, and
    their derivatives with respect to (R_V, d, Re m, Im m) at the converged
    orders (forward-mode tangents; the quadratic scattered-light term by the
    exact symmetric difference).  One step does all of it."""
    import numpy as np
    import torch
    import make_inputs
    from costuum_amd.tmatrix import (Context, observables,
                                     observables_derivative, tangents,
                                     tmatrix_batch)
    g = make_inputs.load(1, 1024)
    ctx = Context(0)
    dev = ctx.device
    keys = ("radius", "axis_ratio", "wavelength", "mr_re", "mr_im")
    dg = [torch.as_tensor(g[k], dtype=torch.float64, device=dev) for k in keys]
    xg, _ = np.polynomial.legendre.leggauss(16)
    theta = np.arccos(xg)
    n = len(g["radius"])

    def step():
        r = tmatrix_batch(*dg, ctx=ctx, min_order=4, max_order=100, glfac=2,
                          tolerance=1e-4, sync=False)
        q = observables(theta, ctx=ctx, ngauss=20, n_grains=n)
        dqs, dqe = tangents(r)
        dq = [observables_derivative(theta, j, ctx, n, radius=g["radius"],
                                     ngauss=20) for j in range(4)]
        return r, q, dqs, dqe, dq

    def forward():
        r = tmatrix_batch(*dg, ctx=ctx, min_order=4, max_order=100, glfac=2,
                          tolerance=1e-4, sync=False)
        q = observables(theta, ctx=ctx, ngauss=20, n_grains=n)
        return r, q

    def timed(fn):
        for _ in range(a.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            res = fn()
        ctx.synchronize()
        torch.cuda.synchronize()
        return res, (time.perf_counter() - t0) * 1e3 / max(a.steps, 1)

    _, ms_fwd = timed(forward)
    (r, q, dqs, dqe, dq), ms = timed(step)
    st = r.status.cpu().numpy()
    cpu = None
    if a.cpu_baseline:
        threads = a.cpu_threads or len(os.sched_getaffinity(0)) or 1
        cpu = gradients_cpu_baseline(g, theta, 32, threads)
    out = {
        "metric": "converged T-matrices with gradients/s on 1 GPU",
        "value": round(int((st == 0).sum()) / (ms / 1e3), 2),
        "unit": "converged T-matrices with d/d(R_V, d, Re m, Im m) per s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "1024 spheroids, seed 1, orders 4..99; "
                               "Q_ext, Q_abs (20x20 scattered light), Q_sca "
                               "at 16 angles and full-T sums with their "
                               "derivatives along (R_V, d, Re m, Im m)",
                   "grains": n, "converged": int((st == 0).sum()),
                   "forward_only_ms_per_step": round(ms_fwd, 3),
                   "with_gradients_over_forward": round(ms / ms_fwd, 3),
                   "mean_dqext_dR_converged": float(
                       dq[0]["qext"][torch.as_tensor(st == 0, device=dev)]
                       .mean().item())},
        "roofline": None, "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def run_single(a, emit=True):
    """The single-grain case through the reference's module surface:
    R_V = 0.2 um, lambda = 0.55 um, m = 1.5 + 0.02i, for the prolate (d =
    0.5) and oblate (d = 2) reading of axis ratio 2; TMatrixCalculator
    convergence (tolerance 1e-4), isotropic ensemble, Q_ext and Q_abs at 30
    deg, Z at 91 scattering angles 0..180 deg.  A step is both grains."""
    import math
    import numpy as np
    from costuum_amd import CosTuuM as C
    rv, lam, mr = 0.2, 0.55, 1.5 + 0.02j
    th = math.radians(30.0)
    Theta = np.radians(np.arange(0, 181, 2.0))
    area = math.pi * rv * rv

    def step():
        out = []
        for d in (0.5, 2.0):
            tm = C.TMatrix(rv, d, lam, mr, cos2beta=1.0 / 3.0)
            qabs = tm.get_absorption_cross_section(th)[0] / area
            qext = tm.get_extinction_matrix(th, 0.0)[0, 0] / area
            Z = tm.get_scattering_matrix(0.0, 0.0, Theta, 0.0)
            out.append((tm.get_nmax(), tm.get_ngauss(), qext, qabs,
                        float(Z[45, 0, 0])))
            tm.close()
        return out

    for _ in range(a.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = step()
    ms = (time.perf_counter() - t0) * 1e3 / max(a.steps, 1)
    cpu = None
    if a.cpu_baseline:
        # the oracle: the calculator's two-phase convergence, the isotropic
        # ensemble, Q_ext and Q_abs at 30 deg and Z at 91 angles, per grain
        import oracle
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            for d in (0.5, 2.0):
                stc, n, ng, rve = oracle.converge_calculator(1.0, d, rv, lam,
                                                             mr)
                T = oracle.tmatrix_all(rve, d, lam, mr, n, ng)
                pN = np.zeros(2 * n + 1)
                pN[0] = 1.0
                Ta = oracle.orientation_average(n, T, pN)
                oracle.extinction(n, Ta, lam, rve, th)
                oracle.absorption(n, Ta, lam, rve, th, ngauss=100)
                for t_out in Theta:
                    oracle.scattering_query(n, Ta, lam, 0., 0., 0., 0.,
                                            t_out, 0.)
        dt = (time.perf_counter() - t0) / reps
        cpu = {"value": round(2 / dt, 2), "unit": "converged T-matrices/s "
               "(with Q and Z at 91 angles)", "cores": 1, "kind": "port",
               "host": host_info(),
               "sample": "both grains, %d repetitions, %.1f ms per step on "
                         "one thread" % (reps, dt * 1e3)}
    out = {
        "metric": "single grains through the calculator path per second",
        "value": round(2 / (ms / 1e3), 2),
        "unit": "converged T-matrices/s (with Q and Z at 91 angles)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single grain R_V 0.2 um, lambda 0.55 um, "
                               "m 1.5+0.02i, d 0.5 and 2 (one step = both)",
                   "grains": [{"d": d, "nmax": int(r[0]), "ngauss": int(r[1]),
                               "qext_30deg": r[2], "qabs_30deg": r[3],
                               "z11_90deg": r[4]}
                              for d, r in zip((0.5, 2.0), res)]},
        "roofline": None, "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.workload in ("grid", "gradients", "single"):
        if a.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        {"grid": run_grid, "gradients": run_gradients,
         "single": run_single}[a.workload](a)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: grains shard across GPUs "
              "with no communication, and no multi-GPU path is built",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import make_inputs
    from costuum_amd.tmatrix import (BatchPlan, Context, observables,
                                     tmatrix_batch)

    g = make_inputs.load(0, 4096)
    if a.grains != 4096:
        g = {k: v[:a.grains] for k, v in g.items()}
    n = len(g["radius"])
    ctx = Context(0)
    dev = ctx.device
    keys = ("radius", "axis_ratio", "wavelength", "mr_re", "mr_im")
    dg = [torch.as_tensor(g[k], dtype=torch.float64, device=dev) for k in keys]

    # 32 Gauss-Legendre incidence angles (cos theta nodes), alignment off
    xg, wg = np.polynomial.legendre.leggauss(32)
    theta = np.arccos(xg)
    wgt = torch.as_tensor(0.5 * wg, dtype=torch.float64, device=dev)

    def host_step(prec):
        r = tmatrix_batch(*dg, ctx=ctx, sync=False, min_order=4,
                          max_order=100, glfac=2, tolerance=1e-4,
                          precision=prec)
        obs = observables(theta, ctx=ctx, ngauss=20, n_grains=n)
        return r, obs["qext"] @ wgt, obs["qabs"] @ wgt

    def timed(prec, driver):
        if driver == "plan":
            plan = BatchPlan(*dg, min_order=4, max_order=100, glfac=2,
                             tolerance=1e-4, precision=prec, theta=theta,
                             ngauss=20, ctx=ctx)
            tuning = plan.tune(3) if a.tune else None

            def step(_):
                plan.run()
                return plan, plan.obs["qext"] @ wgt, plan.obs["qabs"] @ wgt
        else:
            step = host_step
        for _ in range(a.warmup):
            step(prec)
        ctx.profile(True)  # enables and resets the counters
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            res = step(prec)
        ctx.synchronize()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        stats = ctx.profile_stats()
        ctx.profile(False)
        r, qe, qa = res
        if driver == "plan":
            nmax = r.nmax.cpu().numpy().astype(np.uint32)
            status = r.status.cpu().numpy()
            info = r.info()
            blocks = (info["mixed_blocks"], info["mixed_fallback"])
            extra = {"plan_reruns": info["reruns"],
                     "plan_pool_bytes": info["pool_bytes"],
                     "window_round_orders": info["window_items"],
                     "tuning": tuning}
        else:
            nmax, _, status = r.orders()
            blocks = ctx.precision_stats()
            extra = {}
        out = {"ms": (t1 - t0) * 1e3 / max(a.steps, 1), "stats": stats,
               "nmax": nmax, "status": status,
               "qext": qe.cpu().numpy(), "qabs": qa.cpu().numpy(),
               "blocks": blocks, "extra": extra}
        if driver == "plan":
            plan.close()
        return out

    precs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    runs = {p: timed(p, a.driver) for p in precs}
    # the other driver's FP64 step, for comparison (same work, same results)
    other = "host" if a.driver == "plan" else "plan"
    other_run = timed("fp64", other) if a.compare_driver else None
    fx = load_fixture() if n == 4096 else None

    # the plain FP64 path one grain per call (the caller that does not
    # batch): every k-th grain in size order, each call synchronous
    per_grain = None
    if a.per_grain > 0:
        x = 2 * np.pi * g["radius"] / g["wavelength"]
        order = np.argsort(x)
        stride = max(1, n // a.per_grain)
        sel = order[stride // 2::stride]
        single = [[t[i:i + 1] for t in dg] for i in sel]
        step_one = (lambda args: (tmatrix_batch(
            *args, ctx=ctx, min_order=4, max_order=100, glfac=2,
            tolerance=1e-4), observables(theta, ctx=ctx, ngauss=20,
                                         n_grains=1)))
        step_one(single[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nconv1 = 0
        for args in single:
            r1, _ = step_one(args)
            nconv1 += int(r1.status.cpu().item() == 0)
        torch.cuda.synchronize()
        dt1 = time.perf_counter() - t0
        per_grain = {"value": round(nconv1 / dt1, 2),
                     "unit": "converged T-matrices/s",
                     "sample": "%d grains (every %dth in size-parameter "
                               "order), one ctm_tmatrix_batch + "
                               "ctm_observables call each, FP64, %.3f s"
                               % (len(sel), stride, dt1)}

    summary = {}
    for p, rr in runs.items():
        nconv = int((rr["status"] == 0).sum())
        entry = {"ms_per_step": round(rr["ms"], 3), "converged": nconv,
                 "blocks_fp32_elimination": int(rr["blocks"][0]),
                 "blocks_fp64_fallback": int(rr["blocks"][1])}
        entry.update(rr["extra"])
        if fx is not None:
            ver, ok6 = verify(fx, rr["nmax"], rr["status"], rr["qext"],
                              rr["qabs"])
            entry.update(ver)
            entry["value"] = round(int(ok6.sum()) / (rr["ms"] / 1e3), 2)
        else:
            entry["value"] = round(nconv / (rr["ms"] / 1e3), 2)
            entry["unverified"] = "no oracle fixture: convergence test only"
        summary[p] = entry
    best = max(summary, key=lambda p: summary[p]["value"])
    rr = runs[best]
    stats = rr["stats"]
    ms_per_step = rr["ms"]

    # rooflines: the assembly kernel with the most algorithmic work (FP64
    # compute bound), the solve and observable families, and the whole step
    symbol, pname, summ = dominant_kernel(stats)
    launches, ms, flops, nbytes = stats[pname]
    avg_ms = ms / max(launches, 1)
    achieved = flops / (ms / 1e3) / 1e12 if ms > 0 else 0.0
    roofline = {"kernel": symbol, "bound": "mfma",
                "launches_counted": "the all-m phase's launches (the kernel "
                                    "runs alone); with the plan driver its "
                                    "convergence-window launches run beside "
                                    "the other assembly parts and are listed "
                                    "under families.assemble_window"
                if pname == "assemble_wide" and a.driver == "plan" else
                "every launch with work in the timed steps",
                "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS,
                                                  4),
                "traffic": measured_traffic(summ) if n == 4096 else None,
                "avg_launch_ms": round(avg_ms, 4),
                "launches": int(launches),
                "algorithmic_flops_per_launch": flops / max(launches, 1),
                "algorithmic_bytes_per_launch": nbytes / max(launches, 1)}
    families = {}
    for fam in ("solve", "refine", "obs_x", "obs_sum", "assemble_small",
                "assemble_pack", "assemble", "assemble_window", "tables"):
        l, fm, ff, fb = stats.get(fam, (0, 0., 0., 0.))
        if l == 0:
            continue
        families[fam] = {"launches": int(l), "ms_per_step": round(
            fm / max(a.steps, 1), 3)}
        if ff > 0:
            tf = ff / (fm / 1e3) / 1e12
            families[fam].update({"tflops": round(tf, 3), "frac_fp64": round(
                tf / FP64_PEAK_TFLOPS, 4)})
        if fb > 0:
            families[fam]["gbs"] = round(fb / (fm / 1e3) / 1e9, 1)
    roofline["families"] = families
    nref = fx["nmax"] if fx is not None else rr["nmax"]
    cref = (fx["status"] == 0) if fx is not None else (rr["status"] == 0)
    step_flops = sum(flops_per_tmatrix(int(v)) for v in nref[cref])
    roofline["step"] = {
        "algorithmic_flops": step_flops,
        "achieved_tflops": round(step_flops / (ms_per_step / 1e3) / 1e12, 3),
        "frac": round(step_flops / (ms_per_step / 1e3) / 1e12 /
                      FP64_PEAK_TFLOPS, 4),
        "note": "sum over converged grains of the F_T(N) model at the "
                "oracle's orders, per step time"}
    kernels = {k: {"launches": int(v[0]), "ms": round(v[1], 3),
                   "tflops": round(v[2] / (v[1] / 1e3) / 1e12, 3)
                   if v[1] > 0 and v[2] > 0 else None}
               for k, v in stats.items()}
    cpu = None
    if a.cpu_baseline:
        cpu, sel, n_ref, qe_ref, qa_ref = cpu_baseline(
            g, a.cpu_stride, a.cpu_threads, theta, 0.5 * wg)
    plain = summary.get("fp64", {}).get("value")
    mixed = summary.get("mixed", {}).get("value")
    out = {
        "metric": "converged T-matrices/s on 1 GPU, batched mixed vs plain "
                  "FP64, at \u22641e-6 rel Q_ext err",
        "value": summary[best]["value"],
        "unit": "converged T-matrices/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "batched %d spheroids, seed 0, task-path "
                               "convergence (orders 4..99, ngauss = 2 order, "
                               "tol 1e-4), Q_ext/Q_abs (20x20 scattered "
                               "light) at 32 GL angles" % n,
                   "grains": n, "path_in_value": best,
                   "counted": "converged, same order as the oracle, "
                              "averaged Q_ext within 1e-6 of the oracle"
                              if fx is not None else
                              "converged (no oracle fixture)",
                   "orientations": 32, "absorption_quadrature": "20x20"},
        "paths": summary,
        "driver": a.driver,
        "other_driver_fp64": {
            "driver": other, "ms_per_step": round(other_run["ms"], 3),
            "same_orders": bool(np.array_equal(
                np.asarray(other_run["nmax"], dtype=np.int64),
                np.asarray(runs.get("fp64", other_run)["nmax"],
                           dtype=np.int64))),
            "max_rel_diff_qext": float(np.max(
                np.abs(other_run["qext"] - runs.get("fp64", other_run)["qext"])
                / np.maximum(np.abs(other_run["qext"]), 1e-300)))}
        if other_run else None,
        "mixed_over_plain": round(mixed / plain, 4)
        if plain and mixed else None,
        "per_grain_fp64_gpu": per_grain,
        "batched_over_per_grain": round(summary[best]["value"] /
                                        per_grain["value"], 1)
        if per_grain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    if a.secondary:
        # the other configurations of the workload, one short measurement
        # each (their own modes: --workload single | gradients | grid)
        import copy
        sec = {}
        for name, fn, steps, warm in (("single", run_single, 10, 1),
                                      ("gradients", run_gradients, 2, 1),
                                      ("grid", run_grid, 1, 1)):
            aa = copy.copy(a)
            aa.steps, aa.warmup, aa.workload = steps, warm, name
            r = fn(aa, emit=False)
            sec[name] = {k: r[k] for k in ("metric", "value", "unit",
                                           "ms_per_step", "steps", "warmup",
                                           "config", "roofline",
                                           "cpu_baseline")}
            if "parity_vs_oracle" in r:
                sec[name]["parity_vs_oracle"] = r["parity_vs_oracle"]
        out["secondary"] = sec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
