"""The batch plan (work lists built on the GPU, no host round trip between
the convergence rounds, the all-m blocks and the observables) against the
host-driven path (ctm_tmatrix_batch + ctm_observables) on the same inputs.

Both paths run the same kernels on the same blocks, so the converged orders
and statuses must be identical and the values equal to rounding (the plan
changes only which workgroup takes a block and the window schedule; the
test bar is 1e-13 relative, and the maximum difference is printed).  A tile
ratio other than the host path's assembles some blocks on tiles of another
height; edge tiles sum their quadrature nodes in two interleaved groups, so
the node sums of those blocks differ in rounding, which the inversion
amplifies: the bar is then 50 times each grain's sensitivity to a one-ulp
change of its inputs (the oracle fixture's values at perturbed inputs),
at least 1e-11 (measured 2.8e-13 on well-conditioned grains, 4.8e-8 on one
whose sensitivity is 1.2e-6).  Cases:
a 48-grain sample spread over the seed-0 batch (orders 5 to 99) in FP64 and
in mixed precision, run three times in a row on the same plan (outputs
cleared in between, every run checked); and a plan created for small
grains and then fed larger ones, which outgrow its first memory pools (the
run is repeated with grown pools and must still equal the host path)."""
import os

import numpy as np
import pytest

import make_inputs

pytestmark = pytest.mark.gpu

KEYS = ("radius", "axis_ratio", "wavelength", "mr_re", "mr_im")


def _sample_idx(lo, hi, k):
    g = make_inputs.load(0, 4096)
    x = 2 * np.pi * g["radius"] / g["wavelength"]
    order = np.argsort(x)
    return order[np.linspace(lo, hi, k).astype(int)]


def _sample(lo, hi, k):
    g = make_inputs.load(0, 4096)
    sel = _sample_idx(lo, hi, k)
    return {key: g[key][sel] for key in KEYS}


def _spread_idx():
    # 32 grains over the size parameter, 16 of the 400 largest
    return np.concatenate([_sample_idx(0, 3687, 32),
                           _sample_idx(3688, 4095, 16)])


def _spread():
    g = make_inputs.load(0, 4096)
    sel = _spread_idx()
    return {key: g[key][sel] for key in KEYS}


FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "oracle_batch_seed0_n4096.npz")


def _spread_sensitivity():
    """Per grain of the spread: the largest relative change of the oracle's
    Q_sca, Q_ext and averaged Q_ext / Q_abs under a one-ulp change of the
    grain's inputs, and whether its converged order moves under it."""
    sel = _spread_idx()
    with np.load(FIXTURE) as f:
        s = np.maximum.reduce([
            np.abs(f[k + "_ulp"] - f[k]) / np.abs(f[k])
            for k in ("qsca_sum", "qext_sum", "qext_avg", "qabs_avg")])
        moved = f["nmax"] != f["nmax_ulp"]
    return s[sel], moved[sel]


THETA = np.arccos(np.polynomial.legendre.leggauss(8)[0])


def _host(g, precision, ctx):
    from costuum_amd.tmatrix import observables, tmatrix_batch
    n = len(g["radius"])
    r = tmatrix_batch(*[g[k] for k in KEYS], ctx=ctx, min_order=4,
                      max_order=100, glfac=2, tolerance=1e-4,
                      precision=precision)
    nmax, _, status = r.orders()
    obs = observables(THETA, ctx=ctx, ngauss=20, n_grains=n)
    return dict(nmax=nmax.astype(np.int64), status=status.astype(np.int64),
                qsca=r.qsca.cpu().numpy(), qext=r.qext.cpu().numpy(),
                oqext=obs["qext"].cpu().numpy(),
                oqabs=obs["qabs"].cpu().numpy(),
                oqabspol=obs["qabspol"].cpu().numpy())


def _plan_out(p):
    return dict(nmax=p.nmax.cpu().numpy().astype(np.int64),
                status=p.status.cpu().numpy().astype(np.int64),
                qsca=p.qsca.cpu().numpy(), qext=p.qext.cpu().numpy(),
                oqext=p.obs["qext"].cpu().numpy(),
                oqabs=p.obs["qabs"].cpu().numpy(),
                oqabspol=p.obs["qabspol"].cpu().numpy())


MICRO_RATIO_DEFAULT = 0.5
BAR_SAME_TILES, BAR_OTHER_TILES, SENS_FACTOR = 1e-13, 1e-11, 50.


def _bar(micro_ratio):
    """Per-grain bars and the grains whose order may move, for a plan with
    this tile ratio against the host path (default ratio) on the spread."""
    n = len(_spread_idx())
    if micro_ratio == MICRO_RATIO_DEFAULT:
        return np.full(n, BAR_SAME_TILES), np.zeros(n, bool)
    s, moved = _spread_sensitivity()
    return np.maximum(BAR_OTHER_TILES, SENS_FACTOR * s), moved


def _compare(a, b, what, bar=None):
    """bar: None (1e-13 for every grain) or (per-grain bars, grains whose
    order may move under rounding: left out)."""
    if bar is None:
        bar = (np.full(len(a["nmax"]), BAR_SAME_TILES),
               np.zeros(len(a["nmax"]), bool))
    bars, free = bar
    keep = ~free
    assert np.array_equal(a["nmax"][keep], b["nmax"][keep]), (
        what, a["nmax"], b["nmax"])
    assert np.array_equal(a["status"][keep], b["status"][keep]), what
    worst = 0.
    for k in ("qsca", "qext", "oqext", "oqabs", "oqabspol"):
        x, y = a[k][keep], b[k][keep]
        den = np.abs(x) + np.abs(y)
        if k == "oqabspol":
            # a difference of the two polarised parts, often far smaller
            # than either: measured against Q_abs
            den = np.maximum(den, np.abs(a["oqabs"][keep]) +
                             np.abs(b["oqabs"][keep]))
        den = np.maximum(den, 1e-300)
        rel = np.abs(x - y) / den
        if rel.ndim > 1:
            rel = rel.max(axis=1)
        worst = max(worst, float(rel.max()))
        assert np.all(rel <= bars[keep]), (
            what, k, rel.max(), int(np.argmax(rel / bars[keep])))
    print("%s: orders identical, max rel diff %.2e" % (what, worst),
          flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plan_equals_host_path(precision):
    from costuum_amd.tmatrix import BatchPlan, Context
    g = _spread()
    ctx = Context(0)
    try:
        print("48-grain spread, %s: host path" % precision, flush=True)
        ref = _host(g, precision, ctx)
        print("48-grain spread, %s: plan" % precision, flush=True)
        p = BatchPlan(*[g[k] for k in KEYS], min_order=4, max_order=100,
                      glfac=2, tolerance=1e-4, precision=precision,
                      theta=THETA, ngauss=20, ctx=ctx)
        try:
            p.run()
            _compare(ref, _plan_out(p), "%s run 0" % precision)
            info = p.info()
            print("plan info", info, flush=True)
            assert info["allm_items"] == int(np.sum(ref["nmax"] + 1))
            if precision == "mixed":
                assert info["mixed_blocks"] > 0
            for rep in (1, 2):
                p.nmax.zero_()
                p.qext.zero_()
                p.obs["qabs"].zero_()
                import torch
                torch.cuda.synchronize()
                p.run()
                _compare(ref, _plan_out(p), "%s run %d" % (precision, rep))
        finally:
            p.close()
    finally:
        ctx.close()


def test_plan_tuned_equals_host_path():
    """The autotuner changes only the round and tile split of the work:
    after tuning, the plan's results still equal the host path's."""
    from costuum_amd.tmatrix import BatchPlan, Context
    g = _spread()
    ctx = Context(0)
    try:
        print("48-grain spread: host path", flush=True)
        ref = _host(g, "fp64", ctx)
        p = BatchPlan(*[g[k] for k in KEYS], min_order=4, max_order=100,
                      glfac=2, tolerance=1e-4, theta=THETA, ngauss=20,
                      ctx=ctx)
        try:
            print("48-grain spread: plan, tuning", flush=True)
            t = p.tune(2)
            print("tuning", t, flush=True)
            assert t["trials"] >= 10 and t["ms_chosen"] <= t["ms_default"]
            p.nmax.zero_()
            p.qext.zero_()
            import torch
            torch.cuda.synchronize()
            p.run()
            _compare(ref, _plan_out(p), "tuned plan",
                     _bar(t["micro_ratio"]))
        finally:
            p.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("knobs", [
    dict(obs_lean=1, big_first=1),
    dict(obs_lean=0, big_first=0, window_add=1.0, micro_ratio=0.8),
    dict(big_first=2),
])
def test_plan_tuning_set_equals_host_path(knobs):
    """Tunables applied by ctm_plan_tuning_set (the low-LDS or the staged
    observable kernel for the largest blocks, the launch order of the
    largest solves, the first-window margin and tile ratio) change where
    the work runs, not its result: three runs in a row equal the host
    path."""
    from costuum_amd.tmatrix import BatchPlan, Context
    g = _spread()
    ctx = Context(0)
    try:
        print("48-grain spread: host path", flush=True)
        ref = _host(g, "fp64", ctx)
        p = BatchPlan(*[g[k] for k in KEYS], min_order=4, max_order=100,
                      glfac=2, tolerance=1e-4, theta=THETA, ngauss=20,
                      ctx=ctx)
        try:
            p.set_tuning(**knobs)
            for rep in range(3):
                print("plan with %s, run %d" % (knobs, rep), flush=True)
                p.nmax.zero_()
                p.qext.zero_()
                p.obs["qabs"].zero_()
                import torch
                torch.cuda.synchronize()
                p.run()
                _compare(ref, _plan_out(p), "tuning %s run %d" % (knobs, rep),
                         _bar(knobs.get("micro_ratio", MICRO_RATIO_DEFAULT)))
        finally:
            p.close()
    finally:
        ctx.close()


def test_plan_pools_grow():
    from costuum_amd.tmatrix import BatchPlan, Context
    small = _sample(0, 2000, 24)     # orders up to about 8
    large = _sample(3900, 4095, 24)  # orders 30 and above
    ctx = Context(0)
    try:
        print("plan created on 24 small grains, run on 24 large ones",
              flush=True)
        p = BatchPlan(*[small[k] for k in KEYS], min_order=4, max_order=100,
                      glfac=2, tolerance=1e-4, theta=THETA, ngauss=20,
                      ctx=ctx)
        try:
            p.run()
            first = p.info()
            _compare(_host(small, "fp64", ctx), _plan_out(p), "small grains")
            for i, k in enumerate(KEYS):
                p.inputs[i].copy_(p.inputs[i].new_tensor(large[k]))
            import torch
            torch.cuda.synchronize()
            p.run()
            info = p.info()
            print("pools", first["pool_bytes"], "->", info["pool_bytes"],
                  "reruns", info["reruns"], flush=True)
            assert info["reruns"] > first["reruns"]
            assert info["pool_bytes"] > first["pool_bytes"]
            _compare(_host(large, "fp64", ctx), _plan_out(p), "large grains")
        finally:
            p.close()
    finally:
        ctx.close()
