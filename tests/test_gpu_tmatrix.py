"""GPU parity of the batched FP64 T-matrix path against the CPU oracle.

Tolerances: the reference's own relative difference |a-b|/|a+b|
(test/Assert.hpp:67-72); the FP64 path must agree with the oracle to 1e-8
on Q_ext/Q_sca, and T blocks to 1e-9 in Frobenius norm (both computed in
FP64; the inversion differs from the reference's explicit PLU inverse only
by rounding of order cond(Q) * 1e-16).
"""
import numpy as np
import pytest

import oracle as O
import make_inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / abs(a + b)


@pytest.fixture(scope="module")
def ctx():
    from costuum_amd.tmatrix import Context
    c = Context(0)
    yield c
    c.close()


FIXED_CASES = [
    # radius, d, wavelength, m, nmax, ngauss
    (0.2, 0.5, 0.55, 1.5 + 0.02j, 10, 20),     # prolate grain, x = 2.28
    (0.2, 2.0, 0.55, 1.5 + 0.02j, 10, 20),     # oblate twin
    (0.3, 1.0, 0.63, 1.7 + 0.1j, 12, 24),      # sphere (d = 1), kr off the zeros of j_0
    (1.0, 1.4, 1.1, 1.33 + 0.001j, 16, 33),    # odd ngauss
    (0.05, 0.7, 2.0, 1.9 + 0.3j, 5, 10),       # small grain
]


def test_fixed_orders_match_oracle(ctx):
    from costuum_amd.tmatrix import tmatrix_fixed
    cases = FIXED_CASES
    print("fixed-order cases:", cases, flush=True)
    r = tmatrix_fixed([c[0] for c in cases], [c[1] for c in cases],
                      [c[2] for c in cases], [c[3].real for c in cases],
                      [c[3].imag for c in cases], [c[4] for c in cases],
                      [c[5] for c in cases], ctx=ctx)
    qs = r.qsca.cpu().numpy()
    qe = r.qext.cpu().numpy()
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(cases):
        T_ref = O.tmatrix_all(rv, d, lam, mr, nmax, ng)
        T_gpu = r.tmatrix(g)
        for m, (a, b) in enumerate(zip(O.tmatrix_blocks(T_gpu, nmax),
                                       O.tmatrix_blocks(T_ref, nmax))):
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert err < 1e-9, (g, m, err)
        s_ref, e_ref = O.qcoeffs(nmax, T_ref)
        assert rel(qs[g], s_ref) < 1e-10, (g, qs[g], s_ref)
        assert rel(qe[g], e_ref) < 1e-10, (g, qe[g], e_ref)


def test_sphere_is_diagonal(ctx):
    """d = 1: T reduces to Mie coefficients -- diagonal in n, independent of
    m, zero between the two halves."""
    from costuum_amd.tmatrix import tmatrix_fixed
    nmax = 8
    r = tmatrix_fixed([0.3], [1.0], [0.63], [1.7], [0.1], [nmax], [16],
                      ctx=ctx)
    blocks = O.tmatrix_blocks(r.tmatrix(0), nmax)
    d11 = np.diag(blocks[0])[:nmax]
    for m in range(1, nmax + 1):
        B = blocks[m]
        nm = nmax + 1 - m
        off = B - np.diag(np.diag(B))
        assert np.abs(off).max() < 1e-12 * np.abs(B).max()
        assert np.allclose(np.diag(B)[:nm], d11[m - 1:], rtol=1e-10,
                           atol=1e-14)


def test_batch_convergence_matches_oracle(ctx):
    from costuum_amd.tmatrix import tmatrix_batch
    g = make_inputs.load(0, 4096)
    x = 2 * np.pi * g["radius"] / g["wavelength"]
    # a spread of sizes that the oracle finishes in seconds
    order = np.argsort(x)
    sel = order[np.linspace(0, 3600, 48).astype(int)]
    args = [g[k][sel] for k in ("radius", "axis_ratio", "wavelength",
                                 "mr_re", "mr_im")]
    print("batch of", len(sel), "grains, x in", x[sel].min(), x[sel].max(),
          flush=True)
    r = tmatrix_batch(*args, min_order=4, max_order=100, glfac=2,
                      tolerance=1e-4, ctx=ctx)
    nmax_gpu, _, status_gpu = r.orders()
    nmax_ref, st_ref, qs_ref, qe_ref = O.batch_task(*args, min_order=4,
                                                    max_order=100, glfac=2,
                                                    tol=1e-4, nthreads=4)
    assert np.array_equal(nmax_gpu, nmax_ref), (nmax_gpu, nmax_ref)
    assert np.array_equal(status_gpu, st_ref)
    qs = r.qsca.cpu().numpy()
    qe = r.qext.cpu().numpy()
    es = np.abs(qs - qs_ref) / np.abs(qs + qs_ref)
    ee = np.abs(qe - qe_ref) / np.abs(qe + qe_ref)
    print("max rel err Qsca %.3e Qext %.3e" % (es.max(), ee.max()))
    assert ee.max() < 1e-8 and es.max() < 1e-8


def test_orders_above_lds_capacity_match_oracle(ctx):
    """Blocks of order above 99 are inverted in global scratch; parity with
    the oracle at order 104 (x = 50), bar 1e-8 or 50 times the oracle's own
    sensitivity to one rounding of R_V."""
    from costuum_amd.tmatrix import tmatrix_fixed
    rv, d, lam, mr, N, G = 0.8, 1.1, 0.1, 1.5 + 0.01j, 104, 208
    print("order %d, ngauss %d" % (N, G), flush=True)
    r = tmatrix_fixed([rv], [d], [lam], [mr.real], [mr.imag], [N], [G],
                      ctx=ctx)
    T = r.tmatrix(0)
    Tref = O.tmatrix_all(rv, d, lam, mr, N, G)
    T2 = O.tmatrix_all(rv * (1 + 4e-16), d, lam, mr, N, G)
    scale = np.abs(Tref).max()
    bar = max(1e-8, 50 * np.abs(T2 - Tref).max() / scale)
    err = np.abs(T - Tref).max() / scale
    print("max rel err %.3e (bar %.3e)" % (err, bar))
    assert err < bar
    s_ref, e_ref = O.qcoeffs(N, Tref)
    assert rel(r.qsca.cpu().numpy()[0], s_ref) < 1e-8
    assert rel(r.qext.cpu().numpy()[0], e_ref) < 1e-8



def test_large_grains_converge_like_oracle(ctx):
    """Task-path convergence with maximum order 160 for x = 90 and 100
    (orders 108 and 131, blocks inverted in global scratch): same converged
    order as the oracle, Q_ext/Q_sca to 1e-8."""
    from costuum_amd.tmatrix import tmatrix_batch
    cases = [(1.0, 0.9, 0.07, 1.5 + 0.01j), (1.2, 1.3, 0.075, 1.35 + 0.05j)]
    print("large grains", cases, flush=True)
    r = tmatrix_batch([c[0] for c in cases], [c[1] for c in cases],
                      [c[2] for c in cases], [c[3].real for c in cases],
                      [c[3].imag for c in cases], min_order=4, max_order=160,
                      glfac=2, tolerance=1e-4, ctx=ctx)
    nmax, _, status = r.orders()
    for g, (rv, d, lam, mr) in enumerate(cases):
        st, n, ng, _, _ = O.converge_task(rv, d, lam, mr, 4, 160, 2, 1e-4)
        assert (int(nmax[g]), int(status[g])) == (n, st)
        Tref = O.tmatrix_all(rv, d, lam, mr, n, ng)
        s_ref, e_ref = O.qcoeffs(n, Tref)
        assert rel(r.qsca.cpu().numpy()[g], s_ref) < 1e-8
        assert rel(r.qext.cpu().numpy()[g], e_ref) < 1e-8
    # every entry of the order-131 T-matrix (m = 1 blocks are 131 wide:
    # the product covers columns beyond 128), bar as above
    rv, d, lam, mr = cases[1]
    n = int(nmax[1])
    ng = 2 * n
    Tref = O.tmatrix_all(rv, d, lam, mr, n, ng)
    T2 = O.tmatrix_all(rv * (1 + 4e-16), d, lam, mr, n, ng)
    scale = np.abs(Tref).max()
    bar = max(1e-8, 50 * np.abs(T2 - Tref).max() / scale)
    assert np.abs(r.tmatrix(1) - Tref).max() / scale < bar


def test_unknown_precision_is_rejected(ctx):
    import ctypes
    import torch
    from costuum_amd import _native
    dev = torch.device("cuda", 0)
    a = [torch.tensor([v], dtype=torch.float64, device=dev)
         for v in (0.1, 1.5, 0.5, 1.5, 0.01)]
    out = torch.zeros(4, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    conv = _native.CtmConvergence(4, 100, 2, 7, 1e-4)
    rc = ctx._lib.ctm_tmatrix_batch(
        ctx.handle, 1, *[ctypes.c_void_p(t.data_ptr()) for t in a],
        ctypes.byref(conv), ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(out.data_ptr() + 8),
        ctypes.c_void_p(out.data_ptr() + 16),
        ctypes.c_void_p(out.data_ptr() + 24))
    assert rc == _native.CTM_ERR_ARGUMENT
