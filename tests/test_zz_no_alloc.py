"""ctm_observables_noalloc (fixed device scratch, no allocation) against
ctm_observables on the same T-matrices, through the C ABI: two calls in a
row, the second on a smaller batch with fewer directions, then a request
whose scratch exceeds the fixed array (rejected on the host)."""
import ctypes

import numpy as np
import pytest
import torch

from costuum_amd import _native
from costuum_amd.tmatrix import Context, tmatrix_batch

pytestmark = pytest.mark.gpu

KEYS = ("qext", "qextpol", "qextcpol", "qabs", "qabspol", "Z")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _outputs(n, nt, nst, dev):
    out = {k: torch.full((n, nt), -7.0, dtype=torch.float64, device=dev)
           for k in KEYS[:5]}
    out["Z"] = (torch.full((n, nst, 2 * nst, 4, 4), -7.0, dtype=torch.float64,
                           device=dev) if nst > 0 else None)
    torch.cuda.synchronize()
    return out


def _call(ctx, fn, theta, ngauss, nst, n):
    th = np.ascontiguousarray(theta, dtype=np.float64)
    spec = _native.CtmObservableSpec(len(th), th.ctypes.data, 1, 1, 1, ngauss,
                                     nst, 1)
    out = _outputs(n, len(th), nst, ctx.device)
    rc = getattr(ctx._lib, fn)(ctx.handle, ctypes.byref(spec),
                               *[_ptr(out[k]) for k in KEYS])
    return rc, out


def _batch(ctx, rv, d, lam, mre, mim):
    r = tmatrix_batch(np.array(rv), np.array(d), np.array(lam),
                      np.array(mre), np.array(mim), ctx=ctx)
    return len(rv), r


def _compare(ctx, theta, ngauss, nst, n):
    print("noalloc: %d grains, %d angles, ngauss %d, Z grid %d"
          % (n, len(theta), ngauss, nst), flush=True)
    rc, ref = _call(ctx, "ctm_observables", theta, ngauss, nst, n)
    assert rc == 0, ctx._lib.ctm_last_error(ctx.handle)
    rc, got = _call(ctx, "ctm_observables_noalloc", theta, ngauss, nst, n)
    assert rc == 0, ctx._lib.ctm_last_error(ctx.handle)
    torch.cuda.synchronize()
    for k in KEYS:
        if ref[k] is None:
            continue
        assert torch.isfinite(ref[k]).all(), k
        # same kernels, same launch order: identical bits
        assert torch.equal(ref[k], got[k]), (k, (ref[k] - got[k]).abs().max())


def test_noalloc_matches_observables_twice():
    ctx = Context(0)
    try:
        # first call: six grains up to order ~20, eight incidence angles,
        # 12 x 12 scattered-light quadrature, Z on a 6 x 12 grid
        n, _ = _batch(ctx, [0.2, 0.2, 0.05, 0.5, 0.8, 0.3],
                      [0.5, 2.0, 0.7, 1.3, 1.6, 0.9],
                      [0.55, 0.55, 2.0, 1.2, 0.9, 0.4],
                      [1.5, 1.5, 1.9, 1.33, 1.7, 1.4],
                      [0.02, 0.02, 0.3, 0.001, 0.05, 0.1])
        theta = np.linspace(0.1, 3.0, 8)
        _compare(ctx, theta, 12, 6, n)
        # second call, smaller: three grains, four angles, 8 x 8, Z 4 x 8
        n, _ = _batch(ctx, [0.1, 0.3, 0.02], [1.2, 0.6, 1.9],
                      [0.5, 1.0, 0.3], [1.4, 1.8, 1.6], [0.01, 0.2, 0.001])
        _compare(ctx, np.array([0.3, 1.0, 1.7, 2.9]), 8, 4, n)
    finally:
        ctx.close()


def test_noalloc_rejects_oversized_scratch():
    ctx = Context(0)
    try:
        n, r = _batch(ctx, [0.2, 0.8], [0.5, 1.6], [0.55, 0.9], [1.5, 1.7],
                      [0.02, 0.05])
        nmax, _, _ = r.orders()
        # 64 angles with a quadrature large enough that the bilinear forms
        # alone (n_items x n_pairs x 64 bytes, n_pairs = 64 (1 + ngauss))
        # need twice the 64 MiB array
        theta = np.linspace(0.05, 3.1, 64)
        n_items = int((nmax.astype(np.int64) + 1).sum())
        ngauss = 2 * (64 << 20) // (n_items * 64 * 64) + 1
        need = n_items * 64 * (1 + ngauss) * 64
        assert need > 64 << 20, need
        print("noalloc: oversized request, X needs %d bytes" % need,
              flush=True)
        rc, _ = _call(ctx, "ctm_observables_noalloc", theta, ngauss, 0, n)
        assert rc == _native.CTM_ERR_ARGUMENT
        msg = ctx._lib.ctm_last_error(ctx.handle).decode()
        assert "bytes of scratch" in msg, msg
        # the context still works afterwards
        rc, _ = _call(ctx, "ctm_observables_noalloc", theta[:4], 8, 0, n)
        assert rc == 0
    finally:
        ctx.close()
