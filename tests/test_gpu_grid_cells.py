"""Two cells of the averaged (wavelength, radius, shape) grid against the
oracle (tests/golden/oracle_grid_cells.npz, written by
tests/golden/make_grid_fixture.py): all 20 Draine-Hensley-weighted shapes,
m(lambda) from the Lorentz dielectric function, task-path convergence to
order 159 (one cell's shapes converge above order 100, where the solve runs
in global scratch), Mishchenko alignment, Q_ext and Q_abs (20 x 20
scattered light) at 8 angles and Z at 91 scattering angles, shape-averaged
on the GPU (ctm_weighted_reduce).

Bars: every shape's converged order equals the oracle's; the shape-averaged
Q within max(1e-8, 50 x the oracle's own change under one ulp of R_V); Z
within 1e-8 of its largest element (1e-6 on the cell above order 100)."""
import os

import numpy as np
import pytest

import make_inputs

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(os.path.join(
                  os.path.dirname(os.path.abspath(__file__)), "golden",
                  "oracle_grid_cells.npz")),
                  reason="oracle fixture not generated "
                         "(tests/golden/make_grid_fixture.py)")]

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "oracle_grid_cells.npz")


@pytest.mark.parametrize("cell", [0, 1])
def test_grid_cell_matches_oracle(cell):
    import torch
    from costuum_amd.tmatrix import (Context, observables, query,
                                     tmatrix_batch, weighted_sums)
    with np.load(FIXTURE) as f:
        fx = {k: f[k].copy() for k in f.files}
    il, ia = fx["cells"][cell]
    lam, rad, d, w = make_inputs.grid_axes()
    mr = complex(make_inputs.lorentz_index(lam[il]))
    ns = len(d)
    rv = np.full(ns, rad[ia])
    lm = np.full(ns, lam[il])
    ctx = Context(0)
    try:
        print("grid cell", (il, ia), "20 shapes", flush=True)
        r = tmatrix_batch(rv, d, lm, np.full(ns, mr.real),
                          np.full(ns, mr.imag), ctx=ctx, min_order=4,
                          max_order=160, glfac=2, tolerance=1e-4)
        nmax, _, status = r.orders()
        assert np.array_equal(nmax, fx["nmax_%d" % cell]), (
            nmax, fx["nmax_%d" % cell])
        assert np.array_equal(status, fx["status_%d" % cell])
        pN = np.zeros((2, 321))
        pN[:, 0] = 1.0
        pN[0, 2] = 1.5 * 0.6 - 0.5   # oblate (d > 1)
        pN[1, 2] = 1.5 * 0.2 - 0.5   # prolate
        r.average(pN, np.where(d > 1.0, 0, 1).astype(np.int32))
        theta = np.arccos(np.polynomial.legendre.leggauss(8)[0])
        obs = observables(theta, ctx=ctx, ngauss=20, n_grains=ns)
        q = torch.stack([obs["qext"], obs["qabs"]], dim=2)
        wt = torch.as_tensor(w, device=ctx.device)
        qa = weighted_sums(q, wt, ns, normalise=True, ctx=ctx).cpu().numpy()[0]
        Theta = np.radians(np.arange(0.0, 181.0, 2.0))
        _, Z, _ = query(ctx, np.arange(ns)[:, None], 0.0, 0.0, 0.0, 0.0,
                        Theta[None, :], 0.0)
        za = weighted_sums(torch.as_tensor(Z, device=ctx.device), wt, ns,
                           normalise=True, ctx=ctx).cpu().numpy()[0]
    finally:
        ctx.close()
    ws = w / w.sum()
    for k, key in enumerate(("qext", "qabs")):
        ref = ws @ fx["%s_%d" % (key, cell)]
        ulp = ws @ fx["%s_ulp_%d" % (key, cell)]
        bar = np.maximum(1e-8, 50. * np.abs(ulp - ref) / np.abs(ref))
        e = np.abs(qa[:, k] - ref) / np.abs(ref)
        print(key, "max rel err %.2e (bar %.2e)" % (e.max(), bar.min()),
              flush=True)
        assert np.all(e <= bar), (key, e, bar)
    zr = fx["z_avg_%d" % cell]
    ez = np.abs(za - zr).max() / np.abs(zr).max()
    print("Z max rel err %.2e" % ez, flush=True)
    assert ez < (1e-6 if nmax.max() > 100 else 1e-8)
