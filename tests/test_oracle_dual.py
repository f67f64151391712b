"""The dual-number restatement of the oracle's T-matrix chain
(oracle/dual.py) against the C oracle (oracle/ctm_oracle.c), on the CPU:

* its value part equals the C oracle's T (the sums over nodes and the
  inverse are evaluated in another order; measured at most 7e-15 of max |T|,
  bar 1e-13);
* its tangent along each of (R_V, d, Re m, Im m) equals the
  Richardson-extrapolated central differences (4 D(h) - D(2h)) / 3 of the C
  oracle's T at relative step 1e-4 (measured at most 1.2e-10 of max |dT|;
  the differences' own error is larger at steps 1e-3, truncation, and 1e-5,
  rounding; bar 1e-9);
* a planted mistake (the tangent of (1/r) dr/dtheta dropped) fails that
  comparison (measured 0.58), so the check can catch a lost term.
Grains at forced orders 6 and 8, prolate and oblate, weak and strong
absorption."""
import numpy as np
import pytest

import oracle as O
from oracle import dual

CASES = [
    (0.2, 0.5, 0.55, 1.5 + 0.02j, 6, 12),
    (0.3, 1.6, 0.7, 1.7 + 0.1j, 8, 16),
    (0.05, 0.8, 2.0, 1.9 + 0.3j, 6, 12),
]


def _params(c):
    return np.array([c[0], c[1], c[3].real, c[3].imag])


def _oracle_T(c, p):
    return O.tmatrix_all(p[0], p[1], c[2], complex(p[2], p[3]), c[4], c[5])


def _richardson(c, j, h=1e-4):
    p0 = _params(c)
    step = h * max(abs(p0[j]), 1e-3)

    def cd(s):
        pp, pm = p0.copy(), p0.copy()
        pp[j] += s
        pm[j] -= s
        return (_oracle_T(c, pp) - _oracle_T(c, pm)) / (2. * s)

    return (4. * cd(step) - cd(2. * step)) / 3.


@pytest.mark.parametrize("case", range(len(CASES)))
def test_dual_value_equals_oracle(case):
    c = CASES[case]
    T, _ = dual.tmatrix_all_dual(*c, direction=0)
    ref = O.tmatrix_all(*c)
    err = np.abs(T - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("direction", range(4))
def test_dual_tangent_equals_oracle_differences(case, direction):
    c = CASES[case]
    _, dT = dual.tmatrix_all_dual(*c, direction=direction)
    fd = _richardson(c, direction)
    err = np.abs(dT - fd).max() / np.abs(fd).max()
    assert err < 1e-9, err


def test_planted_mistake_is_caught(monkeypatch):
    c = CASES[0]
    fd = _richardson(c, 1)
    _, good = dual.tmatrix_all_dual(*c, direction=1)
    monkeypatch.setattr(dual, "_PLANT", "dr")
    _, bad = dual.tmatrix_all_dual(*c, direction=1)
    e_good = np.abs(good - fd).max() / np.abs(fd).max()
    e_bad = np.abs(bad - fd).max() / np.abs(fd).max()
    assert e_good < 1e-9 and e_bad > 1e-3, (e_good, e_bad)
