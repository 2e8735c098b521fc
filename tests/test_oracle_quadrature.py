"""Pins the oracle's quadrature (gauleg.hpp:35-121) against numpy and exactness."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 12, 16, 20])
def test_gauss_matches_numpy(O, n):
    x, w = O.gauss_rule(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - 0.5 * (xr + 1))) < 2e-16
    assert np.max(np.abs(w - 0.5 * wr)) < 1e-15
    assert abs(w.sum() - 1.0) < 1e-15


@pytest.mark.parametrize("n", [4, 11, 16])
def test_gauss_exact_on_polynomials(O, n):
    x, w = O.gauss_rule(n, -0.3, 2.0)
    for deg in range(2 * n):
        exact = (2.0 ** (deg + 1) - (-0.3) ** (deg + 1)) / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-12 * max(1.0, abs(exact))


@pytest.mark.parametrize("n", [2, 6, 11, 12, 16])
def test_composite_rule_structure(O, n):
    x, w = O.composite_rule(n)
    sig = (np.sqrt(2) - 1) ** 2
    assert x.size == n * (n + 1) // 2 - 1
    # [0, sigma^(n-1)] is dropped: sum w = 1 - sigma^(n-1)
    assert abs(w.sum() - (1.0 - sig ** (n - 1))) < 1e-14
    assert np.all(np.diff(x) > 0) and x[0] > 0 and x[-1] < 1
    # per sub-interval [sigma^k, sigma^(k-1)] with n-k+1 points: exact to degree 2(n-k+1)-1
    pos = x.size
    hi = 1.0
    for k in range(1, n):
        npt = n - k + 1
        lo = sig ** k
        xs, ws = x[pos - npt:pos], w[pos - npt:pos]
        assert np.all((xs > lo) & (xs < hi))
        for deg in range(2 * npt):
            exact = (hi ** (deg + 1) - lo ** (deg + 1)) / (deg + 1)
            assert abs(np.dot(ws, xs ** deg) - exact) < 1e-13 * max(hi ** deg, 1e-300) * (hi - lo) + 1e-300
        pos -= npt
        hi = lo
    assert pos == 0
