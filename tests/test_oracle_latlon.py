"""The lat/lon quadratures of the oracle (oracle/equiangular.py) against
textbook weights and polynomial exactness, and the product's host-side nodes
and weights (cuhpx_amd/hpx_regrid.py) against the oracle."""

import numpy as np
import pytest

from cuhpx_amd.hpx_regrid import latlon_nodes
from oracle.equiangular import clenshaw_curtis_by_moments, gauss_lobatto_by_moments, latlon_grid

# (nodes north to south, weights) from the standard tables
TEXTBOOK = {
    ("equiangular", 3): ([1, 0, -1], [1 / 3, 4 / 3, 1 / 3]),
    ("equiangular", 5): ([1, np.sqrt(0.5), 0, -np.sqrt(0.5), -1], [1 / 15, 8 / 15, 4 / 5, 8 / 15, 1 / 15]),
    ("lobatto", 3): ([1, 0, -1], [1 / 3, 4 / 3, 1 / 3]),
    ("lobatto", 4): ([1, 1 / np.sqrt(5), -1 / np.sqrt(5), -1], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
    ("lobatto", 5): ([1, np.sqrt(3 / 7), 0, -np.sqrt(3 / 7), -1], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
    ("legendre-gauss", 2): ([1 / np.sqrt(3), -1 / np.sqrt(3)], [1, 1]),
    ("legendre-gauss", 3): ([np.sqrt(0.6), 0, -np.sqrt(0.6)], [5 / 9, 8 / 9, 5 / 9]),
}


def oracle_rule(grid, n):
    g = latlon_grid(n, 8, grid, csphase=False)
    return np.cos(g.theta), g.weight * 8 / (2 * np.pi)


@pytest.mark.parametrize("grid,n", sorted(TEXTBOOK))
def test_textbook_weights(grid, n):
    x, w = oracle_rule(grid, n)
    xt, wt = TEXTBOOK[(grid, n)]
    np.testing.assert_allclose(x, xt, atol=1e-14)
    np.testing.assert_allclose(w, wt, rtol=1e-13)


# degree of polynomial exactness: n - 1 (Clenshaw-Curtis; n for odd n),
# 2n - 1 (Gauss), 2n - 3 (Lobatto)
@pytest.mark.parametrize("grid,degree", [("equiangular", lambda n: n - 1), ("legendre-gauss", lambda n: 2 * n - 1),
                                         ("lobatto", lambda n: 2 * n - 3)])
@pytest.mark.parametrize("n", [4, 7, 12, 17])
def test_polynomial_exactness(grid, degree, n):
    x, w = oracle_rule(grid, n)
    for k in range(degree(n) + 1):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-12, (grid, n, k)
    # and not beyond: the next even power is integrated with an error
    k = degree(n) + 1 + (degree(n) + 1) % 2
    assert abs(np.dot(w, x ** k) - 2.0 / (k + 1)) > 1e-12


def test_lobatto_nodes_are_endpoints_and_derivative_roots():
    n = 9
    x, _ = gauss_lobatto_by_moments(n)
    assert x[0] == 1.0 and x[-1] == -1.0
    dp = np.polynomial.legendre.Legendre.basis(n - 1).deriv()
    assert np.max(np.abs(dp(x[1:-1]))) < 1e-10
    assert np.all(np.diff(x) < 0)


@pytest.mark.parametrize("grid", ["equiangular", "legendre-gauss", "lobatto"])
@pytest.mark.parametrize("n", [2, 3, 8, 33, 64, 129])
def test_product_nodes_match_oracle(grid, n):
    if grid == "lobatto" and n == 2:
        x, w = latlon_nodes(grid, n)
        np.testing.assert_allclose(x, [1, -1])
        np.testing.assert_allclose(w, [1, 1])
        return
    x, w = latlon_nodes(grid, n)
    xo, wo = oracle_rule(grid, n)
    np.testing.assert_allclose(x, xo, atol=1e-13)
    np.testing.assert_allclose(w, wo, rtol=1e-11, atol=1e-15)


def test_clenshaw_curtis_symmetric_positive():
    x, w = clenshaw_curtis_by_moments(21)
    np.testing.assert_allclose(w, w[::-1], rtol=1e-12)
    assert np.all(w > 0) and abs(w.sum() - 2) < 1e-13
