"""Quadrature rules of the C ABI (hf_quadrature) without a GPU: the
Chebyshev and Gauss-Lobatto rules equal the oracle's (which is pinned to
closed forms and numpy's Legendre roots), and the Gauss-Legendre rule of the
configured DFT-grid variant integrates polynomials of degree 2n - 1 exactly
and matches numpy's leggauss."""
import numpy as np
import pytest

import oracle as O
from helfem_amd.dftgrid import quadrature


@pytest.mark.parametrize("n", [1, 2, 3, 7, 15, 36, 75])
def test_gauss_legendre_exactness(n):
    x, w = quadrature("legendre", n)
    assert np.all(np.diff(x) > 0)
    for d in range(2 * n):
        exact = (1.0 - (-1.0) ** (d + 1)) / (d + 1)
        assert abs(np.dot(w, x ** d) - exact) < 1e-13 * max(1.0, n / 10)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 1e-14 and np.max(np.abs(w - wr)) < 1e-14


@pytest.mark.parametrize("n", [2, 5, 15, 75])
def test_lobatto_and_chebyshev_equal_oracle(n):
    x, w = quadrature("lobatto", n)
    xo, wo = O.lobatto(n)
    assert np.max(np.abs(x - xo)) < 1e-15 and np.max(np.abs(w - wo)) < 1e-15
    x, w = quadrature("chebyshev", n)
    xo, wo = O.chebyshev(n)
    assert np.max(np.abs(x - xo)) < 1e-15 and np.max(np.abs(w - wo)) < 1e-15
