"""The direct-solve oracle's sketch (tests/direct_oracle.py), checked where
a full direct solution is at hand: on a 65x65-node member of the
degraded-elasticity family the sketch of SuperLU's solution reports zero
error for that solution, and for perturbed vectors it estimates the true
relative error within the chi-square spread of a 64-row Gaussian sketch;
the committed 2.1M-DOF fixture belongs to the headline system and its
samples hold the prescribed boundary displacements."""
import os

import numpy as np
import pytest
from scipy.sparse.linalg import splu

from fem_amd import problems as P
from tests import direct_oracle as D


def test_sketch_estimates_the_true_error():
    s = P.degraded_elasticity(65)
    A = s.scipy()
    x = splu(A.tocsc()).solve(s.b)
    g = D.make(s, x, float(np.linalg.norm(s.b - A @ x) / np.linalg.norm(s.b)), "splu")
    rng = np.random.default_rng(4)
    trials = [x, x + 1e-9 * np.linalg.norm(x) / np.sqrt(s.n) * rng.normal(size=s.n),
              x * (1 + 1e-6), x + 1e-4 * np.abs(x).max() * (np.arange(s.n) % 7 == 0)]
    errs, info = D.errors(s, trials, g=g)
    assert errs[0]["sketch"] < 1e-14 and errs[0]["sample"] == 0.0
    for t, e in zip(trials[1:], errs[1:]):
        true = np.linalg.norm(t - x) / np.linalg.norm(x)
        assert 0.5 * true <= e["sketch"] <= 1.6 * true
    # a sign slip in the sketch (|S x + S x_d|) would report ~2 for the exact solution
    assert errs[0]["sketch"] < 1e-14
    with pytest.raises(ValueError):
        D.errors(P.degraded_elasticity(65, E=200e3), [x], g=g)


def test_headline_fixture_matches_the_system():
    g = D.fixture(1025)
    assert g is not None
    s = P.degraded_elasticity(1025)
    assert int(g["n"]) == s.n == 2101250 and int(g["nnz"]) == s.nnz
    assert float(g["b_norm"]) == float(np.linalg.norm(s.b))
    assert float(g["relres"]) < 1e-12
    # the prescribed displacements are exact in the direct solution: u_y on
    # the top edge is 1e-3, both components on the bottom edge 0
    nn = 1025 * 1025
    idx = g["idx"]
    top_y = (idx >= nn + 1024 * 1025)
    bottom = (idx < 1025) | ((idx >= nn) & (idx < nn + 1025))
    assert top_y.any() or bottom.any()
    assert np.all(np.abs(g["x_idx"][top_y] - 1e-3) <= 1e-15) and np.all(g["x_idx"][bottom] == 0.0)
