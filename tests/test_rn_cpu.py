"""Reflective Newton, host parts, without a GPU: the scaled Hessian
M = D H D + diag(|g| J) (RN_get_M_D, reflective_newton.c:250-262) bit for
bit against the compiled reference, including sparse_product's deletion of
zero entries (iterates sitting exactly on a bound give D_i = 0, so whole rows
and columns of M vanish) and explicit zeros in H."""
import numpy as np
import pytest

from fem_amd import lina
from fem_amd import problems as P


def _case(kind, rng):
    s, dold, ones = P.damage_system(17, h_scale=100.0)
    rp, ci, v = s.rowptr, s.col, s.val.copy()
    n = s.n
    lower, upper = dold.copy(), ones.copy()
    x = rng.uniform(dold, 1.0)
    g = rng.normal(size=n)
    if kind == "on_bounds":
        at = rng.random(n) < 0.3
        x[at] = dold[at]
        g[at] = np.abs(g[at])           # pushes against the lower bound: D = 0
        up = (~at) & (rng.random(n) < 0.2)
        x[up] = 1.0
        g[up] = -np.abs(g[up])
    elif kind == "zeros_in_H":
        z = rng.random(len(v)) < 0.1
        v[z] = 0.0
        row = 5
        v[rp[row]:rp[row + 1]] = 0.0    # a row of zeros keeps (i, 0.0)
    elif kind == "unbounded_sides":
        lower[rng.random(n) < 0.3] = np.nan
        upper[rng.random(n) < 0.3] = np.nan
    return rp, ci, v, x, g, lower, upper


@pytest.mark.parametrize("kind", ["interior", "on_bounds", "zeros_in_H", "unbounded_sides"])
def test_scaled_hessian_matches_reference(ref, kind):
    rng = np.random.default_rng(["interior", "on_bounds", "zeros_in_H", "unbounded_sides"].index(kind) + 1)
    rp, ci, v, x, g, lower, upper = _case(kind, rng)
    H = ref.matrix(rp, ci, v)
    Mr, Dr = ref.rn_get_M_D(H, g, x, lower, upper)
    ref.free(H)
    # RN_set_D_J: the bound in the direction of -g
    n = len(x)
    D = np.ones(n)
    J = np.zeros(n)
    for i in range(n):
        b = upper[i] if g[i] < 0 else lower[i]
        if not np.isnan(b):
            D[i] = np.sqrt(abs(x[i] - b))
            J[i] = 1.0
    assert np.array_equal(D, Dr)
    mrp, mci, mv = lina.rn_scaled_hessian(rp, ci, v, D, g, J)
    assert np.array_equal(mrp, Mr.rowptr)
    assert np.array_equal(mci, Mr.col)
    assert np.array_equal(mv, Mr.val)
    if kind == "on_bounds":
        assert len(mci) < len(ci)      # rows/columns of D_i = 0 were deleted
