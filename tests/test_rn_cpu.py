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


def _rules_row(i, cols, h, D, G, J):
    """one row of M by the per-row rules the device assembly applies over
    F = H's pattern plus the diagonal (h: H's entries of the row by column):
    HD present off the diagonal iff h_ik != 0 and d_k != 0, the diagonal
    always (0.0 + h_ii d_i, or a 0.0 seed); zeros deleted when the row has
    more than one entry, the diagonal 0.0 kept when nothing else is left; M
    the same over d_i times HD's kept entries; then |g_i| J_i added to (or
    inserted as) the diagonal"""
    def hd(k):
        a = h.get(k, 0.0)
        val = a * D[k] if (k in h and a != 0 and D[k] != 0) else 0.0
        return (0.0 + val if k in h else 0.0) if k == i else val
    pres = [k for k in cols if k == i or (k in h and h[k] != 0 and D[k] != 0)]
    hd_nz = sum(hd(k) != 0 for k in pres)
    hd_diag = hd(i)
    keep_hd_diag = len(pres) == 1 or hd_diag != 0 or hd_nz == 0
    di = D[i]
    m_diag = 0.0 + di * hd_diag if (keep_hd_diag and di != 0 and hd_diag != 0) else 0.0
    offs = [k for k in pres if k != i and hd(k) != 0 and di != 0]
    m_nz = (m_diag != 0) + sum(di * hd(k) != 0 for k in offs)
    keep_diag = len(offs) == 0 or m_diag != 0 or m_nz == 0
    row = []
    for k in cols:
        if k == i:
            if keep_diag:
                row.append((i, m_diag + abs(G[i]) * J[i] if J[i] != 0 else m_diag))
            elif J[i] != 0:
                row.append((i, abs(G[i]) * J[i]))
        elif k in offs and di * hd(k) != 0:
            row.append((k, di * hd(k)))
    return row


def test_device_assembly_rules_reproduce_host_assembly():
    """the rules k_rn_scaled_hessian applies row by row give the host
    assembly's M entry for entry, on random patterns with missing
    diagonals, explicit zeros in H, zero scalings D and J adds"""
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = 30
        rp, ci, v = [0], [], []
        for i in range(n):
            cols = set(rng.choice(n, size=5, replace=False).tolist())
            (cols.add if rng.random() > 0.3 else cols.discard)(i)
            cols = sorted(cols)
            vals = rng.standard_normal(len(cols))
            vals[rng.random(len(cols)) < 0.2] = 0.0
            ci += cols
            v += vals.tolist()
            rp.append(len(ci))
        rp, ci, v = np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)
        D = rng.standard_normal(n)
        D[rng.random(n) < 0.3] = 0.0
        if trial % 5 == 0:
            D[:] = 0.0
        G = rng.standard_normal(n)
        J = np.where(rng.random(n) < 0.4, rng.random(n), 0.0)
        mrp, mci, mv = lina.rn_scaled_hessian(rp, ci, v, D, G, J)
        for i in range(n):
            h = dict(zip(ci[rp[i]:rp[i + 1]].tolist(), v[rp[i]:rp[i + 1]].tolist()))
            cols = sorted(set(h) | {i})
            got = _rules_row(i, cols, h, D, G, J)
            want = list(zip(mci[mrp[i]:mrp[i + 1]].tolist(), mv[mrp[i]:mrp[i + 1]].tolist()))
            assert got == want, (trial, i, got, want)
