"""GPU parity of the Krylov drivers beside AMG-PCG, against the oracle's C
restatement (oracle/liboracle.so, itself pinned bit for bit to the compiled
reference by tests/test_oracle_krylov_cpu.py) on the identical systems:

* ILU(0) triangular solves (PCG_preconlr) -- same factors bit for bit
  (tests/test_krylov_cpu.py), level-scheduled device solve;
* pcg with the ILU(0) preconditioner of PCG_default;
* GMRES(30) + ILU(0) as FEM.c drives it (restart-only convergence test);
* BiCGStab_solve with BI_set_AMG_preconditioner's AMG.
"""
import numpy as np
import pytest

from fem_amd import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def lina_mod(gpu):
    return gpu


@pytest.fixture(scope="module")
def orc():
    from oracle.bind import Oracle
    return Oracle()


def _systems():
    return [("poisson33", P.poisson(33)), ("damage33", P.damage_system(33)[0]),
            ("elast17", P.degraded_elasticity(17))]


@pytest.mark.parametrize("name,s", _systems(), ids=lambda v: v if isinstance(v, str) else "")
def test_ilu_apply_matches_reference(lina_mod, orc, name, s):
    say("ilu apply", name)
    ok, Lo, Uo = orc.ilu0(s.rowptr, s.col, s.val)
    assert ok
    z = np.cos(np.arange(s.n) * 0.3)
    w_ref = orc.ilu_apply(Lo, Uo, z)
    il = lina_mod.ILU(s.rowptr, s.col, s.val)
    w = il.apply(z)
    il.close()
    # the sweeps perform PCG_preconl/r's operations in their order: bit for bit
    assert np.array_equal(w, w_ref), rel(w, w_ref)


def _ilu_stress(n=20000, seed=5):
    """A pattern for the level-packed sweeps' edge paths: rows 0..2999 have
    no lower neighbours (a first level of 3000 rows, wider than one
    workgroup), later rows couple to 0..12 random earlier rows (rows with more
    off-diagonal entries than the register slots hold), some of them more than
    the LDS ring's 8192 rows back (results read through global memory);
    symmetric pattern, diagonally dominant."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(3000, n):
        k = int(rng.integers(0, 13))
        js = rng.integers(0, i, size=k)
        if i > 12000 and k:
            js[0] = int(rng.integers(0, i - 9000))
        for j in set(int(j) for j in js):
            rows += [i, j]
            cols += [j, i]
    vals = -rng.uniform(0.1, 1.0, len(rows))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A = A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def test_ilu_apply_wide_levels_long_rows_far_sources(lina_mod, orc):
    rp, ci, v = _ilu_stress()
    n = len(rp) - 1
    say("ilu apply stress", n, "nnz", len(ci))
    ok, Lo, Uo = orc.ilu0(rp, ci, v)
    assert ok
    z = np.sin(np.arange(n) * 0.37)
    w_ref = orc.ilu_apply(Lo, Uo, z)
    il = lina_mod.ILU(rp, ci, v)  # host factorisation, value-based levels
    w1 = il.apply(z)
    assert il.refactor(rp, ci, v)  # device factorisation, pattern levels
    w2 = il.apply(z)
    il.close()
    assert np.array_equal(w1, w_ref), rel(w1, w_ref)
    assert np.array_equal(w2, w_ref), rel(w2, w_ref)


@pytest.mark.parametrize("name,s", _systems()[:2], ids=lambda v: v if isinstance(v, str) else "")
def test_ilu_pcg_matches_reference(lina_mod, orc, name, s):
    say("ilu pcg", name)
    r = orc.pcg_ilu(s.rowptr, s.col, s.val, s.b, tol=1e-10, max_iter=2000)
    il = lina_mod.ILU(s.rowptr, s.col, s.val)
    g = il.pcg(s.b, tol=1e-10, max_iter=2000)
    say("  iters oracle", r["iters"], "gpu", g["iters"], "rel", rel(g["x"], r["x"]))
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel(g["x"], r["x"]) < 1e-12
    assert abs(g["res"] / r["res"] - 1) < 1e-6


@pytest.mark.parametrize("name,s,m", [("poisson33", P.poisson(33), 30), ("elast17", P.degraded_elasticity(17), 30),
                                      ("damage33", P.damage_system(33)[0], 10)])
def test_gmres_ilu_matches_reference(lina_mod, orc, name, s, m):
    say("gmres", name, m)
    r = orc.gmres_ilu(s.rowptr, s.col, s.val, s.b, m=m, tol=1e-10, max_iter=3000)
    il = lina_mod.ILU(s.rowptr, s.col, s.val)
    g = il.gmres(s.b, m=m, tol=1e-10, max_iter=3000)
    A = s.scipy()
    rr = np.linalg.norm(s.b - A @ g["x"]) / np.linalg.norm(s.b)
    say("  iters oracle", r["iters"], "gpu", g["iters"], "rel", rel(g["x"], r["x"]), "relres", rr)
    assert g["iters"] == r["iters"]
    assert rel(g["x"], r["x"]) < 1e-9
    assert rr < 1e-8


@pytest.mark.parametrize("name,s,depth", [("poisson33", P.poisson(33), 4), ("elast17", P.degraded_elasticity(17), 4)])
def test_bicgstab_amg_matches_reference(lina_mod, orc, name, s, depth):
    say("bicgstab", name)
    r = orc.bicgstab_amg(s.rowptr, s.col, s.val, s.b, dof=s.dof, depth=depth, smoother="sor", alpha=1.0, eps=1e-8,
                         max_iter=200)
    prm = lina_mod.params("reference", s.dof, depth, omega=1.0, amg_eps=1e-4, amg_max_iter=20)
    amg = lina_mod.AMG(s.rowptr, s.col, s.val, prm)
    g = lina_mod.bicgstab(amg, s.b, eps=1e-8, max_iter=200)
    say("  iters oracle", r["iters"], "gpu", g["iters"], "rel", rel(g["x"], r["x"]))
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel(g["x"], r["x"]) < 1e-10


def _sorted_rows(rp, ci):
    return all(np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0) for i in range(len(rp) - 1))


@pytest.mark.parametrize("name,s", _systems() + [("damage129", P.damage_system(129)[0])],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_ilu_refactor_device_matches_reference(lina_mod, orc, name, s):
    """lina_ilu_refactor runs ILU(0) on the device (rows sorted, so the device
    path is taken): the factors are bit-identical to the reference's
    incomplete_LU_factorization on a first pattern, on new values of the same
    pattern (values-only upload), and back on the first values."""
    say("ilu refactor", name)
    assert _sorted_rows(s.rowptr, s.col)
    rng = np.random.default_rng(11)
    val2 = s.val * (1.0 + 0.2 * rng.uniform(-1, 1, len(s.val)))
    il = lina_mod.ILU(s.rowptr, s.col, s.val)
    for vals in (s.val, val2, s.val):
        assert il.refactor(s.rowptr, s.col, vals)
        ok, Lr, Ur = orc.ilu0(s.rowptr, s.col, vals)
        assert ok
        for which, R in (("L", Lr), ("U", Ur)):
            rp, ci, v = il.factor(which)
            assert np.array_equal(rp, R.rowptr) and np.array_equal(ci, R.col)
            assert np.array_equal(v, R.val), f"{which} factor differs"
    # the device factors drive the triangular solves as the host ones do
    z = np.cos(np.arange(s.n) * 0.3)
    w = il.apply(z)
    host = lina_mod.ILU(s.rowptr, s.col, s.val)
    assert np.array_equal(w, host.apply(z))
    il.close()
    host.close()


def test_ilu_refactor_zero_pivot(lina_mod, orc):
    """A zero pivot is reported as by the reference (lina_ilu_ok 0), and a
    later refactor with good values of the same pattern recovers."""
    s = P.poisson(9)
    rp, ci, v = s.rowptr.copy(), s.col.copy(), s.val.copy()
    v0 = v.copy()
    for p in range(rp[3], rp[4]):
        if ci[p] == 3:
            v0[p] = 0.0  # zero diagonal in row 3, which has upper entries
    il = lina_mod.ILU(rp, ci, v)
    assert il.refactor(rp, ci, v)
    ok_ref, _, _ = orc.ilu0(rp, ci, v0)
    assert not ok_ref
    assert not il.refactor(rp, ci, v0)
    assert il.refactor(rp, ci, v)
    host = lina_mod.ILU(rp, ci, v)
    for which in ("L", "U"):
        for x, y in zip(il.factor(which), host.factor(which)):
            assert np.array_equal(x, y)
    il.close()
    host.close()


def test_ilu_gmres_after_pattern_change(lina_mod):
    """GMRES on an ILU system whose matrix pattern changed by refactor (the
    reflective-Newton steps' use): the re-upload drops the old Krylov basis,
    and the solve equals a fresh system's bit for bit."""
    s1 = P.poisson(33)
    b = np.sin(np.arange(s1.n) * 0.1) + 1.0
    # second pattern: the couplings (i, i+1) of every 7th row removed
    # symmetrically, their magnitude moved onto both diagonals
    rp, ci, v = s1.rowptr, s1.col, s1.val.copy()
    drop = set()
    for i in range(0, s1.n - 1, 7):
        for p in range(rp[i], rp[i + 1]):
            if ci[p] == i + 1:
                drop.add((i, i + 1))
                drop.add((i + 1, i))
    rows, cols, vals = [], [], []
    for i in range(s1.n):
        for p in range(rp[i], rp[i + 1]):
            if (i, ci[p]) not in drop:
                rows.append(i)
                cols.append(ci[p])
                vals.append(v[p])
    import scipy.sparse as sp
    A2 = sp.csr_matrix((vals, (rows, cols)), shape=(s1.n, s1.n))
    A2.sort_indices()
    A2 = A2 + sp.diags(np.asarray(abs(s1.val).max()) * np.isin(np.arange(s1.n), [i for i, _ in drop]))
    A2 = sp.csr_matrix(A2)
    A2.sort_indices()
    rp2, ci2, v2 = A2.indptr.astype(np.int32), A2.indices.astype(np.int32), A2.data.astype(np.float64)
    assert len(ci2) < len(s1.col)
    il = lina_mod.ILU(s1.rowptr, s1.col, s1.val)
    out1 = il.gmres(b, m=30, tol=1e-8, max_iter=500)
    assert out1["eta"] <= 1e-8
    assert il.refactor(rp2, ci2, v2)
    out2 = il.gmres(b, m=30, tol=1e-8, max_iter=500)
    fresh = lina_mod.ILU(rp2, ci2, v2)
    ref2 = fresh.gmres(b, m=30, tol=1e-8, max_iter=500)
    assert out2["iters"] == ref2["iters"]
    assert np.array_equal(out2["x"], ref2["x"])
    il.close()
    fresh.close()
