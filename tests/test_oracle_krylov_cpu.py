"""Pin the oracle's ILU(0), ILU-pcg, GMRES(m)+ILU(0) and BiCGStab+AMG
restatements, and its AMG-pcg with the level-dependent V-cycle smoother.

1. Golden vectors written by the compiled reference (tests/golden/krylov.npz):
   factors, iterates, iteration counts and eta bit for bit.
2. Live comparison with oracle/_ref on systems outside the fixtures.
3. What the mathematics fixes: ILU(0) of a tridiagonal matrix is its exact LU
   (no fill is dropped), with the pivots of an unpivoted Gaussian
   elimination; L U reproduces A on A's pattern; the preconditioner is the two
   triangular solves of scipy; GMRES's eta is the true
   residual reduction of the returned x; the solvers reach the direct
   solution.
4. Planted mistakes (a dropped matrix entry, the Doolittle instead of the
   Crout form of the factors) make the golden checks fail.
"""
import os

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg

from fem_amd import problems as P
from oracle.bind import CSR, Oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "krylov.npz")
NAMES = ["poisson17", "damage17", "elast17"]


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _sys(g, name):
    return g[f"{name}_rowptr"], g[f"{name}_col"], g[f"{name}_val"], g[f"{name}_b"]


def _same(a: CSR, rp, ci, v):
    return np.array_equal(a.rowptr, rp) and np.array_equal(a.col, ci) and np.array_equal(a.val, v)


# ------------------------------------------------------------ golden vectors

@pytest.mark.parametrize("name", NAMES)
def test_ilu0_matches_golden(orc, gold, name):
    rp, ci, v, _ = _sys(gold, name)
    ok, L, U = orc.ilu0(rp, ci, v)
    assert ok
    for w, F in (("L", L), ("U", U)):
        assert _same(F, gold[f"{name}_ilu{w}_rp"], gold[f"{name}_ilu{w}_ci"], gold[f"{name}_ilu{w}_v"]), w


@pytest.mark.parametrize("name", NAMES)
def test_pcg_ilu_matches_golden(orc, gold, name):
    rp, ci, v, b = _sys(gold, name)
    r = orc.pcg_ilu(rp, ci, v, b, tol=1e-10, max_iter=2000)
    assert r["iters"] == int(gold[f"{name}_pcgilu_iters"])
    assert r["res"] == float(gold[f"{name}_pcgilu_res"])
    assert np.array_equal(r["x"], gold[f"{name}_pcgilu_x"])


@pytest.mark.parametrize("name", NAMES)
def test_pcg_vcycle_hybrid_matches_golden(orc, gold, name):
    rp, ci, v, b = _sys(gold, name)
    r = orc.pcg_amg(rp, ci, v, b, dof=int(gold[f"{name}_dof"]), depth=4, smiter=1, aggr_smiter=1, smoother="hybrid",
                    alpha=1.0, amg_eps=1e-2, amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=2000, coarse_sweeps=1)
    assert r["iters"] == int(gold[f"{name}_pcghyb_iters"])
    assert r["res"] == float(gold[f"{name}_pcghyb_res"])
    assert np.array_equal(r["x"], gold[f"{name}_pcghyb_x"])


@pytest.mark.parametrize("name", NAMES)
def test_gmres_matches_golden(orc, gold, name):
    rp, ci, v, b = _sys(gold, name)
    r = orc.gmres_ilu(rp, ci, v, b, m=30, tol=1e-8, max_iter=600)
    assert r["iters"] == int(gold[f"{name}_gmres_iters"])
    assert r["eta"] == float(gold[f"{name}_gmres_eta"])
    assert np.array_equal(r["x"], gold[f"{name}_gmres_x"])


@pytest.mark.parametrize("name", NAMES)
def test_bicgstab_matches_golden(orc, gold, name):
    rp, ci, v, b = _sys(gold, name)
    r = orc.bicgstab_amg(rp, ci, v, b, dof=int(gold[f"{name}_dof"]), depth=4, smoother="sor", alpha=1.0, eps=1e-8,
                         max_iter=300)
    assert r["iters"] == int(gold[f"{name}_bicg_iters"])
    assert np.array_equal(r["x"], gold[f"{name}_bicg_x"])


def test_live_against_reference(orc, ref):
    """Systems outside the fixtures, against the compiled reference."""
    for s, seed in ((P.poisson(33), 4), (P.damage_system(33)[0], 5), (P.degraded_elasticity(25), 6)):
        b = np.random.default_rng(seed).standard_normal(s.n)
        M = ref.matrix(s.rowptr, s.col, s.val)
        ok, Lr, Ur = ref.ilu(M)
        ok2, Lo, Uo = orc.ilu0(s.rowptr, s.col, s.val)
        assert ok == ok2 == 1
        assert _same(Lo, Lr.rowptr, Lr.col, Lr.val) and _same(Uo, Ur.rowptr, Ur.col, Ur.val)
        c1 = ref.pcg_ilu(M, b, tol=1e-9, max_iter=1000)
        c2 = orc.pcg_ilu(s.rowptr, s.col, s.val, b, tol=1e-9, max_iter=1000)
        assert c1["iters"] == c2["iters"] and c1["res"] == c2["res"] and np.array_equal(c1["x"], c2["x"])
        g1 = ref.gmres_ilu(M, b, m=20, tol=1e-9, max_iter=400)
        g2 = orc.gmres_ilu(s.rowptr, s.col, s.val, b, m=20, tol=1e-9, max_iter=400)
        assert g1["iters"] == g2["iters"] and g1["eta"] == g2["eta"] and np.array_equal(g1["x"], g2["x"])
        kw = dict(dof=s.dof, depth=5, smiter=1, aggr_smiter=1, smoother="hybrid", alpha=1.0, amg_eps=1e-2,
                  amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=2000, coarse_sweeps=2)
        h1 = ref.pcg_amg(M, b, **kw)
        h2 = orc.pcg_amg(s.rowptr, s.col, s.val, b, **kw)
        assert h1["iters"] == h2["iters"] and h1["res"] == h2["res"] and np.array_equal(h1["x"], h2["x"])
        b1 = ref.bicgstab_amg(M, b, dof=s.dof, depth=4, smoother="sor", alpha=1.0, eps=1e-8, max_iter=200)
        b2 = orc.bicgstab_amg(s.rowptr, s.col, s.val, b, dof=s.dof, depth=4, smoother="sor", alpha=1.0, eps=1e-8,
                              max_iter=200)
        assert b1["iters"] == b2["iters"] and np.array_equal(b1["x"], b2["x"])
        ref.free(M)


def test_zero_pivot_matches_reference(orc, ref):
    # row 0 has no diagonal but an upper entry: the first pivot is zero
    A = sp.csr_matrix(np.array([[0.0, 1.0, 0.0], [1.0, 2.0, 1.0], [0.0, 1.0, 2.0]]))
    A.sort_indices()
    ok, _, _ = orc.ilu0(A.indptr, A.indices, A.data)
    M = ref.matrix(A.indptr, A.indices, A.data)
    ok_ref, _, _ = ref.ilu(M)
    ref.free(M)
    assert ok == 0 and ok_ref == 0


# ------------------------------------------------- mathematics-fixed checks

def _tridiag(n, seed):
    rng = np.random.default_rng(seed)
    lo, up = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n - 1)
    d = 3.0 + rng.uniform(0, 1, n)
    A = sp.diags([lo, d, up], [-1, 0, 1], format="csr")
    A.sort_indices()
    return A


def test_ilu0_of_tridiagonal_is_exact_lu(orc):
    A = _tridiag(40, 7)
    ok, L, U = orc.ilu0(A.indptr, A.indices, A.data)
    assert ok
    Ld, Ud, Ad = L.dense(40), U.dense(40), A.toarray()
    assert np.allclose(Ld @ Ud, Ad, rtol=0, atol=1e-14)
    # diagonally dominant: no row exchanges, and the Crout L carries the pivots
    Pm, _, Ug = scipy.linalg.lu(Ad)
    assert np.array_equal(Pm, np.eye(40))
    assert np.allclose(np.diag(Ld), np.diag(Ug), rtol=1e-14)
    assert np.all(np.diag(Ud) == 1.0)


def test_ilu_apply_is_the_two_triangular_solves(orc):
    s = P.degraded_elasticity(13)
    ok, L, U = orc.ilu0(s.rowptr, s.col, s.val)
    z = np.cos(np.arange(s.n) * 0.3)
    w = orc.ilu_apply(L, U, z)
    Ls = sp.csr_matrix((L.val, L.col, L.rowptr), shape=(s.n, s.n))
    Us = sp.csr_matrix((U.val, U.col, U.rowptr), shape=(s.n, s.n))
    y = scipy.sparse.linalg.spsolve_triangular(Ls, z, lower=True)
    x = scipy.sparse.linalg.spsolve_triangular(Us, y, lower=False)
    assert np.linalg.norm(w - x) <= 1e-13 * np.linalg.norm(x)


def test_pcg_ilu_reaches_direct_solution(orc):
    s = P.poisson(17)
    r = orc.pcg_ilu(s.rowptr, s.col, s.val, s.b, tol=1e-12, max_iter=500)
    xd = np.linalg.solve(s.scipy().toarray(), s.b)
    assert np.linalg.norm(r["x"] - xd) <= 1e-9 * np.linalg.norm(xd)
    assert np.array_equal(r["b_out"] != 0, s.b != 0)


def test_ilu0_reproduces_a_on_its_pattern(orc):
    s = P.damage_system(17)[0]
    ok, L, U = orc.ilu0(s.rowptr, s.col, s.val)
    assert ok
    A = s.scipy().toarray()
    LU = L.dense(s.n) @ U.dense(s.n)
    mask = A != 0
    assert np.allclose(LU[mask], A[mask], rtol=1e-12, atol=1e-12 * np.abs(A).max())
    assert np.all(np.triu(L.dense(s.n), 1) == 0) and np.all(np.tril(U.dense(s.n), -1) == 0)


def test_gmres_eta_is_the_true_residual_reduction(orc):
    s = P.degraded_elasticity(13)
    A = s.scipy()
    b = np.random.default_rng(3).standard_normal(s.n)
    for m in (5, 12):
        r = orc.gmres_ilu(s.rowptr, s.col, s.val, b, m=m, tol=1e-7, max_iter=2000)
        assert r["iters"] % m == 0  # CheckOnRestart: whole cycles only
        true = np.linalg.norm(b - A @ r["x"]) / np.linalg.norm(b)
        assert r["eta"] <= 1e-7
        assert abs(true - r["eta"]) <= 1e-3 * r["eta"] + 1e-14
        xd = np.linalg.solve(A.toarray(), b)
        assert np.linalg.norm(r["x"] - xd) <= 1e-5 * np.linalg.norm(xd)


def test_bicgstab_reaches_direct_solution(orc):
    s = P.degraded_elasticity(13)
    A = s.scipy()
    b = np.random.default_rng(9).standard_normal(s.n)
    r = orc.bicgstab_amg(s.rowptr, s.col, s.val, b, dof=s.dof, depth=3, eps=1e-10, max_iter=400)
    assert r["res"] <= 1e-10
    xd = np.linalg.solve(A.toarray(), b)
    assert np.linalg.norm(r["x"] - xd) <= 1e-7 * np.linalg.norm(xd)


# ---------------------------------------------------------- planted mistakes

def test_planted_unit_lower_form_is_detected(orc, gold):
    """The reference's factors are Crout's (pivots in L, unit U); a Doolittle
    form (unit L, pivots in U) of the same product fails the golden check."""
    rp, ci, v, _ = _sys(gold, "elast17")
    ok, L, U = orc.ilu0(rp, ci, v)
    assert ok
    d = L.val[L.rowptr[1:] - 1]  # each row's diagonal is its last entry
    Lu = CSR(L.rowptr, L.col, L.val / d[L.col])  # L = Lu D: the pivots move into D U
    assert np.allclose((Lu.dense(len(d)) * d) @ U.dense(len(d)), L.dense(len(d)) @ U.dense(len(d)))
    assert not _same(Lu, gold["elast17_iluL_rp"], gold["elast17_iluL_ci"], gold["elast17_iluL_v"])


def test_planted_dropped_entry_is_detected(orc, gold):
    """An off-diagonal entry of A lost (a dropped term) changes GMRES's iterate."""
    rp, ci, v, b = _sys(gold, "damage17")
    v2 = v.copy()
    off = np.nonzero(ci != np.repeat(np.arange(len(rp) - 1), np.diff(rp)))[0]
    v2[off[len(off) // 2]] = 0.0
    r = orc.gmres_ilu(rp, ci, v2, b, m=30, tol=1e-8, max_iter=600)
    assert not np.array_equal(r["x"], gold["damage17_gmres_x"])
