"""Pin the oracle (oracle/lina_oracle.c) before trusting it.

1. Golden vectors written by the compiled reference (tests/golden/*.npz):
   the oracle must reproduce the hierarchy, AMG_solve and pcg bit for bit.
2. Live comparison against oracle/_ref where it is built.
3. What the mathematics fixes independently of the reference: SpMV against a
   dense product, the coarse solve against numpy.linalg.solve, pcg against a
   direct solve, and the AMG invariants (C/F split, P rows, R = P^T,
   A_c = R A P).
4. Planted mistakes: a perturbed oracle input must fail the golden checks,
   so the checks above can catch a dropped term or a flipped tie rule.
"""
import glob
import os

import numpy as np
import pytest

from fem_amd import problems as P
from oracle.bind import CSR, Oracle

# the hierarchy fixtures (krylov.npz holds the Krylov drivers' vectors)
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not p.endswith("krylov.npz") and not os.path.basename(p).startswith("direct_"))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _levels(orc, g, **kw):
    depth = int(g["depth"])
    h = orc.amg_setup(g["rowptr"], g["col"], g["val"], dof=int(g["dof"]), depth=depth,
                      coarsening=int(g["coarsening"]), **kw)
    return h, orc.amg_levels(h, depth)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_hierarchy_matches_golden(orc, path):
    g = np.load(path)
    h, L = _levels(orc, g)
    try:
        for l, lv in enumerate(L):
            for w in ("A", "P", "R"):
                if w not in lv:
                    continue
                assert np.array_equal(lv[w].rowptr, g[f"L{l}_{w}_rp"]), (l, w)
                assert np.array_equal(lv[w].col, g[f"L{l}_{w}_ci"]), (l, w)
                assert np.array_equal(lv[w].val, g[f"L{l}_{w}_v"]), (l, w)
            if l > 0:
                assert np.array_equal(lv["labels"], g[f"L{l}_labels"]), l
    finally:
        orc.amg_free(h)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
@pytest.mark.parametrize("sm,alpha", [("sor", 1.6), ("jacobi", 1.0), ("mc_sor", 1.2), ("l1_jacobi", 1.0),
                                      ("mc_ssor", 1.0), ("ssor", 1.0)])
def test_oracle_amg_solve_matches_golden(orc, path, sm, alpha):
    g = np.load(path)
    h, _ = _levels(orc, g)
    try:
        k, x = orc.amg_solve(h, g["b"], np.zeros(len(g["b"])), smoother=sm, alpha=alpha, eps=1e-6, max_iter=6)
    finally:
        orc.amg_free(h)
    assert k == int(g[f"amgsolve_{sm}_k"])
    assert np.array_equal(x, g[f"amgsolve_{sm}_x"])


@pytest.mark.parametrize("path", [p for p in GOLDEN if "std" not in p], ids=lambda p: os.path.basename(p))
def test_oracle_pcg_matches_golden(orc, path):
    g = np.load(path)
    dof, depth = int(g["dof"]), int(g["depth"])
    r = orc.pcg_amg(g["rowptr"], g["col"], g["val"], g["b"], dof=dof, depth=depth, tol=1e-10, max_iter=500)
    assert r["iters"] == int(g["pcgref_iters"])
    assert r["vcycles"] == int(np.sum(g["pcgref_vc"]))
    assert np.array_equal(r["x"], g["pcgref_x"])
    assert np.array_equal(r["b_out"], g["pcgref_b"])
    assert r["res"] == float(g["pcgref_res"])
    r = orc.pcg_amg(g["rowptr"], g["col"], g["val"], g["b"], dof=dof, depth=depth, smiter=1, aggr_smiter=1,
                    smoother="mc_ssor", alpha=1.0, amg_eps=1e-2, amg_max_iter=1, zero_init=True, tol=1e-8,
                    max_iter=2000)
    assert r["iters"] == int(g["pcgvc_iters"])
    assert np.array_equal(r["x"], g["pcgvc_x"])


def test_oracle_live_against_reference(orc, ref):
    """Cases outside the fixtures, compared with the compiled reference."""
    for s, depth, coars in ((P.poisson(65), 4, 1), (P.degraded_elasticity(49), 4, 1), (P.poisson(49), 5, 2)):
        M = ref.matrix(s.rowptr, s.col, s.val)
        I = ref.amg_setup(M, dof=s.dof, depth=depth, coarsening=coars)
        R = ref.amg_levels(I, depth)
        h = orc.amg_setup(s.rowptr, s.col, s.val, dof=s.dof, depth=depth, coarsening=coars)
        O = orc.amg_levels(h, depth)
        for l in range(depth):
            for w in ("A", "P", "R"):
                if w in O[l]:
                    assert np.array_equal(O[l][w].val, R[l][w].val) and np.array_equal(O[l][w].col, R[l][w].col)
        k1, x1 = ref.amg_solve(I, s.b, np.zeros(s.n), smoother="sor", alpha=1.6, eps=1e-5, max_iter=5)
        k2, x2 = orc.amg_solve(h, s.b, np.zeros(s.n), smoother="sor", alpha=1.6, eps=1e-5, max_iter=5)
        assert k1 == k2 and np.array_equal(x1, x2)
        orc.amg_free(h)
        ref.amg_free(I)
        ref.free(M)


@pytest.mark.parametrize("sm", ["sor", "jacobi"])
def test_oracle_full_multigrid_against_reference(orc, ref, sm):
    """AMG_solve with the FULL_AMG cycle structure (one full-multigrid pass,
    Multigrid.c:1582-1611): the oracle's pass bit for bit against the compiled
    reference set up with FULL_AMG, from a zero and from a nonzero initial
    guess (the pass replaces the guess: both give the same result), on
    scalar and two-field systems and both coarsenings; a zero right-hand side
    runs no pass.  With SOR on the scalar systems the pass leaves a smaller
    residual than one V-cycle from zero."""
    rng = np.random.default_rng(5)
    for s, depth, coars in ((P.poisson(65), 4, 1), (P.degraded_elasticity(33), 4, 1), (P.poisson(49), 5, 2)):
        M = ref.matrix(s.rowptr, s.col, s.val)
        I = ref.amg_setup(M, dof=s.dof, depth=depth, coarsening=coars, cycle="full")
        h = orc.amg_setup(s.rowptr, s.col, s.val, dof=s.dof, depth=depth, coarsening=coars)
        xs = []
        for x0 in (np.zeros(s.n), rng.standard_normal(s.n)):
            k1, x1 = ref.amg_solve(I, s.b, x0, smoother=sm, alpha=1.3, eps=1e-30, max_iter=50)
            k2, x2 = orc.amg_solve(h, s.b, x0, smoother=sm, alpha=1.3, cycle="full")
            assert k1 == k2 == 1 and np.array_equal(x1, x2)
            xs.append(x2)
        assert np.array_equal(xs[0], xs[1])
        k1, x1 = ref.amg_solve(I, np.zeros(s.n), np.zeros(s.n), smoother=sm, alpha=1.3)
        k2, x2 = orc.amg_solve(h, np.zeros(s.n), np.zeros(s.n), smoother=sm, alpha=1.3, cycle="full")
        assert k1 == k2 == 0 and not x2.any()
        if sm == "sor" and s.dof == 1:
            # (damped Jacobi at omega 1.3 diverges on these systems, V-cycles alike)
            A = s.scipy()
            e_fmg = np.linalg.norm(s.b - A @ xs[0])
            _, xv = orc.amg_solve(h, s.b, np.zeros(s.n), smoother=sm, alpha=1.3, eps=1e-30, max_iter=1)
            assert e_fmg < np.linalg.norm(s.b - A @ xv) < np.linalg.norm(s.b)
        orc.amg_free(h)
        ref.amg_free(I)
        ref.free(M)


# ------------------------------------------------- mathematics-fixed checks

def test_spmv_equals_dense_product(orc):
    s = P.degraded_elasticity(9)
    x = np.cos(np.arange(s.n) * 0.7)
    y = orc.spmv(s.rowptr, s.col, s.val, x)
    D = CSR(s.rowptr, s.col, s.val).dense(s.n)
    assert np.allclose(y, D @ x, rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("case", ["poisson", "elasticity"])
def test_amg_invariants(orc, case):
    s = P.poisson(33) if case == "poisson" else P.degraded_elasticity(25)
    depth = 4
    h = orc.amg_setup(s.rowptr, s.col, s.val, dof=s.dof, depth=depth)
    L = orc.amg_levels(h, depth)
    try:
        for l in range(1, depth):
            lab = L[l]["labels"]
            A = L[l]["A"].dense(L[l]["fine"])
            Pm = L[l]["P"].dense(L[l]["coarse"] + L[l]["fine"])[:, :L[l]["coarse"]]
            Rm = L[l]["R"].dense(L[l]["fine"])
            assert set(np.unique(lab)) <= {1, 2, 3}
            # coarse and isolated points are injected: one unit entry in P
            cmap = np.cumsum(lab != 2) - 1
            for i in np.where(lab != 2)[0][:200]:
                row = Pm[i]
                assert row[cmap[i]] == 1.0 and np.count_nonzero(row) == 1
            # every interpolated row sums to one (normalize_row_sum) or is empty
            sums = Pm.sum(axis=1)
            nz = np.abs(Pm).sum(axis=1) > 0
            assert np.allclose(sums[nz], 1.0, atol=1e-12)
            assert np.array_equal(Rm, Pm.T)
            Ac = L[l - 1]["A"].dense(L[l - 1]["fine"])
            assert np.allclose(Ac, Rm @ A @ Pm, rtol=1e-10, atol=1e-9 * np.abs(Ac).max())
            assert np.allclose(Ac, Ac.T, rtol=1e-9, atol=1e-9 * np.abs(Ac).max())
    finally:
        orc.amg_free(h)


def test_coarse_solve_is_exact(orc):
    """With depth 2 the V-cycle's coarse solve is the dense LU of the Galerkin
    matrix: one AMG_solve step with 0 smoothing iterations returns P A_c^-1 R f."""
    s = P.poisson(17)
    h = orc.amg_setup(s.rowptr, s.col, s.val, dof=1, depth=2, smiter=0, aggr_smiter=0)
    L = orc.amg_levels(h, 2)
    k, x = orc.amg_solve(h, s.b, np.zeros(s.n), smoother="sor", alpha=1.0, eps=1e-30, max_iter=1)
    orc.amg_free(h)
    Ac = L[0]["A"].dense(L[0]["fine"])
    Pm = L[1]["P"].dense(L[1]["coarse"] + s.n)[:, :L[1]["coarse"]]
    f = s.b.copy()
    lab = L[1]["labels"]
    iso = np.where(lab == 3)[0]
    x0 = np.zeros(s.n)
    A = s.scipy()
    x0[iso] = f[iso] / A.diagonal()[iso]
    r = f - A @ x0
    xc = np.linalg.solve(Ac, Pm.T @ r)
    assert k == 1
    assert np.allclose(x, x0 + Pm @ xc, rtol=1e-10, atol=1e-12)


def test_pcg_converges_to_direct_solution(orc):
    s = P.poisson(33)
    r = orc.pcg_amg(s.rowptr, s.col, s.val, s.b, dof=1, depth=4, tol=1e-12, max_iter=200)
    x_direct = np.linalg.solve(s.scipy().toarray(), s.b)
    assert np.linalg.norm(r["x"] - x_direct) / np.linalg.norm(x_direct) < 1e-9


def test_colouring_is_proper(orc):
    s = P.degraded_elasticity(17)
    k, c = orc.colouring(s.rowptr, s.col, s.val)
    for i in range(s.n):
        for p in range(s.rowptr[i], s.rowptr[i + 1]):
            j = s.col[p]
            if j != i and s.val[p] != 0:
                assert c[i] != c[j]
    assert k == c.max() + 1


# --------------------------------------------------------- planted mistakes

def test_planted_threshold_change_is_detected(orc):
    """A wrong strength threshold (a plausible transcription slip) changes the
    C/F split and is caught by the golden hierarchy."""
    g = np.load([p for p in GOLDEN if "poisson33.npz" in p][0])
    h, L = _levels(orc, g, eps_plus=0.3, eps_minus=0.5)
    try:
        same = all(np.array_equal(L[l]["labels"], g[f"L{l}_labels"]) for l in range(1, int(g["depth"])))
        same_vals = all(np.array_equal(L[l]["A"].val, g[f"L{l}_A_v"]) for l in range(int(g["depth"])))
    finally:
        orc.amg_free(h)
    assert not (same and same_vals)


def test_planted_truncation_change_is_detected(orc):
    g = np.load([p for p in GOLDEN if "elast33" in p][0])
    h, L = _levels(orc, g, matrix_tol=0.05)
    try:
        same = all(np.array_equal(L[l]["P"].val, g[f"L{l}_P_v"]) for l in range(1, int(g["depth"])))
    finally:
        orc.amg_free(h)
    assert not same


@pytest.mark.parametrize("case,sm,alpha", [("poisson33", "sor", 1.2), ("poisson33", "mc_ssor", 1.0),
                                            ("elast25", "ssor", 1.0), ("elast25", "mc_ssor", 1.0)])
def test_oracle_pcg_full_multigrid_preconditioner_against_reference(orc, ref, case, sm, alpha):
    """pcg (PCGS.c:163-233) preconditioned by AMG_solve of a FULL_AMG
    hierarchy (one full-multigrid pass per preconditioner call): the oracle's
    iterates, iteration count and final residuum bit for bit against the
    compiled reference, on a scalar and a two-field system (smoothers under
    which the pass is a usable preconditioner there: forward SOR leaves the
    two-field system unconverged after 300 iterations in both)."""
    s, depth = {"poisson33": (P.poisson(33), 4), "elast25": (P.degraded_elasticity(25), 4)}[case]
    M = ref.matrix(s.rowptr, s.col, s.val)
    r = ref.pcg_amg(M, s.b, dof=s.dof, depth=depth, smoother=sm, alpha=alpha, tol=1e-10, max_iter=300, cycle="full")
    ref.free(M)
    o = orc.pcg_amg(s.rowptr, s.col, s.val, s.b, dof=s.dof, depth=depth, smoother=sm, alpha=alpha, tol=1e-10,
                    max_iter=300, cycle="full")
    assert r["iters"] == o["iters"] and 0 < o["iters"] < 300
    assert np.array_equal(r["x"], o["x"]) and r["res"] == o["res"]
    assert np.linalg.norm(s.b - s.scipy() @ o["x"]) <= 1e-8 * np.linalg.norm(s.b)


def test_assembled_hierarchy_runs_the_reference_pcg(ref):
    """The CPU baseline's timing path (oracle/ref_harness.c
    refh_amg_assemble + refh_pcg_assembled): the reference's pcg and V-cycle
    on a hierarchy handed over from liblina_hip's host setup, the coarsest
    level solved with a dense inverse.  On a 65x65-node member of the
    headline family it takes the reference's own AMG_setup + pcg iterations
    (the same symmetric SOR V(1,1)) and iterates to 1e-10: the timed work is
    the reference's algorithm on the reference's hierarchy."""
    from fem_amd import lina
    s = P.degraded_elasticity(65)
    depth = 5
    amg = lina.AMG(s.rowptr, s.col, s.val, lina.params("vcycle", 2, depth), upload=False)
    lv = [amg.level(l) for l in range(amg.depth)]
    h, keep = ref.assemble(lv, 2, amg.signs())
    M = ref.matrix(s.rowptr, s.col, s.val)
    got = ref.pcg_assembled(h, M, s.b, tol=1e-8, max_iter=2000)
    want = ref.pcg_amg(M, s.b, dof=2, depth=depth, smiter=1, aggr_smiter=1, smoother="ssor", alpha=1.0, amg_eps=1e-2,
                       amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=2000)
    ref.free(M)
    assert 0 < want["iters"] < 2000 and got["iters"] == want["iters"]
    assert np.linalg.norm(got["x"] - want["x"]) <= 1e-10 * np.linalg.norm(want["x"])
