"""The oracle's own smoothers, pinned without trusting a second copy of them.

The reference holds one smoother of the SOR family, AMG_SOR_smoother
(Multigrid.c:1345-1365): a forward sweep over the rows in their stored
order.  The multicolour sweeps, the backward SSOR sweep and l1-Jacobi of the
oracle (oracle/lina_oracle.c, smooth_csr) have no counterpart there.  Each
is pinned by an identity that the mathematics fixes:

* a multicolour SOR sweep is the reference's lexicographic SOR on the
  matrix whose rows (and columns) are renumbered in colour order -- rows of
  one colour do not couple, so visiting a colour's rows in turn changes
  nothing -- with every row's entries kept in their stored order, so each
  row sum is formed in the same order: the two agree bit for bit;
* the post-smoothing call of the symmetric multicolour sweep is the same
  with the colours in reverse order, and the backward SSOR sweep is the
  reference's SOR on the matrix with the row order reversed;
* undamped parallel Jacobi is the reference's AMG_parallel_Jacobi_smoother
  (Multigrid.c:1367-1400), bit for bit;
* l1-Jacobi is x + D^-1 (f - A x) with D the row sums of |A|, computed
  densely with numpy;
* a V-cycle with pre- and post-smoothers adjoint to each other (the
  symmetric multicolour SSOR, SSOR, l1-Jacobi below) and R = P^T is a
  symmetric operator, u.M^-1 v = v.M^-1 u to rounding, and a positive one;
  the forward-only multicolour SOR is not symmetric.

Planted mistakes in the oracle (orc_plant) show the checks have teeth: a
stale gather in the multicolour sweep fails the first identity, forward
colour order in the symmetric sweep's post call fails the symmetry check.
The harness smoothers that generated the golden V-cycle vectors
(oracle/ref_harness.c) are held to the same identities."""
import ctypes as C

import numpy as np
import pytest

from fem_amd import problems as P
from oracle.bind import Oracle, dptr, iptr

SYSTEMS = {
    "poisson33": lambda: P.poisson(33),
    "elasticity17": lambda: P.degraded_elasticity(17),
    "damage17": lambda: P.damage_system(17)[0],
}


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.L.orc_smooth.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                               C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.c_int, C.c_double, C.c_int, C.c_int]
    o.L.orc_plant.argtypes = [C.c_int]
    yield o
    o.L.orc_plant(0)


def _sys(name):
    s = SYSTEMS[name]()
    rp = np.ascontiguousarray(s.rowptr, np.int32)
    ci = np.ascontiguousarray(s.col, np.int32)
    v = np.ascontiguousarray(s.val, np.float64)
    rng = np.random.default_rng(len(name))
    return s, rp, ci, v, rng.uniform(-1, 1, s.n), rng.uniform(-1, 1, s.n)


def _colouring(orc, rp, ci, v):
    col = np.zeros(len(rp) - 1, np.int32)
    nc = orc.L.orc_colouring(len(rp) - 1, iptr(rp), iptr(ci), dptr(v), iptr(col))
    return col, nc


def _renumbered(rp, ci, v, order):
    """rows in `order` (new row k = old row order[k]), columns renumbered the
    same way, each row's entries kept in their stored order"""
    n = len(rp) - 1
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    lens = np.diff(rp)[order]
    rp2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    take = np.concatenate([np.arange(rp[i], rp[i + 1]) for i in order])
    return rp2, inv[ci[take]].astype(np.int32), v[take].copy()


def _orc_smooth(orc, rp, ci, v, F, X, sm, post, w=1.0, sweeps=1, colour=None, nc=0, dof=1):
    X = X.copy()
    cp = iptr(colour) if colour is not None else None
    orc.L.orc_smooth(len(rp) - 1, iptr(rp), iptr(ci), dptr(v), cp, nc, dof, dptr(F), dptr(X), sweeps, w, sm, post)
    return X


def _ref_call(ref, fn, rp, ci, v, F, X, w=1.0, sweeps=1, dof=1):
    """a reference smoother (Multigrid.c signature) on the row list of (rp,
    ci, v) as stored; every row labelled COARSE, fix = UNDECIDED"""
    M = ref.matrix(rp, ci, v)
    X = X.copy()
    F = np.ascontiguousarray(F, np.float64)
    lab = np.ones(len(rp) - 1, np.int32)
    f = getattr(ref.L, fn)
    f.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_double,
                  C.c_int, C.c_int, C.c_int]
    f.restype = None
    f(M, dptr(F), dptr(X), iptr(lab), w, 0, sweeps, dof)
    ref.free(M)
    return X


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("w,sweeps", [(1.0, 1), (1.3, 2)])
def test_multicolour_sor_is_reference_sor_in_colour_order(orc, ref, name, w, sweeps):
    s, rp, ci, v, F, X = _sys(name)
    col, nc = _colouring(orc, rp, ci, v)
    assert nc >= 2
    order = np.argsort(col, kind="stable")
    rp2, ci2, v2 = _renumbered(rp, ci, v, order)
    want = _ref_call(ref, "AMG_SOR_smoother", rp2, ci2, v2, F[order], X[order], w, sweeps)
    got = _orc_smooth(orc, rp, ci, v, F, X, 2, 0, w, sweeps, col, nc)
    assert np.array_equal(got[order], want)
    # the symmetric sweep's pre-smoothing call is the same forward sweep
    assert np.array_equal(_orc_smooth(orc, rp, ci, v, F, X, 4, 0, w, sweeps, col, nc), got)
    # the harness smoother behind the golden V-cycle vectors
    assert np.array_equal(_ref_call(ref, "refh_mc_sor_smoother", rp, ci, v, F, X, w, sweeps), got)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_symmetric_multicolour_post_sweep_is_reference_sor_in_reverse_colour_order(orc, ref, name):
    s, rp, ci, v, F, X = _sys(name)
    col, nc = _colouring(orc, rp, ci, v)
    order = np.lexsort((np.arange(s.n), -col))      # colours descending, rows ascending inside
    rp2, ci2, v2 = _renumbered(rp, ci, v, order)
    want = _ref_call(ref, "AMG_SOR_smoother", rp2, ci2, v2, F[order], X[order], 1.0, 1)
    got = _orc_smooth(orc, rp, ci, v, F, X, 4, 1, 1.0, 1, col, nc)
    assert np.array_equal(got[order], want)
    # the harness form alternates direction on successive calls for one
    # matrix object: forward (pre), then backward (post)
    ref.L.refh_sym_reset()
    M = ref.matrix(rp, ci, v)
    lab = np.ones(s.n, np.int32)
    f = ref.L.refh_mc_ssor_smoother
    f.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_double,
                  C.c_int, C.c_int, C.c_int]
    Xa, Xb = X.copy(), X.copy()
    f(M, dptr(F), dptr(Xa), iptr(lab), 1.0, 0, 1, 1)             # pre call: forward
    f(M, dptr(F), dptr(Xb), iptr(lab), 1.0, 0, 1, 1)             # post call: backward
    ref.free(M)
    ref.L.refh_sym_reset()
    assert np.array_equal(Xa, _orc_smooth(orc, rp, ci, v, F, X, 4, 0, 1.0, 1, col, nc))
    assert np.array_equal(Xb, got)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_backward_ssor_sweep_is_reference_sor_on_reversed_rows(orc, ref, name):
    s, rp, ci, v, F, X = _sys(name)
    order = np.arange(s.n)[::-1].copy()
    rp2, ci2, v2 = _renumbered(rp, ci, v, order)
    want = _ref_call(ref, "AMG_SOR_smoother", rp2, ci2, v2, F[order], X[order], 1.2, 2)
    got = _orc_smooth(orc, rp, ci, v, F, X, 5, 1, 1.2, 2)
    assert np.array_equal(got[order], want)
    fwd = _ref_call(ref, "AMG_SOR_smoother", rp, ci, v, F, X, 1.2, 2)
    assert np.array_equal(_orc_smooth(orc, rp, ci, v, F, X, 5, 0, 1.2, 2), fwd)
    assert np.array_equal(_orc_smooth(orc, rp, ci, v, F, X, 0, 0, 1.2, 2), fwd)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_parallel_jacobi_matches_reference(orc, ref, name):
    s, rp, ci, v, F, X = _sys(name)
    for dof in sorted({1, s.dof}):
        want = _ref_call(ref, "AMG_parallel_Jacobi_smoother", rp, ci, v, F, X, 0.7, 3, dof)
        got = _orc_smooth(orc, rp, ci, v, F, X, 1, 0, 0.7, 3, dof=dof)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_l1_jacobi_is_the_dense_step(orc, ref, name):
    s, rp, ci, v, F, X = _sys(name)
    A = s.scipy().toarray()
    d = np.abs(A).sum(axis=1)
    want = X.copy()
    for _ in range(2):
        want = want + (F - A @ want) / d
    got = _orc_smooth(orc, rp, ci, v, F, X, 3, 0, 1.0, 2)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    # the harness form behind the golden vectors
    assert np.array_equal(_ref_call(ref, "refh_l1_jacobi_smoother", rp, ci, v, F, X, 1.0, 2), got)
    # the l1 row sums, not the diagonal: plain damped Jacobi is a different step
    jac = X + (F - A @ X) / np.diag(A)
    assert np.linalg.norm(_orc_smooth(orc, rp, ci, v, F, X, 3, 0, 1.0, 1) - jac) > 1e-3 * np.linalg.norm(jac)


def _vcycle_operator(orc, s, sm, coarse_sm=-1, coarse_sweeps=1, depth=4):
    h = orc.amg_setup(s.rowptr, s.col, s.val, dof=s.dof, depth=depth, smiter=1, aggr_smiter=1)
    orc.L.orc_set_coarse_smoother(h, coarse_sm, coarse_sweeps)

    def apply(r):
        k, x = orc.amg_solve(h, r, np.zeros(s.n), smoother=sm, alpha=1.0, eps=0.0, max_iter=1)
        assert k == 1
        return x
    return h, apply


@pytest.mark.parametrize("name", ["poisson33", "elasticity17"])
@pytest.mark.parametrize("sm,coarse", [("mc_ssor", 3), ("ssor", -1), ("mc_ssor", -1), ("l1_jacobi", -1)])
def test_symmetric_vcycle_is_a_symmetric_positive_operator(orc, name, sm, coarse):
    s = SYSTEMS[name]()
    h, M = _vcycle_operator(orc, s, sm, coarse)
    rng = np.random.default_rng(3)
    try:
        for _ in range(3):
            u, w = rng.normal(size=s.n), rng.normal(size=s.n)
            Mu, Mw = M(u), M(w)
            assert abs(u @ Mw - w @ Mu) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Mw)
            assert u @ Mu > 0
    finally:
        orc.amg_free(h)


def test_forward_multicolour_vcycle_is_not_symmetric(orc):
    """the check above separates: the forward-only sweep fails it"""
    s = SYSTEMS["poisson33"]()
    h, M = _vcycle_operator(orc, s, "mc_sor")
    rng = np.random.default_rng(3)
    u, w = rng.normal(size=s.n), rng.normal(size=s.n)
    Mu, Mw = M(u), M(w)
    orc.amg_free(h)
    assert abs(u @ Mw - w @ Mu) > 1e-6 * np.linalg.norm(u) * np.linalg.norm(Mw)


# ---------------------------------------------------------- planted mistakes

def test_planted_stale_gather_is_detected(orc, ref):
    s, rp, ci, v, F, X = _sys("poisson33")
    col, nc = _colouring(orc, rp, ci, v)
    order = np.argsort(col, kind="stable")
    rp2, ci2, v2 = _renumbered(rp, ci, v, order)
    want = _ref_call(ref, "AMG_SOR_smoother", rp2, ci2, v2, F[order], X[order])
    orc.L.orc_plant(1)
    try:
        got = _orc_smooth(orc, rp, ci, v, F, X, 2, 0, 1.0, 1, col, nc)
    finally:
        orc.L.orc_plant(0)
    assert not np.array_equal(got[order], want)


def test_planted_forward_post_sweep_is_detected(orc):
    s = SYSTEMS["poisson33"]()
    orc.L.orc_plant(2)
    try:
        h, M = _vcycle_operator(orc, s, "mc_ssor", 3)
        rng = np.random.default_rng(3)
        u, w = rng.normal(size=s.n), rng.normal(size=s.n)
        Mu, Mw = M(u), M(w)
        orc.amg_free(h)
    finally:
        orc.L.orc_plant(0)
    assert abs(u @ Mw - w @ Mu) > 1e-6 * np.linalg.norm(u) * np.linalg.norm(Mw)
