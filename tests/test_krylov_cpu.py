"""ILU(0) on the host, checked without a GPU: the factors are the
reference's bit for bit, and L*U reproduces A on A's pattern (the defining
property of ILU(0))."""
import numpy as np
import pytest

from fem_amd import lina
from fem_amd import problems as P
from oracle.bind import CSR


@pytest.mark.parametrize("kind", ["poisson", "damage", "elasticity"])
def test_ilu_factors_match_reference(ref, kind):
    s = {"poisson": lambda: P.poisson(33), "damage": lambda: P.damage_system(33)[0],
         "elasticity": lambda: P.degraded_elasticity(17)}[kind]()
    M = ref.matrix(s.rowptr, s.col, s.val)
    ok, Lr, Ur = ref.ilu(M)
    ref.free(M)
    il = lina.ILU(s.rowptr, s.col, s.val, upload=False)
    assert il.ok == bool(ok)
    for w, R in (("L", Lr), ("U", Ur)):
        rp, ci, v = il.factor(w)
        assert np.array_equal(rp, R.rowptr) and np.array_equal(ci, R.col) and np.array_equal(v, R.val)


def test_ilu_reproduces_pattern_of_A():
    s = P.damage_system(17)[0]
    il = lina.ILU(s.rowptr, s.col, s.val, upload=False)
    L = CSR(*il.factor("L")).dense(s.n)
    U = CSR(*il.factor("U")).dense(s.n)
    A = s.scipy().toarray()
    LU = L @ U
    mask = A != 0
    assert np.allclose(LU[mask], A[mask], rtol=1e-12, atol=1e-12 * np.abs(A).max())
    assert np.allclose(np.diag(U), 1.0)
    assert np.all(np.triu(L, 1) == 0) and np.all(np.tril(U, -1) == 0)


@pytest.mark.parametrize("field,value", [("cycle", 2), ("amg_max_iter", 100), ("zero_init", 0), ("smoother", 9)])
def test_batch_rejects_unsupported_cycle_settings(field, value):
    """lina_batch_create validates its parameters before touching a device:
    the batch preconditions with one V-cycle from a zero guess, so a full
    multigrid cycle, several V-cycles or a warm start are refused instead of
    silently ignored."""
    import ctypes as C
    sp_ = [P.poisson(9), P.poisson(11)]
    row_off, rp, ci, v, _ = P.block_diagonal(sp_)
    prm = lina.params("vcycle", 1, 3)
    setattr(prm, field, value)
    L = lina.lib()
    L.lina_batch_create.restype = C.c_void_p
    ro = np.ascontiguousarray(row_off, dtype=np.int32)
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    h = L.lina_batch_create(len(sp_), ro.ctypes.data_as(C.POINTER(C.c_int)), rp.ctypes.data_as(C.POINTER(C.c_int)),
                            ci.ctypes.data_as(C.POINTER(C.c_int)), v.ctypes.data_as(C.POINTER(C.c_double)),
                            C.byref(prm))
    assert not h
    msg = L.lina_last_error().decode()
    assert ("V-cycle" in msg) if field != "smoother" else ("smoother" in msg)


def _permuted(rp, ci, v, perm):
    """P A P^T: row r is row perm[r] of A, columns renumbered and sorted"""
    n = len(rp) - 1
    iperm = np.empty(n, np.int64)
    iperm[perm] = np.arange(n)
    prp = np.zeros(n + 1, np.int32)
    pci, pv = [], []
    for r in range(n):
        i = perm[r]
        c = iperm[ci[rp[i]:rp[i + 1]]]
        o = np.argsort(c, kind="stable")
        pci.append(c[o])
        pv.append(v[rp[i]:rp[i + 1]][o])
        prp[r + 1] = prp[r] + len(o)
    return prp, np.concatenate(pci).astype(np.int32), np.concatenate(pv)


def with_explicit_zeros(s, frac=0.15, seed=2):
    """s with the same pattern and a seeded share of its off-diagonal values
    set to stored zeros (the ILU(0) pattern keeps them; elimination fills them)"""
    import dataclasses
    rows = np.repeat(np.arange(s.n), np.diff(s.rowptr))
    off = np.flatnonzero(s.col != rows)
    v = s.val.copy()
    v[np.random.default_rng(seed).choice(off, int(frac * len(off)), replace=False)] = 0.0
    return dataclasses.replace(s, val=v)


@pytest.mark.parametrize("kind", ["poisson", "damage", "elasticity", "damage_zeros"])
def test_multicolour_ilu_is_the_reference_ilu_of_the_permuted_matrix(ref, kind):
    """LINA_ILU_MULTICOLOUR: the permutation sorts the rows by a proper
    colouring of A's pattern, explicit zeros included (no two rows of one
    colour coupled: every colour's diagonal block of P A P^T is diagonal), the factors are the compiled reference's
    incomplete_LU_factorization of P A P^T bit for bit, and L's level
    schedule has no more levels than there are colours."""
    from oracle.bind import Oracle
    s = {"poisson": lambda: P.poisson(33), "damage": lambda: P.damage_system(33)[0],
         "elasticity": lambda: P.degraded_elasticity(17),
         "damage_zeros": lambda: with_explicit_zeros(P.damage_system(33)[0])}[kind]()
    il = lina.ILU(s.rowptr, s.col, s.val, upload=False, ordering="multicolour")
    perm = il.permutation()
    ncol, col = Oracle().colouring(s.rowptr, s.col, np.ones(len(s.col)))
    assert np.array_equal(perm, np.argsort(col, kind="stable"))
    prp, pci, pv = _permuted(s.rowptr, s.col, s.val, perm)
    cs = col[perm]
    for r in range(s.n):
        same = cs[pci[prp[r]:prp[r + 1]]] == cs[r]
        assert np.array_equal(pci[prp[r]:prp[r + 1]][same], [r])
    M = ref.matrix(prp, pci, pv)
    ok, Lr, Ur = ref.ilu(M)
    ref.free(M)
    assert ok and il.ok
    for w, R in (("L", Lr), ("U", Ur)):
        rp, ci, v = il.factor(w)
        assert np.array_equal(rp, R.rowptr) and np.array_equal(ci, R.col) and np.array_equal(v, R.val)
    rp, ci, _ = il.factor("L")
    lev = np.zeros(s.n, np.int64)
    for i in range(s.n):
        lower = ci[rp[i]:rp[i + 1] - 1]
        lev[i] = 1 + lev[lower].max() if len(lower) else 0
    assert lev.max() + 1 <= ncol


def test_ordered_ilu_and_rn_reject_unknown_orderings():
    """an unknown ILU ordering is an argument error (no handle, a message),
    for the ILU system and for the reflective Newton's inner solver"""
    s = P.poisson(9)
    lib = lina.lib()
    lina._bind_krylov(lib)
    h = lib.lina_ilu_create_ordered(s.n, lina._i(s.rowptr), lina._i(s.col), lina._d(s.val), 7, 0)
    assert not h
    assert b"ordering" in lib.lina_last_error()
    prm = lina.rn_params("pcg", 1, 3, ilu_ordering=7)
    with pytest.raises(lina.LinaError, match="ordering"):
        lina.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, np.zeros(s.n), np.ones(s.n), prm)
