"""ILU(0) in multicolour order (lina_ilu_create_ordered with
LINA_ILU_MULTICOLOUR): the rows and columns of A permuted by the greedy
colouring of its pattern, colour by colour.  Pinned to the reference through
the oracle (itself bit-identical to the compiled reference's
incomplete_LU_factorization, PCG_preconlr, pcg and gmres):

* the permutation is the stable sort of the oracle's colouring of A's
  pattern (explicit zeros included; one case stores 15 % of the damage
  matrix's off-diagonal values as zeros);
* the factors are exactly the reference's incomplete LU of P A P^T, from the
  host factorisation and from the device one (a launch per level);
* the preconditioner applied to z in A's order is P^T U^-1 L^-1 P z of the
  oracle's sweeps on P A P^T, bit for bit;
* pcg / GMRES(30) on A equal the oracle's on the permuted system.

The damage step's reflective Newton with the multicolour ILU-PCG inner
solver is checked against the reference's reflective_newton on the permuted
system in test_gpu_rn.py, and at full size in test_gpu_fullsize.py.
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
def orc():
    from oracle.bind import Oracle
    return Oracle()


def _permuted(rp, ci, v, perm):
    """P A P^T in CSR: row r is row perm[r] of A, columns renumbered, each
    row sorted by its new column"""
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


def _case(name):
    if name == "damage257_zeros":
        from tests.test_krylov_cpu import with_explicit_zeros
        return with_explicit_zeros(P.damage_system(257)[0])
    return {"poisson65": P.poisson(65), "damage129": P.damage_system(129)[0],
            "elast33": P.degraded_elasticity(33), "damage257": P.damage_system(257)[0]}[name]


@pytest.mark.parametrize("name", ["poisson65", "damage129", "elast33", "damage257", "damage257_zeros"])
def test_multicolour_factors_and_sweeps_match_reference(gpu, orc, name):
    s = _case(name)
    ncol, col = orc.colouring(s.rowptr, s.col, np.ones(len(s.col)))
    perm = np.argsort(col, kind="stable").astype(np.int32)
    il = gpu.ILU(s.rowptr, s.col, s.val, ordering="multicolour")
    assert np.array_equal(il.permutation(), perm)
    rng = np.random.default_rng(5)
    val2 = s.val * (1.0 + 0.2 * rng.uniform(-1, 1, len(s.val)))
    z = np.cos(np.arange(s.n) * 0.3)
    for step, vals in enumerate((s.val, val2)):
        if step > 0:
            assert il.refactor(s.rowptr, s.col, vals)   # the device factorisation
        prp, pci, pv = _permuted(s.rowptr, s.col, vals, perm)
        ok, Lo, Uo = orc.ilu0(prp, pci, pv)
        assert ok
        for which, R in (("L", Lo), ("U", Uo)):
            rp, ci, v = il.factor(which)
            assert np.array_equal(rp, R.rowptr) and np.array_equal(ci, R.col)
            assert np.array_equal(v, R.val), f"{which} factor differs ({'device' if step else 'host'})"
        w = il.apply(z)
        w_ref = np.empty(s.n)
        w_ref[perm] = orc.ilu_apply(Lo, Uo, z[perm])
        say(name, s.n, "colours", ncol, "step", step, "apply rel", rel(w, w_ref))
        assert np.array_equal(w, w_ref)
        # into a device vector holding NaN: every entry a sweep reads was
        # written earlier in that sweep (explicit zeros of the pattern too)
        dz, dw = gpu.DeviceVector(s.n), gpu.DeviceVector(s.n)
        dz.upload(z)
        dw.fill(float("nan"))
        il.apply_device(dz.ptr, dw.ptr)
        assert np.array_equal(dw.download(), w_ref)
        dz.free()
        dw.free()
    il.close()


@pytest.mark.parametrize("name", ["damage129", "poisson65"])
def test_multicolour_pcg_and_gmres_match_reference(gpu, orc, name):
    s = _case(name)
    _, col = orc.colouring(s.rowptr, s.col, np.ones(len(s.col)))
    perm = np.argsort(col, kind="stable").astype(np.int32)
    prp, pci, pv = _permuted(s.rowptr, s.col, s.val, perm)
    il = gpu.ILU(s.rowptr, s.col, s.val, ordering="multicolour")
    r = orc.pcg_ilu(prp, pci, pv, s.b[perm], tol=1e-10, max_iter=3000)
    g = il.pcg(s.b, tol=1e-10, max_iter=3000)
    x_ref = np.empty(s.n)
    x_ref[perm] = r["x"]
    nat = gpu.ILU(s.rowptr, s.col, s.val).pcg(s.b, tol=1e-10, max_iter=3000)
    say(name, "pcg iters oracle", r["iters"], "gpu", g["iters"], "natural order", nat["iters"], "rel",
        rel(g["x"], x_ref))
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel(g["x"], x_ref) < 1e-12
    rg = orc.gmres_ilu(prp, pci, pv, s.b[perm], m=30, tol=1e-10, max_iter=3000)
    gg = il.gmres(s.b, m=30, tol=1e-10, max_iter=3000)
    xg = np.empty(s.n)
    xg[perm] = rg["x"]
    say(name, "gmres iters oracle", rg["iters"], "gpu", gg["iters"], "rel", rel(gg["x"], xg))
    assert gg["iters"] == rg["iters"]
    assert rel(gg["x"], xg) < 1e-9
    il.close()
