"""Pass-through levels of the V-cycle (lina_amg_params.skip_levels, and any
intermediate level whose sweep count is zero): such a level reads no matrix
of its own; its right-hand side goes down through R and the coarse
correction comes back through P.  With no smoothing anywhere, one V-cycle
from zero is therefore the composite transfer P_top ... P_1 A_0^-1 R_1 ...
R_top f, which numpy forms from the hierarchy's own P, R and coarsest A."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from fem_amd import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _csr(t, shape):
    rp, ci, v = t
    return sp.csr_matrix((v, ci, rp), shape=shape)


@pytest.mark.parametrize("kind,prec", [("poisson65", "fp64"), ("elasticity33", "fp64"), ("elasticity33", "mixed")])
def test_unsmoothed_cycle_is_the_composite_transfer(gpu, kind, prec):
    s = P.poisson(65) if kind == "poisson65" else P.degraded_elasticity(33)
    dof = 1 if kind == "poisson65" else 2
    depth = 4
    prm = gpu.params("vcycle", dof, depth, precision=prec, smoothing_iterations=0, aggr_smoothing_iterations=0,
                     coarse_sweeps=0)
    amg = gpu.AMG(s.rowptr, s.col, s.val, prm)
    f = np.cos(np.arange(s.n) * 0.21) + 0.5
    k, x = amg.amg_solve(f, np.zeros(s.n))
    say(kind, prec, "V-cycles", k)
    assert k == 1
    lv = [amg.level(l) for l in range(depth)]
    y = f
    for l in range(depth - 1, 0, -1):
        y = _csr(lv[l]["R"], (lv[l]["coarse"], lv[l]["fine"])) @ y
    A0 = _csr(lv[0]["A"], (lv[0]["fine"], lv[0]["fine"])).toarray()
    e = np.linalg.solve(A0, y)
    for l in range(1, depth):
        e = _csr(lv[l]["P"], (lv[l]["fine"], lv[l]["coarse"])) @ e
    err = np.linalg.norm(x - e) / np.linalg.norm(e)
    say("  vs composite", err)
    assert err < (1e-10 if prec == "fp64" else 1e-5)
    amg.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_skipped_levels_keep_pcg_accurate(gpu, prec):
    """pcg with levels 1 and 3 passed through reaches relres 1e-10, and its
    solution agrees with a direct solve as the fully smoothed cycle's does"""
    s = P.degraded_elasticity(65)
    A = s.scipy().tocsc()
    xd = spla.spsolve(A, s.b)
    out = {}
    for skip in (0, 0b1010):
        amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("vcycle", 2, 6, precision=prec, skip_levels=skip))
        r = amg.pcg(s.b, tol=1e-10, max_iter=3000, stop="relres")
        err = np.linalg.norm(r["x"] - xd) / np.linalg.norm(xd)
        relres = np.linalg.norm(s.b - A @ r["x"]) / np.linalg.norm(s.b)
        say(prec, "skip", bin(skip), "iters", r["iters"], "relres", relres, "err vs direct", err)
        assert relres < 2e-10 and err < 1e-7
        out[skip] = r["iters"]
        amg.close()
    assert out[0b1010] < 3 * out[0]
