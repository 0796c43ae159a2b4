"""The five workloads at their full sizes, on the GPU, checked through what
does not depend on size:

* 257^2 Poisson (66,049 DOF), one V(1,1) per CG iteration: iterates and
  iteration count against the oracle's pcg (the C restatement finishes in
  seconds at this size);
* 1025^2 degraded elasticity (2.1M DOF): the true residual of the FP64 and
  the mixed-precision solves, their error against SuperLU's FP64 direct
  solution (committed sketch), the mixed solve's iteration count beside the
  FP64 one;
* 1025^2 damage subproblem (1.05M DOF), with each named inner solver
  (ILU(0)-PCG, GMRES(30)+ILU(0), BiCGStab + FP32 AMG): the reflective
  Newton's result lies in the box, its q is the quadratic's value there, and
  no projected-gradient step from it lowers q;
* 512 batched specimens (10.8M DOF): every specimen's true residual, the
  error of every 64th against a direct solve;
* 513^2 adjoint (526k DOF): the gradient against central differences of the
  loss on three directions.
"""
import numpy as np
import pytest

from fem_amd import problems as P
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("smoother", ["ssor", "mc_ssor"])
def test_poisson257_one_vcycle_matches_oracle(gpu, smoother):
    """One symmetric V(1,1) from a zero guess per CG iteration: lexicographic
    SOR forward before and backward after the coarse correction (ssor), or
    the multicolour form (mc_ssor); omega 1.  (A forward-only SOR V-cycle is
    not symmetric: CG then stalls, in the oracle as on the GPU.)"""
    s = P.poisson(257)
    depth = 5
    say("poisson257", s.n, "nnz", s.nnz, smoother)
    r = Oracle().pcg_amg(s.rowptr, s.col, s.val, s.b, dof=1, depth=depth, smiter=1, aggr_smiter=1,
                         smoother=smoother, alpha=1.0, amg_eps=1e-2, amg_max_iter=1, zero_init=True, tol=1e-10,
                         max_iter=500)
    prm = gpu.params("reference", 1, depth, smoothing_iterations=1, aggr_smoothing_iterations=1, amg_max_iter=1,
                     zero_init=1, smoother=smoother, omega=1.0)
    amg = gpu.AMG(s.rowptr, s.col, s.val, prm)
    g = amg.pcg(s.b, tol=1e-10, max_iter=500)
    amg.close()
    say("  iters oracle", r["iters"], "gpu", g["iters"], "rel", rel(g["x"], r["x"]))
    assert r["iters"] < 100
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel(g["x"], r["x"]) < 1e-12
    A = s.scipy()
    assert np.linalg.norm(s.b - A @ g["x"]) <= 1e-8 * np.linalg.norm(s.b)


def test_elasticity1025_mixed_and_fp64(gpu):
    """Both precisions to relative residual 1e-10 (at 1e-8 the error of this
    ill-conditioned system, element stiffness down to 1e-6, is still
    ~5e-7), each within 1e-8 of SuperLU's FP64 direct solution (its
    committed sketch: 64 Gaussian projections and 20,000 sampled entries);
    the same for the mixed solve whose cycle passes levels 1, 2 and 4
    through unsmoothed."""
    from tests import direct_oracle as D
    s = P.degraded_elasticity(1025)
    say("elasticity1025", s.n, "nnz", s.nnz)
    prm64 = gpu.params("vcycle", 2, 8, autotune=1)
    prm32 = gpu.params("vcycle", 2, 8, precision="mixed", autotune=1)
    prm32s = gpu.params("vcycle", 2, 8, precision="mixed", autotune=1, skip_levels=0b010110)
    amg = gpu.AMG(s.rowptr, s.col, s.val, prm64)
    A = s.scipy()
    nb = np.linalg.norm(s.b)
    out = {}
    names = ("fp64", "mixed", "mixed_skip")
    for name, prm in zip(names, (prm64, prm32, prm32s)):
        amg.set_params(prm)
        out[name] = amg.pcg(s.b, tol=1e-10, max_iter=20000, stop="relres")
        rr = np.linalg.norm(s.b - A @ out[name]["x"]) / nb
        say("  ", name, "iters", out[name]["iters"], "true relres", rr)
        assert rr <= 1.01e-10
    amg.close()
    errs, info = D.errors(s, [out[k]["x"] for k in names])
    for name, e in zip(names, errs):
        say("  ", name, "error vs direct solve: sketch", e["sketch"], "sampled", e["sample"], "max", e["sample_max"])
        assert e["sketch"] <= 1e-8 and e["sample"] <= 1e-8
    assert out["mixed"]["iters"] <= int(1.1 * out["fp64"]["iters"]) + 2


@pytest.mark.parametrize("solver,prec,order", [("pcg", "fp64", "natural"), ("gmres", "fp64", "natural"),
                                              ("bicgstab", "mixed", "natural"), ("pcg", "fp64", "multicolour"),
                                              ("gmres", "fp64", "multicolour")])
def test_damage1025_reflective_newton_kkt(gpu, solver, prec, order):
    """The damage subproblem's reflective Newton with each inner solver:
    ILU(0)-PCG; GMRES(30) + ILU(0) (FEM.c:737-744); BiCGStab + AMG
    (reflective_newton.c:205-217) with the FP32 hierarchy (multicolour SOR,
    the reference's AMG_solve settings); ILU(0)-PCG and GMRES(30) with the
    ILU(0) in multicolour order --
    the scaling, reflections and active-set tests stay FP64 in every
    variant."""
    s, dold, ones = P.damage_system(1025, h_scale=1.0)
    say("damage1025", s.n, "nnz", s.nnz, "inner", solver, prec, order)
    amg = {"precision": 1 if prec == "mixed" else 0}
    if solver == "bicgstab":
        amg["smoother"] = 2  # MC-SOR
    prm = gpu.rn_params(solver, 1, 6, delta=float(np.sqrt(s.n) / 2), amg=amg, ilu_ordering=order)
    if solver == "gmres":
        assert prm.gmres_m == 30
    rn = gpu.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, dold, ones, prm)
    g = rn.solve(dold)
    rn.close()
    x = g["x"]
    H = s.scipy()
    grad = H @ x + s.b
    say("  newton", g["iters"], "inner", g["inner_iterations"], "dq", g["dq"], "q", g["q"])
    assert g["dq"] <= 1e-8
    assert np.all(x >= dold - 1e-12) and np.all(x <= 1.0 + 1e-12)
    q = 0.5 * x @ (H @ x) + s.b @ x
    assert abs(q - g["q"]) <= 1e-10 * abs(q)
    # no projected-gradient step x(t) = clip(x - t g) lowers q (beyond the
    # Newton iteration's stopping tolerance)
    scale = np.abs(grad).max()
    for t in (1e-3, 1e-2, 1e-1, 1.0, 10.0):
        xt = np.clip(x - (t / scale) * grad * 1e-3, dold, 1.0)
        qt = 0.5 * xt @ (H @ xt) + s.b @ xt
        say("  projected step", t, "dq", (qt - q) / abs(q))
        assert qt - q >= -1e-8 * abs(q)
    # complementarity (Coleman-Li measure): the distance to the bound the
    # gradient points to, times the gradient
    v = np.where(grad > 0, x - dold, 1.0 - x)
    say("  max |v g| / max |g|", float(np.abs(v * grad).max() / scale))


def test_batch512_every_specimen_converges(gpu):
    from scipy.sparse.linalg import splu
    sp = P.batch_specimens(512)
    row_off, rp, ci, v, b = P.block_diagonal(sp)
    say("batch512", row_off[-1], "nnz", len(ci))
    bt = gpu.Batch(row_off, rp, ci, v, gpu.params("vcycle", 2, 5, precision="mixed", autotune=1))
    out = bt.pcg(b, tol=1e-8, max_iter=3000, graph_iters=10)
    say("  iters", int(out["iters"].min()), "-", int(out["iters"].max()), "active", out["active"])
    assert out["active"] == 0
    worst = 0.0
    for k, s in enumerate(sp):
        x = out["x"][row_off[k]:row_off[k + 1]]
        A = s.scipy()
        rr = np.linalg.norm(s.b - A @ x) / np.linalg.norm(s.b)
        worst = max(worst, rr)
        if k % 64 == 0:
            xd = splu(A.tocsc()).solve(s.b)
            assert rel(x, xd) <= 1e-5, k
    say("  worst true relres", worst)
    assert worst <= 1.01e-8


def test_adjoint513_gradient_against_finite_differences(gpu):
    from tests import adjoint_ref as R
    case = P.adjoint_case(513)
    s, dofs = case.system(case.theta)
    say("adjoint513", s.n, "elements", len(case.theta))
    amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("vcycle", 2, 7, precision="mixed"))

    so, _ = case.system(case.theta_obs)
    op_obs = gpu.Matrix(so.rowptr, so.col, so.val)
    u_obs = amg.pcg_operator(op_obs, so.b, tol=1e-12, max_iter=20000)["x"]
    op_obs.close()
    u = amg.pcg(s.b, tol=1e-12, max_iter=20000, stop="relres")["x"]
    out = amg.elasticity_adjoint(case.mesh.xy, case.mesh.tri, case.nu, u, case.dloss(u, u_obs, dofs), tol=1e-12,
                                 max_iter=20000)
    h = 1e-6
    for k, vdir in enumerate(R.directions(len(case.theta), 3)):
        r0, rv = R.fd_pieces(case, u, vdir)
        d = []
        for sign in (1, -1):
            sy, _ = case.system(case.theta + sign * h * vdir)
            op = gpu.Matrix(sy.rowptr, sy.col, sy.val)
            d.append(amg.pcg_operator(op, r0 + sign * h * rv, tol=1e-12, max_iter=20000)["x"])
            op.close()
        fd = R.fd_quotient(case, d[0], d[1], u, u_obs, h)
        gv = out["grad"] @ vdir
        say("  direction", k, "fd", fd, "adjoint", gv, "rel", abs(fd - gv) / abs(gv))
        assert abs(fd - gv) <= 1e-6 * abs(gv)
    amg.close()
