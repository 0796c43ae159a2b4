"""GPU reflective Newton against the compiled reference (oracle/_ref):
reflective_newton on the damage subproblem q(d) = d.A.d/2 + d.b with
d_old <= d <= 1, for the three inner solvers the reference offers without
UMFPACK (ILU(0)-PCG, BiCGStab + AMG, AMG_solve).  The driving force is
scaled so that the solution is interior (h_scale 100), pressed towards the
upper bound (1000) or left on the lower bound (1).  Newton iteration counts
must agree exactly and the minimisers to 1e-9 relative; independently of the
reference, the minimum value must not exceed the one L-BFGS-B finds (the
reflective iterates approach active bounds only asymptotically, so a
projected-gradient test is printed, not asserted)."""
import numpy as np
import pytest

from fem_amd import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def kkt(s, x, lo, up):
    g = s.scipy() @ x + s.b
    pg = np.where(x <= lo + 1e-9, np.minimum(g, 0), np.where(x >= up - 1e-9, np.maximum(g, 0), g))
    return float(np.linalg.norm(pg) / np.linalg.norm(g))


def lbfgsb_min(s, lo, up):
    """min q over the box by scipy's L-BFGS-B, independent of both RN codes"""
    from scipy.optimize import minimize
    A = s.scipy()
    f = lambda x: (0.5 * x @ (A @ x) + s.b @ x, A @ x + s.b)
    r = minimize(f, lo.copy(), jac=True, method="L-BFGS-B", bounds=list(zip(lo, up)),
                 options={"maxiter": 20000, "ftol": 1e-15, "gtol": 1e-12})
    return float(r.fun)


@pytest.mark.parametrize("solver", ["pcg", "bicgstab", "amg"])
@pytest.mark.parametrize("h_scale", [1.0, 100.0, 1000.0])
def test_reflective_newton_matches_reference(gpu, ref, solver, h_scale):
    s, dold, ones = P.damage_system(33, h_scale=h_scale)
    say("rn", solver, h_scale)
    delta = np.sqrt(s.n) / 2
    H = ref.matrix(s.rowptr, s.col, s.val)
    code = {"bicgstab": 1, "amg": 2, "pcg": 3}[solver]
    r = ref.reflective_newton(H, s.b, dold, ones, dold.copy(), solver=code, dof=1, depth=4, max_iter=50, eps=1e-8,
                              delta=delta)
    ref.free(H)
    prm = gpu.rn_params(solver, 1, 4, delta=delta)
    rn = gpu.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, dold, ones, prm)
    g = rn.solve(dold)
    rn.close()
    rel = np.linalg.norm(g["x"] - r["x"]) / np.linalg.norm(r["x"])
    say("  iters ref", r["iters"], "gpu", g["iters"], "rel", rel, "q", g["q"], r["q"], "kkt", kkt(s, g["x"], dold, ones))
    assert g["iters"] == r["iters"]
    assert rel < 1e-9
    assert abs(g["q"] - r["q"]) <= 1e-10 * abs(r["q"])
    assert np.all(g["x"] >= dold) and np.all(g["x"] <= 1.0)
    # the minimum value itself, from an unrelated method
    q_ref = lbfgsb_min(s, dold, ones)
    assert g["q"] <= q_ref + 1e-7 * abs(q_ref)


def test_reflective_newton_partial_bounds(gpu, ref):
    """Bounds on a subset of the indices and one-sided bounds (NaN)."""
    s, dold, ones = P.damage_system(33, h_scale=300.0)
    n = s.n
    rng = np.random.default_rng(5)
    lo = dold.copy()
    up = ones.copy()
    up[rng.random(n) < 0.3] = np.nan
    x0 = dold.copy()
    H = ref.matrix(s.rowptr, s.col, s.val)
    r = ref.reflective_newton(H, s.b, lo, up, x0.copy(), solver=3, dof=1, depth=4, max_iter=50, eps=1e-8,
                              delta=np.sqrt(n) / 2)
    ref.free(H)
    prm = gpu.rn_params("pcg", 1, 4, delta=np.sqrt(n) / 2)
    g = gpu.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, lo, up, prm).solve(x0)
    rel = np.linalg.norm(g["x"] - r["x"]) / np.linalg.norm(r["x"])
    say("  iters ref", r["iters"], "gpu", g["iters"], "rel", rel)
    assert g["iters"] == r["iters"]
    assert rel < 1e-9


@pytest.mark.parametrize("order", ["natural", "multicolour"])
@pytest.mark.parametrize("h_scale", [100.0, 1000.0])
def test_reflective_newton_gmres_local_solver(gpu, ref, h_scale, order):
    """GMRES(30)+ILU(0) in the local-solver slot, the ILU in either order: no
    reference counterpart, so the minimiser is compared with the reference's
    ILU-PCG variant and with L-BFGS-B."""
    s, dold, ones = P.damage_system(33, h_scale=h_scale)
    delta = np.sqrt(s.n) / 2
    H = ref.matrix(s.rowptr, s.col, s.val)
    r = ref.reflective_newton(H, s.b, dold, ones, dold.copy(), solver=3, dof=1, depth=4, max_iter=50, eps=1e-8,
                              delta=delta)
    ref.free(H)
    g = gpu.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, dold, ones,
                             gpu.rn_params("gmres", 1, 4, delta=delta, ilu_ordering=order)).solve(dold)
    rel = np.linalg.norm(g["x"] - r["x"]) / np.linalg.norm(r["x"])
    say("  gmres", order, "iters", g["iters"], "ref pcg", r["iters"], "rel", rel, "inner", g["inner_iterations"])
    assert rel < 1e-6
    q_ref = lbfgsb_min(s, dold, ones)
    assert g["q"] <= q_ref + 1e-7 * abs(q_ref)


@pytest.mark.parametrize("case", ["damage", "shifted"])
def test_power_method_matches_reference(gpu, ref, case):
    """power_method (linear_algebra.c:4802-4822), the eigenvalue search of
    the negative-curvature branch, on the device (lina_matrix_power, the
    routine the reflective Newton calls) against the compiled reference's
    power_method from the same start vector (the reference draws it with
    srand(time(NULL)); here it is fixed): the same iteration count (up to
    the rounding of the stopping test) and estimate, the same iterate.  "shifted" is get_neg_curvature_direction's second matrix,
    A - s I with s the largest eigenvalue, whose dominant eigenvalue plus s
    is A's smallest.  Run to a tight tolerance, the estimates agree with
    scipy's eigenvalues."""
    import ctypes as C
    import scipy.sparse as sp
    from scipy.sparse.linalg import eigsh
    from tests.dropin_bind import LinA, dptr
    s = P.damage_system(17, h_scale=1.0)[0]
    A = s.scipy().tocsr()
    lam = eigsh(A, k=1, which="LA", return_eigenvectors=False)[0]
    lmin = eigsh(A, k=1, which="SA", return_eigenvectors=False)[0]
    if case == "shifted":
        A = (A - lam * sp.identity(s.n, format="csr")).tocsr()
        A.sort_indices()
    rng = np.random.default_rng(3)
    v0 = rng.random(s.n)
    v0 /= np.linalg.norm(v0)
    R = LinA(ref.L)
    R.L.power_method.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double,
                                 C.c_void_p, C.c_double]
    R.L.power_method.restype = C.c_int
    M = R.from_csr(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    dM = gpu.Matrix(A.indptr, A.indices, A.data)
    for eps, max_iter in ((1e-6, 1000), (1e-13, 20000)):
        vr = v0.copy()
        er = C.c_double(0.0)
        it_r = R.L.power_method(M, C.byref(er), dptr(vr), max_iter, eps, None, 0.0)
        it, e, v = dM.power(v0, max_iter=max_iter, eps=eps)
        say(case, "eps", eps, "iters ref", it_r, "gpu", it, "eig", er.value, e)
        # near a tight threshold the estimate changes by ~1e-16 per iteration:
        # the stopping iteration may move by a few among 18k
        assert it > 0 and abs(it - it_r) <= max(1, it_r // 2000)
        assert abs(e - er.value) <= 1e-9 * abs(er.value)
        if it == it_r:
            assert np.linalg.norm(v - vr) <= 1e-9
    if case == "damage":
        assert abs(e - lam) <= 1e-6 * abs(lam)
    else:
        assert abs(e + lam - lmin) <= 1e-6 * abs(lam)
    dM.close()
    R.free(M)


@pytest.mark.parametrize("solver,order", [("pcg", "natural"), ("gmres", "natural")])
def test_device_scaled_hessian_matches_host_assembly(gpu, capfd, monkeypatch, solver, order):
    """The scaled Hessian M = D H D + diag(|g| J) assembled on the device
    (while its pattern is the last host assembly's) against the host
    assembly on every step (LINA_RN_HOST_M=1): the same Newton and inner
    iteration counts and the same minimiser, bit for bit.  (The multicolour
    ILU factors M on the fixed superset pattern from the device values on
    every step; test_multicolour_reflective_newton_is_the_reference_on_the_permuted_system
    and test_damage1025_reflective_newton_kkt cover it.)"""
    s, dold, ones = P.damage_system(129, h_scale=100.0)
    monkeypatch.setenv("LINA_RN_TIMING", "1")
    out = {}
    for host in (True, False):
        monkeypatch.setenv("LINA_RN_HOST_M", "1" if host else "0")
        prm = gpu.rn_params(solver, 1, 6, delta=float(np.sqrt(s.n) / 2), ilu_ordering=order)
        rn = gpu.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, dold, ones, prm)
        capfd.readouterr()
        out[host] = rn.solve(dold)
        err = capfd.readouterr().err
        rn.close()
        steps = [l for l in err.splitlines() if l.startswith("[rn] step")]
        dev_steps = sum("device" in l for l in steps)
        say(solver, order, "host M" if host else "device M", "newton", out[host]["iters"], "inner",
            out[host]["inner_iterations"], "device-assembled steps", dev_steps, "of", len(steps))
        assert len(steps) == out[host]["iters"]
        if host:
            assert dev_steps == 0
        else:
            assert dev_steps >= len(steps) // 2
    assert out[True]["iters"] == out[False]["iters"]
    assert out[True]["inner_iterations"] == out[False]["inner_iterations"]
    assert np.array_equal(out[True]["x"], out[False]["x"])


@pytest.mark.parametrize("h_scale", [100.0, 1000.0])
def test_multicolour_reflective_newton_is_the_reference_on_the_permuted_system(gpu, ref, h_scale):
    """ILU-PCG in multicolour order inside the reflective Newton is the
    reference's reflective_newton (ILU-PCG) run on P H P^T, P b and the
    permuted bounds, P the stable sort of the greedy colouring of H's
    pattern: the scaled Hessian keeps H's pattern while the iterates stay
    interior, so its ILU(0) in that order is the reference's ILU of the
    permuted scaled Hessian.  Same Newton iteration count, the minimiser
    mapped back to 1e-9 relative, the same minimum."""
    from oracle.bind import Oracle
    from tests.test_gpu_ilu_multicolour import _permuted
    s, dold, ones = P.damage_system(33, h_scale=h_scale)
    _, col = Oracle().colouring(s.rowptr, s.col, np.ones(len(s.col)))
    perm = np.argsort(col, kind="stable").astype(np.int32)
    prp, pci, pv = _permuted(s.rowptr, s.col, s.val, perm)
    delta = np.sqrt(s.n) / 2
    H = ref.matrix(prp, pci, pv)
    r = ref.reflective_newton(H, s.b[perm], dold[perm], ones[perm], dold[perm].copy(), solver=3, dof=1, depth=4,
                              max_iter=50, eps=1e-8, delta=delta)
    ref.free(H)
    x_ref = np.empty(s.n)
    x_ref[perm] = r["x"]
    prm = gpu.rn_params("pcg", 1, 4, delta=delta, ilu_ordering="multicolour")
    rn = gpu.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, dold, ones, prm)
    g = rn.solve(dold)
    rn.close()
    rel = np.linalg.norm(g["x"] - x_ref) / np.linalg.norm(x_ref)
    say("rn multicolour", h_scale, "iters ref", r["iters"], "gpu", g["iters"], "rel", rel, "q", g["q"], r["q"])
    assert g["iters"] == r["iters"]
    assert rel < 1e-9
    assert abs(g["q"] - r["q"]) <= 1e-10 * abs(r["q"])
    assert np.all(g["x"] >= dold) and np.all(g["x"] <= 1.0)
