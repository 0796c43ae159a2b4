"""The LinA drop-in on the GPU against the compiled reference: the same
caller code (a C program compiled once against include/linalg and linked
against either library, and ctypes calls through the LinA names) gives the
same iteration counts and iterates."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from fem_amd import problems as P
from tests.dropin_bind import DROPIN_SO, LinA, ROOT, dptr

pytestmark = pytest.mark.gpu

from oracle.bind import REF_SO  # noqa: E402


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def libs(gpu, ref):
    return LinA(DROPIN_SO), LinA(ref.L)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_c_caller_links_against_either_library(tmp_path):
    obj = tmp_path / "example.o"
    inc = os.path.join(ROOT, "include", "linalg")
    subprocess.run(["gcc", "-std=c99", "-O2", "-c", "-I", inc, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "dropin_example.c"), "-o", str(obj)], check=True)
    out = {}
    for name, lib in (("dropin", DROPIN_SO), ("reference", REF_SO)):
        exe = tmp_path / f"example_{name}"
        d = os.path.dirname(lib)
        subprocess.run(["gcc", str(obj), "-o", str(exe), lib, f"-Wl,-rpath,{d}", "-lm", "-fopenmp"], check=True)
        r = subprocess.run([str(exe), "65"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        line = [l for l in r.stdout.splitlines() if l.startswith("iterations")][-1]
        out[name] = dict(zip(line.split()[0::2], map(float, line.split()[1::2])))
        say(name, out[name])
    assert abs(out["dropin"]["iterations"] - out["reference"]["iterations"]) <= 1
    assert abs(out["dropin"]["relres"] / out["reference"]["relres"] - 1) < 1e-6
    assert abs(out["dropin"]["residuum"] / out["reference"]["residuum"] - 1) < 1e-9
    assert abs(out["dropin"]["x0"] / out["reference"]["x0"] - 1) < 1e-12


def _solve_both(libs, build, solve):
    res = []
    for lib in libs:
        ctx = build(lib)
        res.append(solve(lib, ctx))
    return res


def test_pcg_amg_matches_reference(libs):
    s = P.degraded_elasticity(17)

    def run(lib):
        A = lib.from_csr(s.rowptr, s.col, s.val)
        lib.L.PCG_default(A, None, None)
        lib.L.PCG_set_AMG_preconditioner(A, 2, 4)
        x = np.zeros(s.n)
        b = s.b.copy()
        it = lib.L.pcg(dptr(x), dptr(b), s.n, 1e-10, 1000)
        return it, x, lib.L.get_pcg_residuum()

    (i1, x1, r1), (i2, x2, r2) = run(libs[0]), run(libs[1])
    say("pcg amg iters", i1, i2, "rel", rel(x1, x2))
    assert abs(i1 - i2) <= 1 and rel(x1, x2) < 1e-12 and abs(r1 / r2 - 1) < 1e-6


def test_pcg_ilu_matches_reference(libs):
    s = P.damage_system(33)[0]

    def run(lib):
        A = lib.from_csr(s.rowptr, s.col, s.val)
        L = lib.L.sparse_zero(s.n)
        U = lib.L.sparse_zero(s.n)
        lib.L.incomplete_LU_factorization(A, L, U)
        lib.L.PCG_default(A, L, U)
        x = np.zeros(s.n)
        b = s.b.copy()
        it = lib.L.pcg(dptr(x), dptr(b), s.n, 1e-10, 1000)
        return it, x

    (i1, x1), (i2, x2) = run(libs[0]), run(libs[1])
    say("pcg ilu iters", i1, i2, "rel", rel(x1, x2))
    assert abs(i1 - i2) <= 1 and rel(x1, x2) < 1e-12


def test_pcg_host_callback_preconditioner(libs):
    """a caller-defined preconditioner (Jacobi) registered with PCG_set_precon"""
    s = P.poisson(33)
    d = s.scipy().diagonal().copy()
    PREC = C.CFUNCTYPE(C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))

    @PREC
    def jacobi(n, r, z):
        for i in range(n):
            z[i] = r[i] / d[i]
        return 1

    def run(lib):
        A = lib.from_csr(s.rowptr, s.col, s.val)
        lib.L.PCG_default(A, None, None)
        lib.PCG_set_precon(C.cast(jacobi, C.c_void_p))
        x = np.zeros(s.n)
        b = s.b.copy()
        return lib.L.pcg(dptr(x), dptr(b), s.n, 1e-10, 2000), x

    (i1, x1), (i2, x2) = run(libs[0]), run(libs[1])
    say("pcg callback iters", i1, i2, "rel", rel(x1, x2))
    assert abs(i1 - i2) <= 1 and rel(x1, x2) < 1e-11


def test_amg_solve_full_multigrid_matches_reference(libs):
    """AMG_setup with FULL_AMG (Multigrid.h:38) and AMG_solve: one
    full-multigrid pass in both libraries, the same result to 1e-12."""
    s = P.degraded_elasticity(17)

    def run(lib):
        A = lib.from_csr(s.rowptr, s.col, s.val)
        lib.L.AMG_print_info(0)
        # the smoother and AMG_solve's controls are library globals: set them
        lib.L.AMG_set_smoother.argtypes = [C.c_void_p]
        lib.L.AMG_set_smoother(C.cast(lib.L.AMG_SOR_smoother, C.c_void_p))
        lib.L.AMG_set_SOR_relaxation_coeff(1.3)
        lib.L.AMG_set_eps(1e-2)
        lib.L.AMG_set_max_iter(100)
        I = lib.L.AMG_setup(A, 2, 4, 2, 3, 1, 2, 1, None, None)
        x = np.sin(np.arange(s.n) * 0.1)
        f = s.b.copy()
        it = lib.L.AMG_solve(I, dptr(f), dptr(x), None)
        lib.L.AMG_free_data(I)
        lib.free(A)
        return it, x

    (i1, x1), (i2, x2) = run(libs[0]), run(libs[1])
    say("full multigrid iters", i1, i2, "rel", rel(x1, x2))
    assert i1 == i2 == 1 and rel(x1, x2) < 1e-12


def test_bicgstab_amg_matches_reference(libs):
    s = P.poisson(33)

    def run(lib):
        A = lib.from_csr(s.rowptr, s.col, s.val)
        lib.L.RN_set_default_matrix(A)
        lib.BI_set_mat_mult(C.cast(lib.RN_default_mat_mult, C.c_void_p))
        lib.L.AMG_set_SOR_relaxation_coeff(1.0)
        lib.L.BI_set_AMG_preconditioner(A, 1, 4)
        lib.L.BI_print_info(0)
        lib.L.BI_set_eps(1e-13)
        lib.L.BI_set_max_iter(200)
        x = np.zeros(s.n)
        b = s.b.copy()
        return lib.L.BiCGStab_solve(dptr(x), dptr(b), s.n), x

    (i1, x1), (i2, x2) = run(libs[0]), run(libs[1])
    say("bicgstab iters", i1, i2, "rel", rel(x1, x2))
    assert abs(i1 - i2) <= 1 and rel(x1, x2) < 1e-10


def test_reflective_newton_matches_reference(libs):
    """set_RN_Hesse / set_RN_linear / set_RN_constraints / RN_set_params /
    RN_set_solver(RN_PCG) / reflective_newton through the LinA names"""
    s, dold, ones = P.damage_system(33, h_scale=100.0)
    n = s.n
    idx = np.arange(n, dtype=np.int32)

    def run(lib):
        L = lib.L
        H = lib.from_csr(s.rowptr, s.col, s.val)
        b = s.b.copy()
        lo, up = dold.copy(), ones.copy()
        L.RN_print_info(0)
        L.set_RN_Hesse.argtypes = [C.c_void_p]
        L.set_RN_Hesse(C.cast(H, C.c_void_p))
        L.set_RN_linear.argtypes = [C.POINTER(C.c_double)]
        L.set_RN_linear(dptr(b))
        L.set_RN_functional.argtypes = [C.c_void_p]
        L.set_RN_functional(None)
        L.RN_set_gradient.argtypes = [C.c_void_p]
        L.RN_set_gradient(None)
        L.set_RN_constraints.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        L.set_RN_constraints(idx.ctypes.data_as(C.POINTER(C.c_int)), dptr(lo), dptr(up), n)
        L.RN_set_params.argtypes = [C.c_int] + [C.c_double] * 10
        L.RN_set_params(50, 1e-8, 0.1, 0.9, 0.1, float(np.sqrt(n) / 2), 1.0, 0.6, 1.1, 1e-3, 1000.0)
        L.RN_set_local_solver.argtypes = [C.c_void_p]
        L.RN_set_local_solver(None)
        L.RN_set_solver(3)
        x = dold.copy()
        L.reflective_newton.argtypes = [C.POINTER(C.c_double)]
        it = L.reflective_newton(dptr(x))
        return it, x

    (i1, x1), (i2, x2) = run(libs[0]), run(libs[1])
    say("reflective newton iters", i1, i2, "rel", rel(x1, x2))
    assert i1 == i2 and rel(x1, x2) < 1e-9


def test_eigenvalue_searches_match_reference(libs):
    """power_method, get_most_negative_eigenvector and
    get_neg_curvature_direction (linear_algebra.c:4802-4865), which Crack
    calls directly (crack.c:8995, 10528, 10648): power_method from the same
    start vector in both libraries (iterations within the rounding of the
    stopping test, estimate and iterate to 1e-9); the two-stage searches
    (start vectors drawn from the time, as in the reference) reach scipy's
    smallest eigenvalue, of a positive definite and of an indefinite matrix,
    in both libraries; matrix_norm bit-identical."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import eigsh
    s = P.damage_system(17, h_scale=1.0)[0]
    A0 = s.scipy().tocsr()
    lmax = eigsh(A0, k=1, which="LA", return_eigenvectors=False)[0]
    lmin = eigsh(A0, k=1, which="SA", return_eigenvectors=False)[0]
    c = 0.5 * (lmin + eigsh(A0, k=2, which="SA", return_eigenvectors=False).max())
    Ai = (A0 - c * sp.identity(s.n, format="csr")).tocsr()
    Ai.sort_indices()
    v0 = np.random.default_rng(7).random(s.n)
    v0 /= np.linalg.norm(v0)
    out = []
    for lib in libs:
        M = lib.from_csr(A0.indptr.astype(np.int32), A0.indices.astype(np.int32), A0.data)
        Mi = lib.from_csr(Ai.indptr.astype(np.int32), Ai.indices.astype(np.int32), Ai.data)
        v = v0.copy()
        e = np.zeros(1)
        it = lib.L.power_method(M, dptr(e), dptr(v), 1000, 1e-8, None, 0.0)
        d = np.zeros(s.n)
        lo = np.zeros(1)
        lib.L.get_most_negative_eigenvector(M, dptr(lo), dptr(d), 100000, 1e-12)
        cu = np.zeros(1)
        lib.L.get_neg_curvature_direction(Mi, dptr(cu), dptr(d), 100000, 1e-12, 1e-10)
        out.append((it, e[0], v, lo[0], cu[0], lib.L.matrix_norm(M)))
        lib.free(M)
        lib.free(Mi)
    (i1, e1, v1, lo1, cu1, q1), (i2, e2, v2, lo2, cu2, q2) = out
    say("power iters", i1, i2, "eig", e1, e2, "lowest", lo1, lo2, lmin, "curvature", cu1, cu2, lmin - c)
    assert q1 == q2
    assert i1 > 0 and abs(i1 - i2) <= 1 and abs(e1 - e2) <= 1e-9 * abs(e2)
    if i1 == i2:
        assert np.linalg.norm(v1 - v2) <= 1e-9
    for lo in (lo1, lo2):
        assert abs(lo - lmin) <= 1e-6 * lmax
    for cu in (cu1, cu2):
        assert cu < 0 and abs(cu - (lmin - c)) <= 1e-6 * lmax


def test_power_method_with_empty_rows_matches_reference(libs):
    """A row list with empty rows (Len[i] == 0, as read_sparse produces from
    a file that holds such rows) is a valid operand of the reference's
    power_method: the rows act as zero rows.  The drop-in's device copy
    gives them the explicit (i, 0.0) entry, so the iteration count, the
    estimate and the iterate agree with the reference's."""
    s = P.damage_system(17, h_scale=1.0)[0]
    empty = [0, 5, s.n // 2, s.n - 1]
    v0 = np.random.default_rng(11).random(s.n)
    v0 /= np.linalg.norm(v0)
    out = []
    for lib in libs:
        M = lib.from_csr(s.rowptr, s.col, s.val)
        for i in empty:
            M.contents.Len[i] = 0     # rows emptied in place (their arrays stay allocated)
        v = v0.copy()
        e = np.zeros(1)
        it = lib.L.power_method(M, dptr(e), dptr(v), 1000, 1e-8, None, 0.0)
        for i in empty:
            M.contents.Len[i] = 1
        lib.free(M)
        out.append((it, e[0], v))
    (i1, e1, v1), (i2, e2, v2) = out
    say("power iters (empty rows)", i1, i2, "eig", e1, e2)
    assert i1 > 0 and abs(i1 - i2) <= 1 and abs(e1 - e2) <= 1e-9 * abs(e2)
    if i1 == i2:
        assert np.linalg.norm(v1 - v2) <= 1e-9
    assert all(abs(v1[i]) <= 1e-300 for i in empty)
