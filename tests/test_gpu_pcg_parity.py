"""GPU parity of the FP64 AMG-PCG path against the reference LinA.

The oracle (oracle/liboracle.so, the C restatement pinned bit for bit to the
reference compiled from LinA's own sources by tests/test_oracle_cpu.py and
tests/test_oracle_krylov_cpu.py) and the HIP path solve the identical CSR
system; iteration counts must agree to +-1 and iterates to 1e-12 relative,
the hierarchy being bit-identical (tests/test_setup_cpu.py).
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


@pytest.mark.parametrize("m", [17, 33])
def test_spmv_matches_sequential_rows(gpu, m):
    s = P.poisson(m)
    say("spmv poisson", m)
    amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("reference", 1, 3))
    x = np.sin(np.arange(s.n) * 0.37)
    y = amg.spmv(x)
    # sequential row sums in column order: exactly the reference's PCG_mult
    y_ref = np.zeros(s.n)
    for i in range(s.n):
        acc = 0.0
        for k in range(s.rowptr[i], s.rowptr[i + 1]):
            acc += s.val[k] * x[s.col[k]]
        y_ref[i] = acc
    assert np.array_equal(y, y_ref)
    yf = amg.spmv(x, precision="mixed")
    assert rel(yf, y_ref) < 1e-6


def _ragged(n, lens, seed):
    """CSR with the given row lengths: diagonal plus distinct random columns."""
    rng = np.random.default_rng(seed)
    rp = [0]
    cols, vals = [], []
    for i, L in enumerate(lens):
        L = int(min(max(L, 1), n))
        c = set(rng.choice(n, size=L - 1, replace=False).tolist()) - {i}
        c = sorted(c | {i})
        cols += c
        vals += rng.uniform(-1, 1, len(c)).tolist()
        rp.append(len(cols))
    return np.array(rp, np.int32), np.array(cols, np.int32), np.array(vals)


def _clusters(n, seed):
    """CSR whose row blocks see one, two or three column clusters: row i
    couples to i, i+1, i+n/2 and i+n/2+7 (mod n), and every 997th row also to
    a far column, so the 16-bit column codes take the single-base, the
    two-base and the 32-bit fallback path."""
    rng = np.random.default_rng(seed)
    rp = [0]
    cols = []
    for i in range(n):
        c = {i, (i + 1) % n, (i + n // 2) % n, (i + n // 2 + 7) % n}
        if i % 997 == 0:
            c.add((i * 7919 + n // 4) % n)
        cols += sorted(c)
        rp.append(len(cols))
    return np.array(rp, np.int32), np.array(cols, np.int32), rng.uniform(-1, 1, len(cols))


@pytest.mark.parametrize("case", ["elast65", "elast257", "clusters", "dense_rows", "long_rows", "one_row"])
def test_matrix_handle_spmv(gpu, case):
    """lina_matrix: bit-identical to sequential row sums where one lane sums a
    row (mean row length <= 24); otherwise within the componentwise bound of
    a reordered sum, |y - y_seq| <= 64 eps sum_j |a_ij x_j|.  Covers blocks
    of many short rows, 16/32-lane rows, rows longer than one LDS block (the
    workgroup-wide path), a 1x1 matrix, and column layouts that take each
    path of the 16-bit column codes (one base, two bases, 32-bit fallback:
    elast257's blocked [ux; uy] ordering and the synthetic clusters)."""
    say("matrix spmv", case)
    rng = np.random.default_rng(7)
    if case in ("elast65", "elast257"):
        s = P.degraded_elasticity(int(case[5:]))
        rp, ci, v = s.rowptr, s.col, s.val
    elif case == "clusters":
        rp, ci, v = _clusters(120000, 3)
    elif case == "dense_rows":
        rp, ci, v = _ragged(3000, rng.integers(60, 400, 3000), 1)
    elif case == "long_rows":
        lens = rng.integers(1, 12, 4000)
        lens[[5, 100, 2047, 3999]] = [2500, 3999, 2049, 3000]
        rp, ci, v = _ragged(4000, lens, 2)
    else:
        rp, ci, v = np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.5])
    n = len(rp) - 1
    x = rng.uniform(-1, 1, n)
    m = gpu.Matrix(rp, ci, v)
    y = m.spmv(x)
    y_seq = Oracle().spmv(rp, ci, v, x)
    absum = Oracle().spmv(rp, ci, np.abs(v), np.abs(x))
    if (rp[-1] / n) <= 24 and np.diff(rp).max() <= 2048:
        assert np.array_equal(y, y_seq)
    else:
        assert np.all(np.abs(y - y_seq) <= 64 * np.finfo(float).eps * absum)
    yf = m.spmv(x, precision="mixed")
    assert np.all(np.abs(yf - y_seq) <= 64 * np.finfo(np.float32).eps * absum + 1e-30)
    m.close()


def _ref_solve(s, depth, **kw):
    """pcg of the oracle (pinned bit for bit to the compiled reference)"""
    return Oracle().pcg_amg(s.rowptr, s.col, s.val, s.b, dof=s.dof, depth=depth, **kw)


@pytest.mark.parametrize("m", [17, 33, 65])
def test_pcg_reference_preconditioner_poisson(gpu, m):
    """PCG_set_AMG_preconditioner: lexicographic SOR(1.6), AMG solve to 1e-2."""
    s = P.poisson(m)
    depth = 4
    say("pcg reference-mode poisson", m)
    r = _ref_solve(s, depth, smoother="sor", alpha=1.6, smiter=2, aggr_smiter=3, amg_eps=1e-2,
                   amg_max_iter=100, tol=1e-10, max_iter=500)
    amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("reference", 1, depth))
    g = amg.pcg(s.b, tol=1e-10, max_iter=500)
    say("  iters oracle", r["iters"], "gpu", g["iters"], "vcycles oracle", r["vcycles"], "gpu", g["vcycles"])
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel(g["x"], r["x"]) < 1e-12
    assert np.array_equal(g["b_out"], r["b_out"]) or rel(g["b_out"], r["b_out"]) < 1e-15
    assert abs(g["residual"] - r["res"]) <= 1e-12 * max(1.0, abs(r["res"])) or \
        abs(g["residual"] / r["res"] - 1) < 1e-6


@pytest.mark.parametrize("kind,m", [("poisson", 33), ("elasticity", 17), ("elasticity", 33)])
def test_pcg_vcycle_mc_ssor(gpu, kind, m):
    """One symmetric multicolour-SOR V(1,1) per iteration; the reference runs
    the same cycle with the smoother plugged into AMG_set_smoother."""
    s = P.poisson(m) if kind == "poisson" else P.degraded_elasticity(m)
    depth = 4
    say("pcg vcycle-mode", kind, m)
    r = _ref_solve(s, depth, smoother="mc_ssor", alpha=1.0, smiter=1, aggr_smiter=1, amg_eps=1e-2,
                   amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=2000)
    amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("vcycle", s.dof, depth, coarse_smoother=-1))
    g = amg.pcg(s.b, tol=1e-8, max_iter=2000)
    say("  iters oracle", r["iters"], "gpu", g["iters"], "rel", rel(g["x"], r["x"]))
    assert abs(g["iters"] - r["iters"]) <= 1
    assert rel(g["x"], r["x"]) < 1e-12


@pytest.mark.parametrize("kind,m", [("poisson", 65), ("elasticity", 33), ("elasticity", 65)])
def test_pcg_vcycle_default(gpu, kind, m):
    """The default V-cycle: MC-SSOR on the finest level, one l1-Jacobi sweep
    below (multi-lane rows on the dense coarse levels)."""
    s = P.poisson(m) if kind == "poisson" else P.degraded_elasticity(m)
    depth = 5
    say("pcg vcycle default", kind, m)
    r = _ref_solve(s, depth, smoother="hybrid", alpha=1.0, smiter=1, aggr_smiter=1, amg_eps=1e-2,
                   amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=3000, coarse_sweeps=1)
    for prec in ("fp64", "mixed"):
        amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("vcycle", s.dof, depth, precision=prec))
        g = amg.pcg(s.b, tol=1e-8, max_iter=3000)
        say("  ", prec, "iters oracle", r["iters"], "gpu", g["iters"], "rel", rel(g["x"], r["x"]))
        if prec == "fp64":
            assert abs(g["iters"] - r["iters"]) <= 1
            assert rel(g["x"], r["x"]) < 1e-12
        else:
            # the FP32 hierarchy is a different (inexact) preconditioner: its
            # iteration count may grow a little; the FP64 outer loop keeps
            # the solution accuracy
            A = s.scipy()
            assert g["iters"] <= int(1.25 * r["iters"]) + 2
            assert np.linalg.norm(s.b - A @ g["x"]) / np.linalg.norm(s.b) < 1e-6


@pytest.mark.parametrize("case,sm,alpha", [("poisson33", "sor", 1.2), ("poisson33", "mc_ssor", 1.0),
                                            ("elast25", "ssor", 1.0), ("elast25", "mc_ssor", 1.0)])
def test_pcg_full_multigrid_preconditioner(gpu, case, sm, alpha):
    """pcg preconditioned by AMG_solve of a FULL_AMG hierarchy (one
    full-multigrid pass per preconditioner call), against the oracle's pcg
    with the same preconditioner (pinned bit for bit to the compiled
    reference): FP64 iterations within 1 and iterates to 1e-12; the mixed
    hierarchy reaches the same true residual."""
    s, depth = {"poisson33": (P.poisson(33), 4), "elast25": (P.degraded_elasticity(25), 4)}[case]
    say("pcg full-multigrid preconditioner", case, sm)
    r = _ref_solve(s, depth, smoother=sm, alpha=alpha, tol=1e-10, max_iter=300, cycle="full")
    A = s.scipy()
    for prec in ("fp64", "mixed"):
        amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("reference", s.dof, depth, smoother=sm, omega=alpha,
                                                          cycle="full", precision=prec))
        g = amg.pcg(s.b, tol=1e-10, max_iter=300)
        amg.close()
        say("  ", prec, "iters oracle", r["iters"], "gpu", g["iters"], "rel", rel(g["x"], r["x"]))
        if prec == "fp64":
            assert abs(g["iters"] - r["iters"]) <= 1
            assert rel(g["x"], r["x"]) < 1e-12
        else:
            assert g["iters"] <= int(1.25 * r["iters"]) + 2
            assert np.linalg.norm(s.b - A @ g["x"]) <= 1e-8 * np.linalg.norm(s.b)


@pytest.mark.parametrize("smoother", ["mc_sor", "l1_jacobi", "ssor"])
def test_amg_solve_other_smoothers(gpu, smoother):
    s = P.poisson(33)
    depth = 4
    say("amg_solve", smoother)
    orc = Oracle()
    I = orc.amg_setup(s.rowptr, s.col, s.val, dof=1, depth=depth)
    x0 = np.zeros(s.n)
    k_ref, x_ref = orc.amg_solve(I, s.b, x0, smoother=smoother, alpha=1.2, eps=1e-6, max_iter=30)
    orc.amg_free(I)
    amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("reference", 1, depth, smoother=smoother, omega=1.2,
                                                      amg_eps=1e-6, amg_max_iter=30))
    k, x = amg.amg_solve(s.b, x0)
    say("  vcycles oracle", k_ref, "gpu", k)
    assert k == k_ref
    assert rel(x, x_ref) < 1e-12


@pytest.mark.parametrize("smoother", ["sor", "ssor", "mc_sor", "l1_jacobi"])
def test_amg_solve_full_multigrid(gpu, smoother):
    """AMG_solve with the FULL_AMG cycle (one full-multigrid pass): the
    device pass against the oracle's, FP64 to 1e-12, on a scalar and a
    two-field system, from a nonzero guess (which the pass replaces); the
    mixed-precision pass within FP32 accuracy of the FP64 one."""
    for s, depth in ((P.poisson(33), 4), (P.degraded_elasticity(25), 4)):
        say("full multigrid", smoother, s.n)
        orc = Oracle()
        I = orc.amg_setup(s.rowptr, s.col, s.val, dof=s.dof, depth=depth)
        x0 = np.cos(np.arange(s.n) * 0.3)
        k_ref, x_ref = orc.amg_solve(I, s.b, x0, smoother=smoother, alpha=1.2, cycle="full")
        orc.amg_free(I)
        prm = gpu.params("reference", s.dof, depth, smoother=smoother, omega=1.2, cycle="full")
        amg = gpu.AMG(s.rowptr, s.col, s.val, prm)
        k, x = amg.amg_solve(s.b, x0)
        amg.set_params(gpu.params("reference", s.dof, depth, smoother=smoother, omega=1.2, cycle="full",
                                  precision="mixed"))
        km, xm = amg.amg_solve(s.b, x0)
        amg.close()
        say("  oracle", k_ref, "gpu", k, "rel", rel(x, x_ref), "mixed rel", rel(xm, x_ref))
        assert k == k_ref == 1 and km == 1
        assert rel(x, x_ref) < 1e-12
        assert rel(xm, x_ref) < 1e-5


def test_mixed_precision_reaches_fp64_accuracy(gpu):
    s = P.degraded_elasticity(33)
    say("mixed precision elasticity 33")
    A = s.scipy()
    f = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("vcycle", 2, 4))
    m = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("vcycle", 2, 4, precision="mixed"))
    g64 = f.pcg(s.b, tol=1e-10, max_iter=3000, stop="relres")
    g32 = m.pcg(s.b, tol=1e-10, max_iter=3000, stop="relres")
    say("  iters fp64", g64["iters"], "mixed", g32["iters"])
    for g in (g64, g32):
        assert np.linalg.norm(s.b - A @ g["x"]) / np.linalg.norm(s.b) < 1e-9
    assert rel(g32["x"], g64["x"]) < 1e-6


@pytest.mark.parametrize("kind,prec", [("reference", "fp64"), ("vcycle", "mixed")])
def test_recompute_values_only_upload(gpu, kind, prec):
    """AMG_recompute_coarse_matrices on an uploaded hierarchy: after a first
    recompute (a full re-upload, from which the hierarchy keeps what a
    values-only one compares), new finest values of the same pattern
    re-upload the levels' values only (every level's pattern, colouring and
    level schedules unchanged).  The pcg afterwards is bit for bit that of a
    hierarchy that reached the same values through full re-uploads (an
    intermediate recompute with zeroed entries changes the coarse patterns,
    so both of its re-uploads are full)."""
    s1 = P.damage_system(65, seed_h=21)[0]
    s2 = P.damage_system(65, seed_h=22)[0]
    assert np.array_equal(s1.rowptr, s2.rowptr) and np.array_equal(s1.col, s2.col)
    mid = s2.val.copy()
    off = np.nonzero(s2.col != np.repeat(np.arange(s2.n), np.diff(s2.rowptr)))[0]
    mid[off[::3]] = 0.0
    prm = gpu.params(kind, 1, 4, precision=prec, autotune=0)
    say("recompute values-only", kind, prec)
    a = gpu.AMG(s1.rowptr, s1.col, s1.val, prm)
    a.recompute(s1.val)
    a.recompute(s2.val)
    b = gpu.AMG(s1.rowptr, s1.col, s1.val, prm)
    b.recompute(mid)
    b.recompute(s2.val)
    for l in range(a.depth):
        assert np.array_equal(a.level(l)["A"][2], b.level(l)["A"][2]), l
    ra = a.pcg(s2.b, tol=1e-10, max_iter=500)
    rb = b.pcg(s2.b, tol=1e-10, max_iter=500)
    say("  iters", ra["iters"], rb["iters"])
    assert ra["iters"] == rb["iters"]
    assert np.array_equal(ra["x"], rb["x"])
    A = s2.scipy()
    assert np.linalg.norm(s2.b - A @ ra["x"]) <= 1e-8 * np.linalg.norm(s2.b)
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["poisson17", "poisson33", "poisson33_std", "pair17"])
def test_amg_solve_reference_jacobi_matches_golden(gpu, case):
    """The reference's undamped AMG_parallel_Jacobi_smoother
    (Multigrid.c:1367-1400) as the V-cycle smoother of AMG_solve on the
    device: equal V-cycle counts, iterates to 1e-12.  Scalar systems against
    the compiled reference's own result (golden vectors: AMG_solve from a
    zero guess, eps 1e-6, at most 6 V-cycles); a two-field system (two
    Poisson fields in the blocked [u; v] order, dof 2, where the smoother
    works on node blocks) against the oracle, whose Jacobi sweep equals the
    reference's bit for bit.  (On the elasticity systems the undamped
    Jacobi V-cycle diverges -- the golden iterates grow to 1e22 in six
    cycles -- so they carry no comparison.)"""
    import os
    import scipy.sparse as sps
    if case == "pair17":
        s1 = P.poisson(17)
        A = sps.block_diag([s1.scipy(), s1.scipy()]).tocsr()
        A.sort_indices()
        rp, ci, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
        b = np.concatenate([s1.b, 0.5 * s1.b[::-1]])
        dof, depth, coarsening = 2, 4, 1
        orc = Oracle()
        I = orc.amg_setup(rp, ci, v, dof=dof, depth=depth, coarsening=coarsening)
        k_ref, x_ref = orc.amg_solve(I, b, np.zeros(len(b)), smoother="jacobi", alpha=1.0, eps=1e-6, max_iter=6)
        orc.amg_free(I)
    else:
        g = np.load(os.path.join(os.path.dirname(__file__), "golden", f"{case}.npz"))
        rp, ci, v, b = g["rowptr"], g["col"], g["val"], g["b"]
        dof, depth, coarsening = int(g["dof"]), int(g["depth"]), int(g["coarsening"])
        k_ref, x_ref = int(g["amgsolve_jacobi_k"]), g["amgsolve_jacobi_x"]
    say("reference jacobi", case, "dof", dof, "depth", depth, "coarsening", coarsening)
    prm = gpu.params("reference", dof, depth, coarsening=coarsening, smoother="jacobi", omega=1.0, amg_eps=1e-6,
                     amg_max_iter=6)
    amg = gpu.AMG(rp, ci, v, prm)
    k, x = amg.amg_solve(b, np.zeros(len(b)))
    amg.close()
    say("  vcycles reference", k_ref, "gpu", k, "rel", rel(x, x_ref), "|x|", float(np.linalg.norm(x_ref)))
    assert np.linalg.norm(x_ref) < 10 * np.linalg.norm(b)      # a convergent case
    assert k == k_ref
    assert rel(x, x_ref) < 1e-12


@pytest.mark.parametrize("kind", ["poisson33", "damage65", "elasticity33"])
def test_device_recompute_matches_oracle(gpu, kind, capfd, monkeypatch):
    """AMG_recompute_coarse_matrices on the device (the Galerkin products
    R (A P) of every level from a per-pattern plan in LinA's summation
    order): every level's matrix equals the oracle's recompute
    (orc_amg_recompute, pinned to the reference) bit for bit, over two
    successive sets of new finest values; the device path is the one taken
    (its progress line), and the pcg with the recomputed hierarchy equals the
    oracle's pcg on the new system (multicolour SOR: a hierarchy with the
    exact lexicographic SOR's packed records recomputes on the host)."""
    monkeypatch.setenv("LINA_RN_TIMING", "1")
    base = {"poisson33": lambda: P.poisson(33), "damage65": lambda: P.damage_system(65)[0],
            "elasticity33": lambda: P.degraded_elasticity(33)}[kind]()
    depth = 4
    orc = Oracle()
    h = orc.amg_setup(base.rowptr, base.col, base.val, dof=base.dof, depth=depth)
    # multicolour SOR: the exact lexicographic SOR's packed records keep the host path
    amg = gpu.AMG(base.rowptr, base.col, base.val, gpu.params("reference", base.dof, depth, smoother="mc_sor"))
    rng = np.random.default_rng(8)
    rows = np.repeat(np.arange(base.n), np.diff(base.rowptr))
    for trial in range(2):
        # S A S with a random positive diagonal S: new values of the same
        # pattern that keep A symmetric positive definite (explicit zeros stay)
        sc = 1.0 + 0.4 * rng.random(base.n)
        v = base.val * sc[rows] * sc[base.col]
        capfd.readouterr()
        amg.recompute(v)
        err = capfd.readouterr().err
        say("  ", kind, "trial", trial, [l for l in err.splitlines() if "recompute" in l])
        assert "recompute on the device: done" in err
        orc.L.orc_amg_recompute(h, v.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)))
        ol = orc.amg_levels(h, depth)
        for l in range(depth):
            g = amg.level(l)["A"]
            o = ol[l]["A"]
            assert np.array_equal(g[0], o.rowptr) and np.array_equal(g[1], o.col), (trial, l)
            assert np.array_equal(g[2], o.val), (trial, l)
    import ctypes as C
    # pcg preconditioned by the recomputed hierarchies (P and R of the first setup)
    xo, bo = np.zeros(base.n), base.b.copy()
    res, vc = C.c_double(), C.c_int()
    dp = C.POINTER(C.c_double)
    it = orc.L.orc_pcg_amg(h, xo.ctypes.data_as(dp), bo.ctypes.data_as(dp), 1e-10, 500, 2, 1.6, 1e-2, 100, 0, 0,
                           C.byref(res), C.byref(vc))
    g = amg.pcg(base.b, tol=1e-10, max_iter=500)
    amg.close()
    orc.amg_free(h)
    say("  pcg iters oracle", it, "gpu", g["iters"], "rel", rel(g["x"], xo))
    assert abs(g["iters"] - it) <= 1 and rel(g["x"], xo) < 1e-12


@pytest.mark.parametrize("m,depth", [(65, 3), (129, 3)])
def test_blocked_coarse_inverse_matches_scalar_steps(gpu, m, depth, monkeypatch):
    """The coarsest level's dense inverse by 64-pivot blocks with the FP64
    MFMA update against the unblocked Gauss-Jordan steps (LINA_GJ=scalar):
    coarse systems of several blocks, pcg iterations equal and iterates to
    1e-10 (the two differ only in rounding)."""
    s = P.poisson(m)
    out = []
    for mode in ("scalar", "blocked"):
        if mode == "scalar":
            monkeypatch.setenv("LINA_GJ", "scalar")
        else:
            monkeypatch.delenv("LINA_GJ", raising=False)
        amg = gpu.AMG(s.rowptr, s.col, s.val, gpu.params("reference", 1, depth))
        nd = amg.level(0)["fine"]
        g = amg.pcg(s.b, tol=1e-10, max_iter=500)
        amg.close()
        out.append((g["iters"], g["x"]))
        say("  coarse inverse", mode, "coarsest rows", nd, "pcg iters", g["iters"])
    assert nd > 64
    assert out[0][0] == out[1][0] and rel(out[1][1], out[0][1]) < 1e-10
