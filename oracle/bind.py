"""ctypes bindings for the test-side CPU libraries under oracle/.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

* ``Ref``    -> oracle/_ref/libreflina.so, the reference LinA compiled from its
               own sources (+ oracle/ref_harness.c glue).
* ``Oracle`` -> oracle/liboracle.so, the C restatement in oracle/lina_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libreflina.so")
REF_FAST_SO = os.path.join(HERE, "_ref", "libreflina_fast.so")  # -O3 build, timings only
ORACLE_SO = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

SMOOTHERS = {"sor": 0, "jacobi": 1, "mc_sor": 2, "l1_jacobi": 3, "mc_ssor": 4, "ssor": 5, "hybrid": 6}


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class CSR:
    def __init__(self, rowptr, col, val, ncols=None):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        self.ncols = ncols

    def dense(self, ncols=None):
        nc = ncols or self.ncols or (int(self.col.max()) + 1 if len(self.col) else self.n)
        D = np.zeros((self.n, nc))
        for i in range(self.n):
            for k in range(self.rowptr[i], self.rowptr[i + 1]):
                D[i, self.col[k]] += self.val[k]
        return D

    def matvec(self, x):
        y = np.zeros(self.n)
        for i in range(self.n):
            s = 0.0
            for k in range(self.rowptr[i], self.rowptr[i + 1]):
                s += self.val[k] * x[self.col[k]]
            y[i] = s
        return y


class Ref:
    """The compiled reference (oracle/_ref/libreflina.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        L = C.CDLL(path)
        self.L = L
        L.refh_init.argtypes = [C.c_int]
        L.refh_from_csr.argtypes = [C.c_int, _ip, _ip, _dp]
        L.refh_from_csr.restype = _vp
        L.refh_free.argtypes = [_vp]
        L.refh_size.argtypes = [_vp]
        L.refh_nnz.argtypes = [_vp]
        L.refh_to_csr.argtypes = [_vp, _ip, _ip, _dp]
        L.refh_amg_setup.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.refh_amg_setup.restype = _vp
        L.refh_amg_free.argtypes = [_vp]
        L.refh_set_cycle.argtypes = [C.c_int]
        L.refh_amg_sizes.argtypes = [_vp, _ip, _ip]
        for f in ("refh_amg_A", "refh_amg_P", "refh_amg_R", "refh_amg_L", "refh_amg_U"):
            getattr(L, f).argtypes = [_vp, C.c_int]
            getattr(L, f).restype = _vp
        L.refh_amg_signs.argtypes = [_vp]
        L.refh_amg_signs.restype = _vp
        L.refh_amg_labels.argtypes = [_vp, C.c_int, _ip]
        L.refh_amg_recompute.argtypes = [_vp, _vp]
        L.refh_amg_solve.argtypes = [_vp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int]
        L.refh_pcg_amg.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _ip, C.c_int, _ip,
                                   _dp, _dp, C.POINTER(_vp)]
        L.refh_pcg_amg_default.argtypes = [_vp, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp]
        L.refh_pcg_ilu.argtypes = [_vp, _dp, _dp, C.c_double, C.c_int, _dp]
        L.refh_ilu.argtypes = [_vp, C.POINTER(_vp), C.POINTER(_vp)]
        L.refh_L_invert.argtypes = [_vp, _dp]
        L.refh_U_invert.argtypes = [_vp, _dp]
        L.refh_sparse_LU.argtypes = [_vp, C.POINTER(_vp), C.POINTER(_vp)]
        L.refh_product.argtypes = [_vp, _vp]
        L.refh_product.restype = _vp
        L.refh_colour_count.argtypes = [_vp, _ip]
        L.refh_bicgstab_amg.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.refh_reflective_newton.argtypes = [_vp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_double, C.c_double, _dp]
        L.refh_set_hybrid.argtypes = [C.c_int]
        L.refh_gmres_ilu.argtypes = [_vp, _dp, _dp, C.c_int, C.c_double, C.c_int, _dp]
        L.refh_amg_assemble.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp),
                                        C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_ip), _dp, C.c_int, _ip, _dp, _dp]
        L.refh_amg_assemble.restype = _vp
        L.refh_set_threads.argtypes = [C.c_int]
        L.refh_pcg_assembled.argtypes = [_vp, _vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp,
                                         C.c_double, C.c_int, _dp]
        L.refh_init(1)

    # -- matrices
    def matrix(self, rowptr, col, val):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        return self.L.refh_from_csr(len(rowptr) - 1, iptr(rowptr), iptr(col), dptr(val))

    def to_csr(self, M) -> CSR:
        n = self.L.refh_size(M)
        nnz = self.L.refh_nnz(M)
        rp = np.zeros(n + 1, np.int32)
        c = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1), np.float64)
        self.L.refh_to_csr(M, iptr(rp), iptr(c), dptr(v))
        return CSR(rp, c[:nnz], v[:nnz])

    def free(self, M):
        self.L.refh_free(M)

    # -- an assembled hierarchy (timings of the full-size benchmark system)
    def assemble(self, levels, dof, signs, smiter=1, aggr_smiter=1, coarsening=1):
        """amg_system_info from level dicts {'A','P','R': (rowptr, col, val),
        'labels'} (index 0 coarsest), the finest row signs; the coarsest
        level is solved with the dense inverse of its dense rows.  Returns
        (handle, keep) -- keep holds what the handle points into."""
        import numpy as np
        depth = len(levels)
        keep = []

        def mat(t):
            return self.matrix(*t)
        A = (_vp * depth)(*[mat(lv["A"]) for lv in levels])
        P = (_vp * depth)(*[mat(lv["P"]) if "P" in lv else None for lv in levels])
        R = (_vp * depth)(*[mat(lv["R"]) if "R" in lv else None for lv in levels])
        labs = [np.ascontiguousarray(lv["labels"], np.int32) if "labels" in lv else None for lv in levels]
        keep.extend(labs)
        Lp = (_ip * depth)(*[iptr(a) if a is not None else None for a in labs])
        rp, ci, v = levels[0]["A"]
        n0 = len(rp) - 1
        nzc = np.array([np.count_nonzero(v[rp[i]:rp[i + 1]]) for i in range(n0)])
        didx = np.flatnonzero(nzc > 1).astype(np.int32)
        pos = -np.ones(n0, np.int64)
        pos[didx] = np.arange(len(didx))
        D = np.zeros((len(didx), len(didx)))
        tdiag = np.zeros(n0)
        for i in range(n0):
            cols, vals = ci[rp[i]:rp[i + 1]], v[rp[i]:rp[i + 1]]
            if pos[i] >= 0:
                m = pos[cols] >= 0
                D[pos[i], pos[cols[m]]] = vals[m]
            else:
                tdiag[i] = vals[cols == i][0] if np.any(cols == i) else 0.0
        inv = np.ascontiguousarray(np.linalg.inv(D)) if len(didx) else np.zeros(1)
        sg = np.ascontiguousarray(signs, np.float64)
        keep.extend([didx, inv, tdiag, sg])
        h = self.L.refh_amg_assemble(depth, dof, smiter, aggr_smiter, coarsening, A, P, R, Lp, dptr(sg), len(didx),
                                     iptr(didx), dptr(inv), dptr(tdiag))
        return h, (keep, A)

    def pcg_assembled(self, h, M, b, smoother="ssor", alpha=1.0, amg_eps=1e-2, amg_max_iter=1, zero_init=True,
                      tol=1e-10, max_iter=5):
        import numpy as np
        X = np.zeros(len(b))
        B = np.array(b, dtype=np.float64)
        t = C.c_double()
        it = self.L.refh_pcg_assembled(h, M, SMOOTHERS[smoother], alpha, amg_eps, amg_max_iter, int(zero_init),
                                       dptr(X), dptr(B), tol, max_iter, C.byref(t))
        return {"iters": it, "x": X, "t_solve": t.value}

    # -- AMG hierarchy
    def amg_setup(self, M, dof=1, depth=4, smiter=2, aggr_smiter=3, coarsening=1, matrix_tol=None, cycle="v"):
        """cycle "v" (STAND_ALONE) or "full" (FULL_AMG, Multigrid.h:38)"""
        self.L.refh_set_cycle({"v": 0, "full": 2}[cycle])
        try:
            return self.L.refh_amg_setup(M, dof, depth, smiter, aggr_smiter, coarsening,
                                         0 if matrix_tol is None else 1, matrix_tol or 0.0)
        finally:
            self.L.refh_set_cycle(0)

    def amg_levels(self, I, depth):
        fine = np.zeros(depth, np.int32)
        coarse = np.zeros(depth, np.int32)
        self.L.refh_amg_sizes(I, iptr(fine), iptr(coarse))
        out = []
        for l in range(depth):
            d = {"fine": int(fine[l]), "coarse": int(coarse[l]), "A": self.to_csr(self.L.refh_amg_A(I, l))}
            if l > 0:
                d["P"] = self.to_csr(self.L.refh_amg_P(I, l))
                d["R"] = self.to_csr(self.L.refh_amg_R(I, l))
                lab = np.zeros(int(fine[l]), np.int32)
                self.L.refh_amg_labels(I, l, iptr(lab))
                d["labels"] = lab
            else:
                d["L"] = self.to_csr(self.L.refh_amg_L(I, 0))
                d["U"] = self.to_csr(self.L.refh_amg_U(I, 0))
            out.append(d)
        return out

    def amg_signs(self, I):
        return self.to_csr(self.L.refh_amg_signs(I))

    def amg_free(self, I):
        self.L.refh_amg_free(I)

    def amg_solve(self, I, F, X, smoother="sor", alpha=1.6, eps=1e-2, max_iter=100, coarse_sweeps=2):
        F = np.array(F, dtype=np.float64)
        X = np.array(X, dtype=np.float64)
        self.L.refh_set_hybrid(coarse_sweeps)
        k = self.L.refh_amg_solve(I, dptr(F), dptr(X), SMOOTHERS[smoother], alpha, eps, max_iter)
        return k, X

    # -- solvers
    def pcg_amg(self, M, b, x0=None, dof=1, depth=4, smiter=2, aggr_smiter=3, coarsening=1, smoother="sor",
                alpha=1.6, amg_eps=1e-2, amg_max_iter=100, zero_init=False, tol=1e-10, max_iter=1000, keep=False,
                coarse_sweeps=2, cycle="v"):
        n = self.L.refh_size(M)
        self.L.refh_set_hybrid(coarse_sweeps)
        self.L.refh_set_cycle({"v": 0, "full": 2}[cycle])
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        res = C.c_double()
        cap = 100000
        calls = np.zeros(cap, np.int32)
        ncalls = C.c_int()
        ts = C.c_double()
        tv = C.c_double()
        kept = _vp()
        it = self.L.refh_pcg_amg(M, dof, depth, smiter, aggr_smiter, coarsening, SMOOTHERS[smoother], alpha,
                                 amg_eps, amg_max_iter, int(zero_init), dptr(X), dptr(B), tol, max_iter, C.byref(res),
                                 iptr(calls), cap, C.byref(ncalls), C.byref(ts), C.byref(tv),
                                 C.byref(kept) if keep else None)
        self.L.refh_set_cycle(0)
        out = {"iters": it, "x": X, "b_out": B, "res": res.value, "vcycles": calls[:min(ncalls.value, cap)].copy(),
               "t_setup": ts.value, "t_solve": tv.value}
        if keep:
            out["amg"] = kept.value
        return out

    def pcg_ilu(self, M, b, x0=None, tol=1e-10, max_iter=1000):
        n = self.L.refh_size(M)
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        res = C.c_double()
        it = self.L.refh_pcg_ilu(M, dptr(X), dptr(B), tol, max_iter, C.byref(res))
        return {"iters": it, "x": X, "b_out": B, "res": res.value}

    def ilu(self, M):
        Lp, Up = _vp(), _vp()
        ok = self.L.refh_ilu(M, C.byref(Lp), C.byref(Up))
        Lc, Uc = self.to_csr(Lp), self.to_csr(Up)
        self.free(Lp)
        self.free(Up)
        return ok, Lc, Uc

    def gmres_ilu(self, M, b, x0=None, m=30, tol=1e-8, max_iter=1000):
        n = self.L.refh_size(M)
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        eta = C.c_double()
        it = self.L.refh_gmres_ilu(M, dptr(X), dptr(B), m, tol, max_iter, C.byref(eta))
        return {"iters": it, "x": X, "eta": eta.value}

    def product(self, A, B):
        P = self.L.refh_product(A, B)
        out = self.to_csr(P)
        self.free(P)
        return out

    def colours(self, M):
        n = self.L.refh_size(M)
        c = np.zeros(n, np.int32)
        k = self.L.refh_colour_count(M, iptr(c))
        return k, c

    def bicgstab_amg(self, M, b, x0=None, dof=1, depth=4, smoother="sor", alpha=1.0, eps=1e-8, max_iter=1000):
        n = self.L.refh_size(M)
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        it = self.L.refh_bicgstab_amg(M, dof, depth, SMOOTHERS[smoother], alpha, eps, max_iter, dptr(X), dptr(B))
        return {"iters": it, "x": X}

    def rn_get_M_D(self, H, G, X, lower, upper):
        n = len(X)
        D = np.zeros(n)
        f = self.L.refh_rn_get_M_D
        f.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
        f.restype = _vp
        M = f(H, dptr(np.ascontiguousarray(G, dtype=np.float64)), dptr(np.ascontiguousarray(X, dtype=np.float64)),
              dptr(np.ascontiguousarray(lower, dtype=np.float64)), dptr(np.ascontiguousarray(upper, dtype=np.float64)),
              dptr(D))
        out = self.to_csr(M)
        self.free(M)
        return out, D

    def reflective_newton(self, H, b, lower, upper, x0, solver=3, dof=1, depth=4, max_iter=50, eps=1e-8, delta=250.0):
        X = np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        lo = np.ascontiguousarray(lower, dtype=np.float64)
        up = np.ascontiguousarray(upper, dtype=np.float64)
        q = C.c_double()
        it = self.L.refh_reflective_newton(H, dptr(b), dptr(lo), dptr(up), dptr(X), solver, dof, depth, max_iter,
                                           eps, delta, C.byref(q))
        return {"iters": it, "x": X, "q": q.value}


class Oracle:
    """The C restatement (oracle/liboracle.so)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle oracle)")
        L = C.CDLL(path)
        self.L = L
        L.orc_amg_setup.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_double]
        L.orc_amg_setup.restype = _vp
        L.orc_amg_free.argtypes = [_vp]
        L.orc_amg_level.argtypes = [_vp, C.c_int, _ip, _ip, _ip, _ip, _ip]
        L.orc_amg_copy.argtypes = [_vp, C.c_int, C.c_char, _ip, _ip, _dp]
        L.orc_amg_labels.argtypes = [_vp, C.c_int, _ip]
        L.orc_amg_recompute.argtypes = [_vp, _dp]
        L.orc_amg_solve.argtypes = [_vp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_amg_solve_fmg.argtypes = [_vp, _dp, _dp, C.c_int, C.c_double]
        L.orc_pcg_amg.argtypes = [_vp, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.c_int, C.c_int, _dp, _ip]
        L.orc_colouring.argtypes = [C.c_int, _ip, _ip, _dp, _ip]
        L.orc_set_coarse_smoother.argtypes = [_vp, C.c_int, C.c_int]
        L.orc_set_cycle.argtypes = [_vp, C.c_int]
        L.orc_spmv.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_ilu0.argtypes = [C.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip, _dp]
        L.orc_gmres_ilu.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, _dp]
        L.orc_pcg_ilu.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, _dp]
        L.orc_ilu_apply.argtypes = [C.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _dp, _dp]
        L.orc_bicgstab_amg.argtypes = [_vp, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                       C.c_int, C.c_double, C.c_int, _dp]

    def amg_setup(self, rowptr, col, val, dof=1, depth=4, smiter=2, aggr_smiter=3, coarsening=1,
                  eps_minus=0.3, eps_plus=0.5, matrix_tol=0.1):
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        h = self.L.orc_amg_setup(len(rp) - 1, iptr(rp), iptr(ci), dptr(v), dof, depth, smiter, aggr_smiter,
                                 coarsening, eps_minus, eps_plus, matrix_tol)
        if not h:
            raise RuntimeError("oracle AMG setup failed")
        return h

    def amg_free(self, h):
        self.L.orc_amg_free(h)

    def amg_levels(self, h, depth):
        out = []
        for l in range(depth):
            f, c, na, np_, nr = (C.c_int() for _ in range(5))
            self.L.orc_amg_level(h, l, C.byref(f), C.byref(c), C.byref(na), C.byref(np_), C.byref(nr))
            d = {"fine": f.value, "coarse": c.value}
            for which, nnz, rows in (("A", na.value, f.value), ("P", np_.value, f.value), ("R", nr.value, c.value)):
                if l == 0 and which != "A":
                    continue
                rp = np.zeros(rows + 1, np.int32)
                ci = np.zeros(max(nnz, 1), np.int32)
                v = np.zeros(max(nnz, 1), np.float64)
                self.L.orc_amg_copy(h, l, which.encode(), iptr(rp), iptr(ci), dptr(v))
                d[which] = CSR(rp, ci[:nnz], v[:nnz])
            if l > 0:
                lab = np.zeros(f.value, np.int32)
                self.L.orc_amg_labels(h, l, iptr(lab))
                d["labels"] = lab
            out.append(d)
        return out

    def amg_solve(self, h, F, X, smoother="sor", alpha=1.6, eps=1e-2, max_iter=100, cycle="v"):
        """cycle "v" (STAND_ALONE) or "full" (FULL_AMG: one full-multigrid
        pass; eps and max_iter unused)"""
        F = np.array(F, dtype=np.float64)
        X = np.array(X, dtype=np.float64)
        if cycle == "full":
            k = self.L.orc_amg_solve_fmg(h, dptr(F), dptr(X), SMOOTHERS[smoother], alpha)
        else:
            k = self.L.orc_amg_solve(h, dptr(F), dptr(X), SMOOTHERS[smoother], alpha, eps, max_iter)
        return k, X

    def pcg_amg(self, rowptr, col, val, b, x0=None, dof=1, depth=4, smiter=2, aggr_smiter=3, coarsening=1,
                smoother="sor", alpha=1.6, amg_eps=1e-2, amg_max_iter=100, zero_init=False, tol=1e-10,
                max_iter=1000, stop="reference", coarse_sweeps=2, cycle="v"):
        """smoother "hybrid": MC-SSOR on the finest level, l1-Jacobi with
        coarse_sweeps sweeps below (the harness's refh_hybrid_smoother);
        cycle "full": every preconditioner call is one full-multigrid pass."""
        h = self.amg_setup(rowptr, col, val, dof, depth, smiter, aggr_smiter, coarsening)
        self.L.orc_set_cycle(h, {"v": 0, "full": 2}[cycle])
        if smoother == "hybrid":
            self.L.orc_set_coarse_smoother(h, SMOOTHERS["l1_jacobi"], coarse_sweeps)
            smoother = "mc_ssor"
        n = len(rowptr) - 1
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        res = C.c_double()
        vc = C.c_int()
        it = self.L.orc_pcg_amg(h, dptr(X), dptr(B), tol, max_iter, SMOOTHERS[smoother], alpha, amg_eps,
                                amg_max_iter, int(zero_init), 1 if stop == "relres" else 0, C.byref(res),
                                C.byref(vc))
        self.amg_free(h)
        return {"iters": it, "x": X, "b_out": B, "res": res.value, "vcycles": vc.value}

    def colouring(self, rowptr, col, val):
        n = len(rowptr) - 1
        c = np.zeros(n, np.int32)
        k = self.L.orc_colouring(n, iptr(np.ascontiguousarray(rowptr, dtype=np.int32)),
                                 iptr(np.ascontiguousarray(col, dtype=np.int32)),
                                 dptr(np.ascontiguousarray(val, dtype=np.float64)), iptr(c))
        return k, c

    def spmv(self, rowptr, col, val, x):
        n = len(rowptr) - 1
        y = np.zeros(n)
        self.L.orc_spmv(n, iptr(np.ascontiguousarray(rowptr, dtype=np.int32)),
                        iptr(np.ascontiguousarray(col, dtype=np.int32)),
                        dptr(np.ascontiguousarray(val, dtype=np.float64)),
                        dptr(np.ascontiguousarray(x, dtype=np.float64)), dptr(y))
        return y

    def ilu0(self, rowptr, col, val):
        """(ok, L, U) as CSR; the factors are meaningful only when ok."""
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        n, cap = len(rp) - 1, len(ci) + len(rp)
        lrp, urp = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
        lci, uci = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        lv, uv = np.zeros(cap), np.zeros(cap)
        ok = self.L.orc_ilu0(n, iptr(rp), iptr(ci), dptr(v), iptr(lrp), iptr(lci), dptr(lv), iptr(urp), iptr(uci),
                             dptr(uv))
        if not ok:
            return 0, None, None
        return 1, CSR(lrp, lci[:lrp[n]], lv[:lrp[n]]), CSR(urp, uci[:urp[n]], uv[:urp[n]])

    def gmres_ilu(self, rowptr, col, val, b, x0=None, m=30, tol=1e-8, max_iter=1000):
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        n = len(rp) - 1
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.ascontiguousarray(b, dtype=np.float64)
        eta = C.c_double()
        it = self.L.orc_gmres_ilu(n, iptr(rp), iptr(ci), dptr(v), dptr(X), dptr(B), m, tol, max_iter, C.byref(eta))
        return {"iters": it, "x": X, "eta": eta.value}

    def bicgstab_amg(self, rowptr, col, val, b, x0=None, dof=1, depth=4, smoother="sor", alpha=1.0, eps=1e-8,
                     max_iter=1000, amg_eps=1e-4, amg_max_iter=20):
        """BiCGStab with the AMG preconditioner set up as BI_set_AMG_preconditioner does."""
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        n = len(rp) - 1
        h = self.amg_setup(rp, ci, v, dof, depth, 2, 3, 1, matrix_tol=0.1)
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.ascontiguousarray(b, dtype=np.float64)
        res = C.c_double()
        try:
            it = self.L.orc_bicgstab_amg(h, n, iptr(rp), iptr(ci), dptr(v), dptr(X), dptr(B), SMOOTHERS[smoother],
                                         alpha, amg_eps, amg_max_iter, eps, max_iter, C.byref(res))
        finally:
            self.amg_free(h)
        return {"iters": it, "x": X, "res": res.value}

    def ilu_apply(self, L: CSR, U: CSR, z):
        """w = U^-1 L^-1 z with the factors of ilu0."""
        n = L.n
        w = np.zeros(n)
        self.L.orc_ilu_apply(n, iptr(L.rowptr), iptr(L.col), dptr(L.val), iptr(U.rowptr), iptr(U.col), dptr(U.val),
                             dptr(np.ascontiguousarray(z, dtype=np.float64)), dptr(w))
        return w

    def pcg_ilu(self, rowptr, col, val, b, x0=None, tol=1e-10, max_iter=1000):
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        n = len(rp) - 1
        X = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        res = C.c_double()
        it = self.L.orc_pcg_ilu(n, iptr(rp), iptr(ci), dptr(v), dptr(X), dptr(B), tol, max_iter, C.byref(res))
        return {"iters": it, "x": X, "b_out": B, "res": res.value}
