"""ctypes binding of liblina_hip.so (the handle API of include/lina_hip.h).

The library is built in-tree by ``fem_amd.build``; importing this module
never falls back to anything else: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LINA_LIB_PATH", os.path.join(HERE, "liblina_hip.so"))

SMOOTHER = {"sor": 0, "jacobi": 1, "mc_sor": 2, "l1_jacobi": 3, "mc_ssor": 4, "ssor": 5}
PREC = {"fp64": 0, "mixed": 1}
STOP = {"reference": 0, "relres": 1}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [("dof", C.c_int), ("depth", C.c_int), ("smoothing_iterations", C.c_int),
                ("aggr_smoothing_iterations", C.c_int), ("coarsening", C.c_int), ("eps_minus", C.c_double),
                ("eps_plus", C.c_double), ("matrix_tol", C.c_double), ("smoother", C.c_int), ("omega", C.c_double),
                ("amg_eps", C.c_double), ("amg_max_iter", C.c_int), ("zero_init", C.c_int),
                ("precision", C.c_int), ("coarse_smoother", C.c_int), ("coarse_sweeps", C.c_int),
                ("autotune", C.c_int), ("cycle", C.c_int), ("skip_levels", C.c_uint)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("precon_calls", C.c_int), ("vcycles", C.c_int), ("residual", C.c_double),
                ("relres", C.c_double), ("t_solve_ms", C.c_double), ("spmv_ms", C.c_double),
                ("spmv_launches", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RNParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("eps", C.c_double), ("sigma_l", C.c_double), ("sigma_u", C.c_double),
                ("rho", C.c_double), ("delta", C.c_double), ("xi", C.c_double), ("trust_lower", C.c_double),
                ("trust_upper", C.c_double), ("radius_min", C.c_double), ("radius_max", C.c_double),
                ("solver", C.c_int), ("pcg_iter", C.c_int), ("bi_eps", C.c_double), ("bi_max_iter", C.c_int),
                ("gmres_m", C.c_int), ("gmres_max_iter", C.c_int), ("gmres_tol", C.c_double),
                ("amg", Params), ("seed", C.c_ulonglong), ("q_offset", C.c_double), ("host_solver", C.c_void_p),
                ("host_ctx", C.c_void_p), ("ilu_ordering", C.c_int)]


class RNStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("line_searches", C.c_int), ("line_search_iter", C.c_int),
                ("neg_curvature", C.c_int), ("inner_iterations", C.c_int), ("dq", C.c_double), ("q", C.c_double),
                ("delta", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


RN_SOLVER = {"bicgstab": 1, "amg": 2, "pcg": 3, "gmres": 4, "host": 5}


class BatchStats(C.Structure):
    _fields_ = [("iterations_launched", C.c_int), ("active", C.c_int), ("t_solve_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LinaError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LinaError(f"{LIB_PATH} is not built: run python -m fem_amd.build")
    L = C.CDLL(LIB_PATH)
    L.lina_last_error.restype = C.c_char_p
    L.lina_device_count.restype = C.c_int
    L.lina_params_reference.argtypes = [C.POINTER(Params), C.c_int, C.c_int]
    L.lina_params_vcycle.argtypes = [C.POINTER(Params), C.c_int, C.c_int]
    L.lina_amg_create.argtypes = [C.c_int, _ip, _ip, _dp, C.POINTER(Params), C.c_int]
    L.lina_amg_create.restype = C.c_void_p
    L.lina_amg_destroy.argtypes = [C.c_void_p]
    L.lina_amg_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
    L.lina_amg_depth.argtypes = [C.c_void_p]
    L.lina_amg_level_info.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip]
    L.lina_amg_level_nnz.argtypes = [C.c_void_p, C.c_int, C.c_char]
    L.lina_amg_level_copy.argtypes = [C.c_void_p, C.c_int, C.c_char, _ip, _ip, _dp]
    L.lina_amg_labels.argtypes = [C.c_void_p, C.c_int, _ip]
    L.lina_amg_colours.argtypes = [C.c_void_p, C.c_int, _ip]
    L.lina_amg_signs.argtypes = [C.c_void_p, _dp]
    L.lina_amg_recompute.argtypes = [C.c_void_p, _dp]
    L.lina_pcg.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_int, C.POINTER(Stats)]
    L.lina_pcg_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int,
                                  C.POINTER(Stats)]
    L.lina_amg_solve.argtypes = [C.c_void_p, _dp, _dp]
    L.lina_spmv.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
    L.lina_time_spmv.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.lina_time_spmv.restype = C.c_double
    L.lina_spmv_bytes.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.lina_spmv_bytes.restype = C.c_double
    L.lina_amg_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    L.lina_amg_tuned.argtypes = [C.c_void_p, _ip, C.c_char_p, _ip, _ip, C.c_int]
    L.lina_dev_alloc.argtypes = [C.c_long]
    L.lina_dev_alloc.restype = C.c_void_p
    L.lina_dev_free.argtypes = [C.c_void_p]
    L.lina_dev_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_int]
    L.lina_dev_fill.argtypes = [C.c_void_p, C.c_long, C.c_double]
    L.lina_matrix_create.argtypes = [C.c_int, _ip, _ip, _dp]
    L.lina_matrix_create.restype = C.c_void_p
    L.lina_matrix_destroy.argtypes = [C.c_void_p]
    L.lina_matrix_spmv.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.lina_matrix_power.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, _dp]
    L.lina_matrix_spmv_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.lina_matrix_time_spmv.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.lina_matrix_time_spmv.restype = C.c_double
    L.lina_matrix_spmv_bytes.argtypes = [C.c_void_p, C.c_int]
    L.lina_matrix_spmv_bytes.restype = C.c_double
    L.lina_pcg_operator.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                    C.POINTER(Stats)]
    L.lina_elasticity_adjoint.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip, C.c_double, _dp, C.c_int, _dp,
                                          _dp, C.c_double, C.c_int, _dp, _dp, C.POINTER(Stats)]
    L.lina_rn_params_default.argtypes = [C.POINTER(RNParams), C.c_int, C.c_int]
    L.lina_rn_create.argtypes = [C.c_int, _ip, _ip, _dp, _dp, C.c_int, _ip, _dp, _dp, C.POINTER(RNParams)]
    L.lina_rn_create.restype = C.c_void_p
    L.lina_rn_destroy.argtypes = [C.c_void_p]
    L.lina_rn_solve.argtypes = [C.c_void_p, _dp, C.POINTER(RNStats)]
    L.lina_rn_set_trust_region.argtypes = [C.c_void_p, C.c_double]
    L.lina_rn_scaled_hessian.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _ip, _ip, _dp]
    L.lina_batch_create.argtypes = [C.c_int, _ip, _ip, _ip, _dp, C.POINTER(Params)]
    L.lina_batch_create.restype = C.c_void_p
    L.lina_batch_destroy.argtypes = [C.c_void_p]
    L.lina_batch_pcg.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_int, _dp, _dp, C.POINTER(BatchStats)]
    L.lina_batch_pcg_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int, _dp, _dp,
                                        C.POINTER(BatchStats)]
    L.lina_batch_tuned.argtypes = [C.c_void_p, _ip, C.c_char_p, _ip, _ip, C.c_int]
    L.lina_amg_recompute_device.argtypes = [C.c_void_p, C.c_void_p]
    L.lina_amg_coarse_rows.argtypes = [C.c_void_p, _ip, _ip]
    L.lina_assembly_create.argtypes = [C.c_int, _ip, _ip, _ip, _dp, _ip, _ip]
    L.lina_assembly_create.restype = C.c_void_p
    L.lina_assembly_destroy.argtypes = [C.c_void_p]
    L.lina_assembly_nnz.argtypes = [C.c_void_p]
    L.lina_assembly_terms.argtypes = [C.c_void_p]
    L.lina_assembly_terms.restype = C.c_long
    L.lina_assembly_set_dirichlet.argtypes = [C.c_void_p, C.c_void_p]
    L.lina_assembly_values.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.lina_dirichlet_symmetric.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        raise LinaError(f"liblina_hip error {rc}: {lib().lina_last_error().decode()}")
    return rc


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def device_count() -> int:
    return lib().lina_device_count()


def params(kind: str = "reference", dof: int = 1, depth: int = 4, **kw) -> Params:
    p = Params()
    if kind == "reference":
        lib().lina_params_reference(C.byref(p), dof, depth)
    elif kind == "vcycle":
        lib().lina_params_vcycle(C.byref(p), dof, depth)
    else:
        raise ValueError(kind)
    for k, v in kw.items():
        if k in ("smoother", "coarse_smoother") and isinstance(v, str):
            v = SMOOTHER[v]
        if k == "precision" and isinstance(v, str):
            v = PREC[v]
        if k == "cycle" and isinstance(v, str):
            v = {"v": 0, "full": 2}[v]
        if k not in (f for f, _ in Params._fields_):
            raise KeyError(f"unknown AMG parameter {k!r}")
        setattr(p, k, v)
    return p


class AMG:
    """A device-resident AMG hierarchy (lina_solver handle)."""

    def __init__(self, rowptr, col, val, prm: Params, upload: bool = True):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        self.prm = prm
        h = lib().lina_amg_create(self.n, _i(self.rowptr), _i(self.col), _d(self.val), C.byref(prm),
                                  1 if upload else 0)
        if not h:
            raise LinaError(f"lina_amg_create failed: {lib().lina_last_error().decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lina_amg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, prm: Params):
        _check(lib().lina_amg_set_params(self.h, C.byref(prm)))
        self.prm = prm

    @property
    def depth(self) -> int:
        return lib().lina_amg_depth(self.h)

    def coarse_rows(self):
        """(dense, trivial) rows of the coarsest level's direct solve"""
        d, t = C.c_int(), C.c_int()
        _check(lib().lina_amg_coarse_rows(self.h, C.byref(d), C.byref(t)))
        return d.value, t.value

    def level_sizes(self, l: int) -> dict:
        """rows of level l and of the next coarser one, and the nonzeros of
        its A, P and R (P, R: 0 on the coarsest level)"""
        fine, coarse, na, np_, nr = (C.c_int() for _ in range(5))
        _check(lib().lina_amg_level_info(self.h, l, C.byref(fine), C.byref(coarse), C.byref(na), C.byref(np_),
                                         C.byref(nr)))
        return {"fine": fine.value, "coarse": coarse.value, "nnz_A": na.value, "nnz_P": np_.value,
                "nnz_R": nr.value}

    def level(self, l: int) -> dict:
        fine, coarse, na, np_, nr = (C.c_int() for _ in range(5))
        _check(lib().lina_amg_level_info(self.h, l, C.byref(fine), C.byref(coarse), C.byref(na), C.byref(np_),
                                         C.byref(nr)))
        out = {"fine": fine.value, "coarse": coarse.value}
        for which, nnz, rows in (("A", na.value, fine.value), ("P", np_.value, fine.value),
                                 ("R", nr.value, coarse.value)):
            if which != "A" and l == 0:
                continue
            rp = np.zeros(rows + 1, np.int32)
            ci = np.zeros(max(nnz, 1), np.int32)
            v = np.zeros(max(nnz, 1), np.float64)
            _check(lib().lina_amg_level_copy(self.h, l, which.encode(), _i(rp), _i(ci), _d(v)))
            out[which] = (rp, ci[:nnz], v[:nnz])
        if l > 0:
            lab = np.zeros(fine.value, np.int32)
            _check(lib().lina_amg_labels(self.h, l, _i(lab)))
            out["labels"] = lab
            col = np.zeros(fine.value, np.int32)
            out["ncolours"] = _check(lib().lina_amg_colours(self.h, l, _i(col)))
            out["colours"] = col
        return out

    def recompute_device(self, d_val: int):
        """AMG_recompute_coarse_matrices from finest values in device memory
        (a DeviceVector's ptr)"""
        _check(lib().lina_amg_recompute_device(self.h, C.c_void_p(d_val)))

    def signs(self):
        """the finest rows' signs applied by make_positive_diag (+1 / -1)"""
        out = np.zeros(self.n)
        _check(lib().lina_amg_signs(self.h, _d(out)))
        return out

    def recompute(self, val):
        v = np.ascontiguousarray(val, dtype=np.float64)
        _check(lib().lina_amg_recompute(self.h, _d(v)))

    def pcg(self, b, x0=None, tol=1e-10, max_iter=1000, stop="reference"):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        st = Stats()
        it = _check(lib().lina_pcg(self.h, _d(x), _d(B), tol, max_iter, STOP[stop], C.byref(st)))
        d = st.as_dict()
        d["iters"] = it
        d["x"] = x
        d["b_out"] = B
        return d

    def pcg_device(self, d_x: int, d_b: int, tol=1e-10, max_iter=1000, stop="reference"):
        st = Stats()
        it = _check(lib().lina_pcg_device(self.h, C.c_void_p(d_x), C.c_void_p(d_b), tol, max_iter, STOP[stop],
                                          C.byref(st)))
        d = st.as_dict()
        d["iters"] = it
        return d

    def pcg_operator(self, op: "Matrix", b, x0=None, tol=1e-10, max_iter=1000, stop="relres"):
        """op x = b by pcg with this hierarchy as the preconditioner"""
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        st = Stats()
        it = _check(lib().lina_pcg_operator(self.h, op.h, _d(x), _d(B), tol, max_iter, STOP[stop], C.byref(st)))
        d = st.as_dict()
        d["iters"] = it
        d["x"] = x
        return d

    def elasticity_adjoint(self, xy, tri, nu, u, dLdu, elem_scale=None, blocked=True, tol=1e-12, max_iter=5000):
        """lambda and dL/dtheta per element (lina_elasticity_adjoint)"""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
        u = np.ascontiguousarray(u, dtype=np.float64)
        dLdu = np.ascontiguousarray(dLdu, dtype=np.float64)
        lam = np.zeros(self.n)
        grad = np.zeros(len(tri))
        sc = None if elem_scale is None else _d(np.ascontiguousarray(elem_scale, dtype=np.float64))
        st = Stats()
        it = _check(lib().lina_elasticity_adjoint(self.h, len(xy), _d(xy), len(tri), _i(tri), nu, sc,
                                                  1 if blocked else 0, _d(u), _d(dLdu), tol, max_iter, _d(lam),
                                                  _d(grad), C.byref(st)))
        return {"iters": it, "lambda": lam, "grad": grad, "stats": st.as_dict()}

    def amg_solve(self, f, x0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        x = np.array(x0, dtype=np.float64)
        k = _check(lib().lina_amg_solve(self.h, _d(f), _d(x)))
        return k, x

    def spmv(self, x, level=None, precision="fp64"):
        l = self.depth - 1 if level is None else level
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(lib().lina_spmv(self.h, l, PREC[precision], _d(x), _d(y)))
        return y

    def time_spmv(self, level=None, precision="fp64", reps=50) -> float:
        l = self.depth - 1 if level is None else level
        ms = lib().lina_time_spmv(self.h, l, PREC[precision], reps)
        if ms < 0:
            raise LinaError(lib().lina_last_error().decode())
        return ms

    def tuned(self):
        """[(level, kind, lanes FP64, lanes FP32)] chosen by the autotuner"""
        cap = 256
        lv = np.zeros(cap, np.int32)
        wh = C.create_string_buffer(cap)
        l64 = np.zeros(cap, np.int32)
        l32 = np.zeros(cap, np.int32)
        m = _check(lib().lina_amg_tuned(self.h, _i(lv), wh, _i(l64), _i(l32), cap))
        return [(int(lv[i]), wh.raw[i:i + 1].decode(), int(l64[i]), int(l32[i])) for i in range(min(m, cap))]

    def kernel_timing(self, on: bool = True):
        _check(lib().lina_amg_kernel_timing(self.h, 1 if on else 0))

    def spmv_bytes(self, level=None, precision="fp64") -> float:
        l = self.depth - 1 if level is None else level
        return lib().lina_spmv_bytes(self.h, l, PREC[precision])


class Matrix:
    """Device-resident CSR matrix (lina_matrix handle): y = A x."""

    def __init__(self, rowptr, col, val):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        h = lib().lina_matrix_create(self.n, _i(self.rowptr), _i(self.col), _d(self.val))
        if not h:
            raise LinaError(f"lina_matrix_create failed: {lib().lina_last_error().decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lina_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def power(self, v, max_iter=1000, eps=1e-6, eig0=0.0):
        """power_method from the start vector v: (iterations or 0 at max_iter,
        eigenvalue estimate, last normalised iterate)"""
        v = np.array(v, dtype=np.float64)
        eig = np.array([eig0], dtype=np.float64)
        it = lib().lina_matrix_power(self.h, _d(v), max_iter, eps, _d(eig))
        _check(it)
        return it, float(eig[0]), v

    def spmv(self, x, precision="fp64"):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(lib().lina_matrix_spmv(self.h, PREC[precision], _d(x), _d(y)))
        return y

    def spmv_device(self, d_x: int, d_y: int):
        _check(lib().lina_matrix_spmv_device(self.h, C.c_void_p(d_x), C.c_void_p(d_y)))

    def time_spmv(self, precision="fp64", reps=50) -> float:
        ms = lib().lina_matrix_time_spmv(self.h, PREC[precision], reps)
        if ms < 0:
            raise LinaError(lib().lina_last_error().decode())
        return ms

    def spmv_bytes(self, precision="fp64") -> float:
        return lib().lina_matrix_spmv_bytes(self.h, PREC[precision])


class DeviceVector:
    """An FP64 vector in device memory allocated by liblina_hip."""

    def __init__(self, n: int):
        self.n = n
        self.ptr = lib().lina_dev_alloc(n)
        if not self.ptr:
            raise LinaError(lib().lina_last_error().decode())

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        _check(lib().lina_dev_copy(self.ptr, a.ctypes.data, self.n, 0))

    def download(self):
        out = np.empty(self.n)
        _check(lib().lina_dev_copy(out.ctypes.data, self.ptr, self.n, 1))
        return out

    def copy_from(self, other: "DeviceVector"):
        _check(lib().lina_dev_copy(self.ptr, other.ptr, self.n, 2))

    def fill(self, v: float):
        _check(lib().lina_dev_fill(self.ptr, self.n, v))

    def free(self):
        if self.ptr:
            lib().lina_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def _bind_krylov(L):
    if getattr(L, "_krylov_bound", False):
        return
    L.lina_ilu_create.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int]
    L.lina_ilu_create.restype = C.c_void_p
    L.lina_ilu_create_ordered.argtypes = [C.c_int, _ip, _ip, _dp, C.c_int, C.c_int]
    L.lina_ilu_create_ordered.restype = C.c_void_p
    L.lina_ilu_permutation.argtypes = [C.c_void_p, _ip]
    L.lina_ilu_apply_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.lina_ilu_destroy.argtypes = [C.c_void_p]
    L.lina_ilu_ok.argtypes = [C.c_void_p]
    L.lina_ilu_refactor.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp]
    L.lina_ilu_factor_nnz.argtypes = [C.c_void_p, C.c_char]
    L.lina_ilu_factor_copy.argtypes = [C.c_void_p, C.c_char, _ip, _ip, _dp]
    L.lina_ilu_apply.argtypes = [C.c_void_p, _dp, _dp]
    L.lina_ilu_pcg.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_int, _dp]
    L.lina_ilu_gmres.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int, _dp]
    L.lina_bicgstab.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _dp]
    L._krylov_bound = True


ILU_ORDER = {"natural": 0, "multicolour": 1}


class ILU:
    """ILU(0) factors of A with device triangular solves (lina_ilu handle)."""

    def __init__(self, rowptr, col, val, upload: bool = True, ordering: str = "natural"):
        """ordering "multicolour": the factors of P A P^T (P from a greedy
        colouring of A's pattern), vectors in A's order"""
        L = lib()
        _bind_krylov(L)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        h = L.lina_ilu_create_ordered(self.n, _i(self.rowptr), _i(self.col), _d(self.val), ILU_ORDER[ordering],
                                      1 if upload else 0)
        if not h:
            raise LinaError(f"lina_ilu_create failed: {L.lina_last_error().decode()}")
        self.h = h

    def permutation(self):
        """perm[r] = the row of A at position r of the factors"""
        p = np.zeros(self.n, np.int32)
        _check(lib().lina_ilu_permutation(self.h, _i(p)))
        return p

    def close(self):
        if getattr(self, "h", None):
            lib().lina_ilu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def ok(self) -> bool:
        return bool(lib().lina_ilu_ok(self.h))

    def refactor(self, rowptr, col, val):
        """lina_ilu_refactor: ILU(0) of a new matrix on the device (same factors as the host)."""
        rp, ci, v = (np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(col, np.int32),
                     np.ascontiguousarray(val, np.float64))
        _check(lib().lina_ilu_refactor(self.h, len(rp) - 1, _i(rp), _i(ci), _d(v)))
        return self.ok

    def factor(self, which: str):
        nnz = lib().lina_ilu_factor_nnz(self.h, which.encode())
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        lib().lina_ilu_factor_copy(self.h, which.encode(), _i(rp), _i(ci), _d(v))
        return rp, ci[:nnz], v[:nnz]

    def apply_device(self, d_z: int, d_w: int):
        """w = U^-1 L^-1 z on device vectors (DeviceVector ptrs)"""
        _check(lib().lina_ilu_apply_device(self.h, C.c_void_p(d_z), C.c_void_p(d_w)))

    def apply(self, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        w = np.zeros_like(z)
        _check(lib().lina_ilu_apply(self.h, _d(z), _d(w)))
        return w

    def pcg(self, b, x0=None, tol=1e-10, max_iter=1000, stop="reference"):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        B = np.array(b, dtype=np.float64)
        res = C.c_double()
        it = _check(lib().lina_ilu_pcg(self.h, _d(x), _d(B), tol, max_iter, STOP[stop], C.byref(res)))
        return {"iters": it, "x": x, "b_out": B, "res": res.value}

    def gmres(self, b, x0=None, m=30, tol=1e-8, max_iter=1000):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        eta = C.c_double()
        it = _check(lib().lina_ilu_gmres(self.h, _d(x), _d(b), m, tol, max_iter, C.byref(eta)))
        return {"iters": it, "x": x, "eta": eta.value}


def bicgstab(amg: "AMG", b, x0=None, eps=1e-8, max_iter=1000):
    L = lib()
    _bind_krylov(L)
    x = np.zeros(amg.n) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    res = C.c_double()
    it = _check(L.lina_bicgstab(amg.h, _d(x), _d(b), eps, max_iter, C.byref(res)))
    return {"iters": it, "x": x, "res": res.value}


def rn_params(solver="bicgstab", dof=1, depth=6, **kw) -> RNParams:
    p = RNParams()
    lib().lina_rn_params_default(C.byref(p), dof, depth)
    p.solver = RN_SOLVER[solver] if isinstance(solver, str) else solver
    for k, v in kw.items():
        if k == "amg":
            for kk, vv in v.items():
                setattr(p.amg, kk, vv)
        elif k == "ilu_ordering" and isinstance(v, str):
            p.ilu_ordering = ILU_ORDER[v]
        else:
            setattr(p, k, v)
    return p


class ReflectiveNewton:
    """reflective_newton on q(x) = x.H.x/2 + x.b with box bounds (lina_rn handle).
    idx: bounded indices (default all), lower/upper per bounded index (NaN:
    unbounded side)."""

    def __init__(self, rowptr, col, val, b, lower, upper, prm: RNParams, idx=None):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        self.idx = np.arange(self.n, dtype=np.int32) if idx is None else np.ascontiguousarray(idx, dtype=np.int32)
        self.lower = np.ascontiguousarray(lower, dtype=np.float64)
        self.upper = np.ascontiguousarray(upper, dtype=np.float64)
        h = lib().lina_rn_create(self.n, _i(self.rowptr), _i(self.col), _d(self.val), _d(self.b), len(self.idx),
                                 _i(self.idx), _d(self.lower), _d(self.upper), C.byref(prm))
        if not h:
            raise LinaError(f"lina_rn_create failed: {lib().lina_last_error().decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lina_rn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def solve(self, x0):
        x = np.array(x0, dtype=np.float64)
        st = RNStats()
        it = _check(lib().lina_rn_solve(self.h, _d(x), C.byref(st)))
        d = st.as_dict()
        d["iters"] = it
        d["x"] = x
        return d

    def set_trust_region(self, delta: float):
        _check(lib().lina_rn_set_trust_region(self.h, delta))


def rn_scaled_hessian(rowptr, col, val, D, G, J):
    """M = D H D + diag(|G| J) as RN_get_M_D assembles it (host)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    n = len(rp) - 1
    cap = int(rp[-1]) + n
    mrp = np.zeros(n + 1, np.int32)
    mci = np.zeros(cap, np.int32)
    mv = np.zeros(cap)
    nnz = _check(lib().lina_rn_scaled_hessian(n, _i(rp), _i(np.ascontiguousarray(col, dtype=np.int32)),
                                              _d(np.ascontiguousarray(val, dtype=np.float64)),
                                              _d(np.ascontiguousarray(D, dtype=np.float64)),
                                              _d(np.ascontiguousarray(G, dtype=np.float64)),
                                              _d(np.ascontiguousarray(J, dtype=np.float64)), _i(mrp), _i(mci), _d(mv)))
    return mrp, mci[:nnz], mv[:nnz]


class Batch:
    """Batched AMG-PCG over independent systems (lina_batch handle): one
    block-diagonal union (row_off, rowptr, col, val) as from
    problems.block_diagonal."""

    def __init__(self, row_off, rowptr, col, val, prm: Params):
        self.row_off = np.ascontiguousarray(row_off, dtype=np.int32)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.nsys = len(self.row_off) - 1
        self.n = int(self.row_off[-1])
        h = lib().lina_batch_create(self.nsys, _i(self.row_off), _i(self.rowptr), _i(self.col), _d(self.val),
                                    C.byref(prm))
        if not h:
            raise LinaError(f"lina_batch_create failed: {lib().lina_last_error().decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lina_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def pcg(self, b, x0=None, tol=1e-8, max_iter=5000, graph_iters=0):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        it = np.zeros(self.nsys)
        rr = np.zeros(self.nsys)
        st = BatchStats()
        launched = _check(lib().lina_batch_pcg(self.h, _d(x), _d(b), tol, max_iter, graph_iters, _d(it), _d(rr),
                                               C.byref(st)))
        return {"x": x, "iters": it.astype(int), "relres": rr, "launched": launched, **st.as_dict()}

    def pcg_device(self, d_x: int, d_b: int, tol=1e-8, max_iter=5000, graph_iters=0):
        it = np.zeros(self.nsys)
        rr = np.zeros(self.nsys)
        st = BatchStats()
        launched = _check(lib().lina_batch_pcg_device(self.h, C.c_void_p(d_x), C.c_void_p(d_b), tol, max_iter,
                                                      graph_iters, _d(it), _d(rr), C.byref(st)))
        return {"iters": it.astype(int), "relres": rr, "launched": launched, **st.as_dict()}

    def tuned(self):
        cap = 256
        lv = np.zeros(cap, np.int32)
        wh = C.create_string_buffer(cap)
        l64 = np.zeros(cap, np.int32)
        l32 = np.zeros(cap, np.int32)
        m = _check(lib().lina_batch_tuned(self.h, _i(lv), wh, _i(l64), _i(l32), cap))
        return [(int(lv[i]), wh.raw[i:i + 1].decode(), int(l64[i]), int(l32[i])) for i in range(min(m, cap))]


class Assembly:
    """The degraded stiffness's values on the device from a third-order
    tensor and nodal rank-4 moduli (lina_assembly handle): the contraction of
    sparse3D_times_rank4_2D_right_use_map, optionally with the symmetric
    Dirichlet treatment."""

    def __init__(self, n, bptr, bi1, bi2, bval, rowptr, col):
        self.n = int(n)
        self._keep = [np.ascontiguousarray(a, dtype=t) for a, t in
                      ((bptr, np.int32), (bi1, np.int32), (bi2, np.int32), (bval, np.float64),
                       (rowptr, np.int32), (col, np.int32))]
        bp, b1, b2, bv, rp, ci = self._keep
        self.h = lib().lina_assembly_create(self.n, _i(bp), _i(b1), _i(b2), _d(bv), _i(rp), _i(ci))
        if not self.h:
            raise LinaError(lib().lina_last_error().decode())
        self.nnz = _check(lib().lina_assembly_nnz(self.h))
        self.terms = int(lib().lina_assembly_terms(self.h))

    def set_dirichlet(self, flags):
        if flags is None:
            _check(lib().lina_assembly_set_dirichlet(self.h, None))
            return
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.size == 2 * self.n
        _check(lib().lina_assembly_set_dirichlet(self.h, f.ctypes.data))

    def values(self, d_C: int, d_val: int, d_udir: int = 0, d_b: int = 0):
        """device pointers: moduli (16 n), values out (nnz), Dirichlet values
        (2 n) and right-hand side (2 n, in/out) when constrained DOFs are set"""
        _check(lib().lina_assembly_values(self.h, C.c_void_p(d_C), C.c_void_p(d_val),
                                          C.c_void_p(d_udir or None), C.c_void_p(d_b or None)))

    def close(self):
        if getattr(self, "h", None):
            lib().lina_assembly_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def dirichlet_symmetric(n: int, d_rowptr: int, d_col: int, d_flags: int, d_u: int, d_val: int, d_b: int = 0):
    """lina_dirichlet_symmetric on device arrays (pointers)"""
    _check(lib().lina_dirichlet_symmetric(n, C.c_void_p(d_rowptr), C.c_void_p(d_col), C.c_void_p(d_flags),
                                          C.c_void_p(d_u), C.c_void_p(d_val), C.c_void_p(d_b or None)))
