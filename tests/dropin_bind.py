"""ctypes access to LinA-named entry points (test helper): the drop-in
liblinalg_hip.so and the compiled reference (oracle/_ref) export the same
symbols with the same types, so one wrapper drives both."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN_SO = os.path.join(ROOT, "fem_amd", "liblinalg_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SparseMatrix(C.Structure):
    _fields_ = [("Values", C.POINTER(_dp)), ("Indices", C.POINTER(_ip)), ("Len", _ip), ("size", C.c_int)]


_SM = C.POINTER(SparseMatrix)


def dptr(a):
    return a.ctypes.data_as(_dp)


class LinA:
    """LinA API of one library (drop-in or reference)."""

    def __init__(self, lib):
        L = lib if isinstance(lib, C.CDLL) else C.CDLL(lib)
        self.L = L
        sig = {
            "sparse_zero": ([C.c_int], _SM),
            "write_sparse": ([_SM, C.c_char_p, C.c_char_p], C.c_int),
            "matrix_norm": ([_SM], C.c_double),
            "power_method": ([_SM, _dp, _dp, C.c_int, C.c_double, _SM, C.c_double], C.c_int),
            "get_most_negative_eigenvector": ([_SM, _dp, _dp, C.c_int, C.c_double], C.c_int),
            "get_neg_curvature_direction": ([_SM, _dp, _dp, C.c_int, C.c_double, C.c_double], C.c_int),
            "read_sparse": ([C.c_char_p, C.c_char_p], _SM),
            "read_sparse_textfile": ([C.c_char_p, C.c_char_p], _SM),
            "write_sparse_textfile": ([C.c_char_p, C.c_char_p, _SM], None),
            "print_to_matrix_market_format": ([_SM, C.c_char_p], None),
            "insert_sparse": ([_SM, C.c_double, C.c_int, C.c_int], C.c_int),
            "set_sparse": ([_SM, C.c_double, C.c_int, C.c_int], None),
            "get_matrix_element": ([_SM, C.c_int, C.c_int], C.c_double),
            "delete_zeros": ([_SM], None),
            "clone": ([_SM], _SM),
            "Free_sparse": ([_SM], None),
            "sparse_mult": ([_SM, _dp], _dp),
            "linear_map": ([_dp, C.c_double, _SM, _dp], None),
            "sparse_multiplication": ([_SM, _dp], None),
            "sparse_bilinear": ([_dp, _SM, _dp], C.c_double),
            "matrix_residuum": ([_SM, _dp, _dp], C.c_double),
            "scalar": ([_dp, _dp, C.c_int], C.c_double),
            "euklid_norm": ([_dp, C.c_int], C.c_double),
            "incomplete_LU_factorization": ([_SM, _SM, _SM], C.c_int),
            "L_triang_invert": ([_SM, _dp], None),
            "U_triang_invert": ([_SM, _dp], None),
            "pcg": ([_dp, _dp, C.c_int, C.c_double, C.c_int], C.c_int),
            "get_pcg_residuum": ([], C.c_double),
            "PCG_default": ([_SM, _SM, _SM], None),
            "PCG_set_default_precon": ([], None),
            "PCG_set_AMG_preconditioner": ([_SM, C.c_int, C.c_int], None),
            "BI_set_AMG_preconditioner": ([_SM, C.c_int, C.c_int], None),
            "BI_set_eps": ([C.c_double], None),
            "BI_set_max_iter": ([C.c_int], None),
            "BI_print_info": ([C.c_int], None),
            "BiCGStab_solve": ([_dp, _dp, C.c_int], C.c_int),
            "RN_set_default_matrix": ([_SM], None),
            "AMG_print_info": ([C.c_int], None),
            "AMG_set_SOR_relaxation_coeff": ([C.c_double], None),
            "AMG_set_eps": ([C.c_double], None),
            "AMG_set_max_iter": ([C.c_int], None),
            "AMG_setup": ([_SM, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                           C.c_void_p], C.c_void_p),
            "AMG_solve": ([C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p], C.c_int),
            "AMG_free_data": ([C.c_void_p], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        self.PCG_mult = L.PCG_mult
        self.RN_default_mat_mult = L.RN_default_mat_mult
        self.PCG_set_mat_mult = L.PCG_set_mat_mult
        self.PCG_set_mat_mult.argtypes = [C.c_void_p]
        self.PCG_set_precon = L.PCG_set_precon
        self.PCG_set_precon.argtypes = [C.c_void_p]
        self.BI_set_mat_mult = L.BI_set_mat_mult
        self.BI_set_mat_mult.argtypes = [C.c_void_p]

    def __getattr__(self, name):
        return getattr(self.L, name)

    def from_csr(self, rowptr, col, val, order=None):
        """row-list matrix built by insert_sparse on sparse_zero (entries in
        `order`, default row-major)"""
        n = len(rowptr) - 1
        A = self.L.sparse_zero(n)
        rows = np.repeat(np.arange(n), np.diff(rowptr))
        idx = range(len(col)) if order is None else order
        for k in idx:
            self.L.insert_sparse(A, float(val[k]), int(rows[k]), int(col[k]))
        return A

    @staticmethod
    def to_lists(A):
        a = A.contents
        out = []
        for i in range(a.size):
            m = a.Len[i]
            out.append(([a.Indices[i][j] for j in range(m)], [a.Values[i][j] for j in range(m)]))
        return out

    def free(self, A):
        self.L.Free_sparse(A)

    def vec(self, p, n):
        return np.ctypeslib.as_array(p, shape=(n,)).copy()
