"""The drop-in's replacements of Crack's UMFPACK drivers on the GPU
(lina_spd_solver for multifrontal_solver, Crack/crack.c:5424; and
lina_spd_local_solver for the RN_set_local_solver slot, :9207): systems
with Set_dirichlet_BC's unit rows (:6445-6459), solved with the
symmetric elimination and AMG-PCG, against scipy's direct solve of the
same (row-reset, non-symmetric) systems; a second call with new values of
the same pattern takes the values-only recompute."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from fem_amd import problems as P
from tests.dropin_bind import DROPIN_SO, LinA, SparseMatrix, dptr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dl(gpu):
    lib = LinA(DROPIN_SO)
    L = lib.L
    SM = C.POINTER(SparseMatrix)
    L.lina_spd_solver.argtypes = [SM, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_char_p,
                                  C.POINTER(C.c_double)]
    L.lina_spd_local_solver.argtypes = [SM, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.lina_spd_last_iterations.argtypes = [C.c_int, C.POINTER(C.c_double)]
    L.reset_row.argtypes = [SM, C.c_int]
    yield lib
    L.lina_spd_solver_free()


def row_reset_system(lib, K, f, dofs, vals):
    """Crack's Set_dirichlet_BC on a row-list copy of K: the constrained rows
    reset to the unit row, F = the prescribed value"""
    A = lib.from_csr(K.indptr.astype(np.int32), K.indices.astype(np.int32), K.data)
    b = f.copy()
    for d, v in zip(dofs, vals):
        lib.L.reset_row(A, int(d))
        lib.L.insert_sparse(A, 1.0, int(d), int(d))
        b[d] = v
    R = K.tolil()
    for d in dofs:
        R.rows[d] = [int(d)]
        R.data[d] = [1.0]
    return A, b, R.tocsc()


def test_elastic_step_solver_replaces_umfpack(dl):
    m = 65
    mesh = P.Mesh(m, m)
    n = mesh.n_nodes
    bot, top = mesh.boundary("bottom"), mesh.boundary("top")
    dofs = np.concatenate([bot, bot + n, top + n])
    vals = np.concatenate([np.zeros(2 * len(bot)), np.full(len(top), 1e-3)])
    a, bb = P.notch_segment(P.SEED_NOTCH, 0.3)
    dist = P._dist_to_segment(mesh.centroids(), a, bb)
    x = np.zeros(2 * n)
    t = C.c_double(0.0)
    for step, l0 in enumerate((0.02, 0.04)):          # the damage grows: same pattern, new values
        g = (1.0 - np.exp(-dist / l0)) ** 2 + 1e-6
        K = P.plane_strain_stiffness(mesh, 210e3 * g, 0.3)
        K.sort_indices()
        f = np.zeros(2 * n)
        A, b, R = row_reset_system(dl, K, f, dofs, vals)
        ok = dl.L.lina_spd_solver(A, dptr(x), dptr(b), b"elastic step", C.byref(t))
        rr = C.c_double()
        it = dl.L.lina_spd_last_iterations(0, C.byref(rr))
        xd = spla.spsolve(R, b)
        e = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        say("step", step, "ok", ok, "iters", it, "relres", rr.value, "err vs direct", e, "time", t.value)
        assert ok == 1 and e < 1e-8
        assert np.allclose(x[dofs], vals, rtol=1e-14, atol=1e-18)   # pcg scales b by 1/|b| and back
        dl.free(A)
    assert t.value > 0


def test_local_solver_replaces_umfpack(dl):
    s, _, _ = P.damage_system(65)
    K = s.scipy().tocsr()
    n = s.n
    rng = np.random.default_rng(3)
    dofs = np.sort(rng.choice(n, 40, replace=False))
    vals = rng.uniform(0, 0.2, 40)
    x = np.zeros(n)
    for step in range(2):
        K2 = K.copy()
        K2.data = K.data * (1.0 + 0.1 * step)
        A, b, R = row_reset_system(dl, K2, s.b.copy(), dofs, vals)
        ok = dl.L.lina_spd_local_solver(A, dptr(x), dptr(b))
        xd = spla.spsolve(R, b)
        e = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        say("local step", step, "ok", ok, "iters", dl.L.lina_spd_last_iterations(1, None), "err", e)
        assert ok == 1 and e < 1e-9
        dl.free(A)
