"""The staggered loop's elastic-step update on the device: the degraded
stiffness re-assembled from the strain-gradient tensor and nodal moduli
(sparse3D_times_rank4_2D_right_use_map, FEM2D.c:8504-8600), the symmetric
Dirichlet treatment, AMG_recompute_coarse_matrices from the device values,
and the AMG-PCG solve -- each against the oracle or a direct solve."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from fem_amd import problems as P
from tests.test_assembly_cpu import damage_scale

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def orc():
    from oracle.bind import Oracle
    o = Oracle()
    o.L.orc_rank4_assembly.argtypes = [C.c_int] + [C.c_void_p] * 8
    o.L.orc_dirichlet_symmetric.argtypes = [C.c_int] + [C.c_void_p] * 6
    return o


def _case(m):
    mesh = P.Mesh(m, m)
    n = mesh.n_nodes
    bptr, bi1, bi2, bv = P.strain_gradient_tensor(mesh)
    rp, ci = P.elastic_pattern(mesh)
    bot, top = mesh.boundary("bottom"), mesh.boundary("top")
    dofs = np.concatenate([bot, bot + n, top + n])
    flags = np.zeros(2 * n, np.uint8)
    flags[dofs] = 1
    u = np.zeros(2 * n)
    u[top + n] = 1e-3
    return mesh, (bptr, bi1, bi2, bv), (rp, ci), flags, u


def _oracle(orc, n, B, pat, Cm, flags=None, u=None, f=None):
    bptr, bi1, bi2, bv = B
    rp, ci = pat
    val = np.zeros(rp[-1])
    p = lambda a: a.ctypes.data
    assert orc.L.orc_rank4_assembly(n, p(bptr), p(bi1), p(bi2), p(bv), p(Cm), p(rp), p(ci), p(val)) == 0
    b = None
    if flags is not None:
        b = f.copy()
        orc.L.orc_dirichlet_symmetric(2 * n, p(rp), p(ci), p(flags), p(u), p(val), p(b))
    return val, b


@pytest.mark.parametrize("m", [17, 65])
def test_assembly_matches_oracle(gpu, orc, m):
    """values bit for bit with the oracle's ordered accumulation, without and
    with the symmetric Dirichlet treatment (matrix and right-hand side)"""
    mesh, B, pat, flags, u = _case(m)
    n = mesh.n_nodes
    say("assembly", m, "nodes", n, "triples", int(B[0][-1]))
    asm = gpu.Assembly(n, *B, *pat)
    assert asm.nnz == pat[0][-1] and asm.terms == 4 * int(B[0][-1])
    Cm = P.nodal_moduli(n, damage_scale(mesh))
    dC, dv = gpu.DeviceVector(16 * n), gpu.DeviceVector(asm.nnz)
    dC.upload(Cm)
    asm.values(dC.ptr, dv.ptr)
    ref, _ = _oracle(orc, n, B, pat, Cm)
    assert np.array_equal(dv.download(), ref)
    f = np.sin(np.arange(2 * n) * 0.11)
    asm.set_dirichlet(flags)
    du, db = gpu.DeviceVector(2 * n), gpu.DeviceVector(2 * n)
    du.upload(u)
    db.upload(f)
    asm.values(dC.ptr, dv.ptr, du.ptr, db.ptr)
    ref, rb = _oracle(orc, n, B, pat, Cm, flags, u, f)
    assert np.array_equal(dv.download(), ref) and np.array_equal(db.download(), rb)
    asm.close()


def test_dirichlet_symmetric_entry(gpu, orc):
    """lina_dirichlet_symmetric on device CSR arrays equals the oracle"""
    mesh, B, pat, flags, u = _case(33)
    n2 = 2 * mesh.n_nodes
    K = P.plane_strain_stiffness(mesh, 210e3 * np.linspace(0.2, 1.0, len(mesh.tri)), 0.3)
    K.sort_indices()
    rp, ci = K.indptr.astype(np.int32), K.indices.astype(np.int32)
    f = np.cos(np.arange(n2) * 0.3)

    def dev_bytes(a):
        by = np.ascontiguousarray(a).tobytes()
        raw = np.frombuffer(by + b"\0" * (8 + (-len(by)) % 8), dtype=np.float64)
        d = gpu.DeviceVector(len(raw))
        d.upload(raw)
        return d
    drp, dci, dfl = dev_bytes(rp), dev_bytes(ci), dev_bytes(flags)
    du, dv, db = gpu.DeviceVector(n2), gpu.DeviceVector(len(K.data)), gpu.DeviceVector(n2)
    du.upload(u)
    dv.upload(K.data)
    db.upload(f)
    gpu.dirichlet_symmetric(n2, drp.ptr, dci.ptr, dfl.ptr, du.ptr, dv.ptr, db.ptr)
    v, b = K.data.copy(), f.copy()
    p = lambda a: a.ctypes.data
    orc.L.orc_dirichlet_symmetric(n2, p(rp), p(ci), p(flags), p(u), p(v), p(b))
    assert np.array_equal(dv.download(), v) and np.array_equal(db.download(), b)


def test_elastic_step_pipeline_on_device(gpu, orc, capfd, monkeypatch):
    """Three damage states of one mesh: stiffness and right-hand side
    assembled on the device, the hierarchy recomputed from the device
    values (every level equal, bit for bit, to a recompute from the same
    values handed over from the host), pcg on the device; each solution
    within 1e-8 of scipy's direct solve of the oracle-assembled system."""
    monkeypatch.setenv("LINA_RN_TIMING", "1")
    m = 65
    mesh, B, pat, flags, u = _case(m)
    n = mesh.n_nodes
    asm = gpu.Assembly(n, *B, *pat)
    asm.set_dirichlet(flags)
    dC, dv, du, db = (gpu.DeviceVector(k) for k in (16 * n, asm.nnz, 2 * n, 2 * n))
    dx = gpu.DeviceVector(2 * n)
    du.upload(u)
    base = damage_scale(mesh)
    amg = amg_host = None
    for state in range(3):
        sc = np.clip(base * (1.0 - 0.15 * state) + 0.05 * state, 1e-6, None)
        Cm = P.nodal_moduli(n, sc)
        dC.upload(Cm)
        db.fill(0.0)
        asm.values(dC.ptr, dv.ptr, du.ptr, db.ptr)
        val, b = _oracle(orc, n, B, pat, Cm, flags, u, np.zeros(2 * n))
        assert np.array_equal(dv.download(), val) and np.array_equal(db.download(), b)
        if amg is None:
            prm = gpu.params("vcycle", 2, 6)
            amg = gpu.AMG(pat[0], pat[1], val, prm)
            amg_host = gpu.AMG(pat[0], pat[1], val, prm)
        else:
            capfd.readouterr()
            amg.recompute_device(dv.ptr)
            err = capfd.readouterr().err
            say("  state", state, [l for l in err.splitlines() if "recompute" in l])
            assert "recompute on the device from device values: done" in err
            amg_host.recompute(val)
            for l in range(amg.depth):
                ga, ha = amg.level(l)["A"], amg_host.level(l)["A"]
                assert all(np.array_equal(x, y) for x, y in zip(ga, ha)), (state, l)
        dx.fill(0.0)
        st = amg.pcg_device(dx.ptr, db.ptr, tol=1e-11, max_iter=3000, stop="relres")
        x = dx.download()
        A = sp.csr_matrix((val, pat[1], pat[0]), shape=(2 * n, 2 * n))
        xd = spla.spsolve(A.tocsc(), b)
        e = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        say("  state", state, "pcg iters", st["iters"], "err vs direct", e)
        assert e < 1e-8
    amg.close()
    amg_host.close()
    asm.close()
