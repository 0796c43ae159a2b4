"""Seeded synthetic FEM systems for the LinA solver path.

The reference assembles its systems with FEM2D (P1 triangles, node-centric,
every mesh edge inserted: ``FEM2D/FEM2D.c:5865-5884,5949-5968``) and hands a
row-list ``sparse_matrix`` to the solvers.  This module builds the same kind
of matrices on a structured triangulated rectangle so that the reference, the
oracle and the HIP path all solve the identical CSR system:

* pattern = every mesh edge + the diagonal (explicit zeros are kept, as an
  assembled FEM matrix keeps them), rows sorted by column;
* Dirichlet rows become identity rows and the matching columns are zeroed
  (symmetric elimination, the lifting moved to the right-hand side), the
  pattern is unchanged;
* vector problems use the reference's blocked ordering ``[u_x(all nodes);
  u_y(all nodes)]`` that ``AMG_init_finest_tupels`` assumes
  (``LinA/Multigrid.c:785-804``).

All randomness comes from a counter-based SplitMix64 stream (never numpy's
generators), so every input is reproducible from the committed seeds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of the SplitMix64 counter stream x_k = mix(seed + (k+1)*golden)."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, lo: float = 0.0, hi: float = 1.0, offset: int = 0) -> np.ndarray:
    u = (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed: int, n: int) -> np.ndarray:
    """Box-Muller on two disjoint halves of the stream."""
    u1 = uniform(seed, n, offset=0)
    u2 = uniform(seed, n, offset=n)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Mesh:
    """Structured triangulation of [0,lx]x[0,ly] with nx*ny nodes.

    Node (r, c) -> r*nx + c (row r along y, column c along x).  Each cell is
    split by its (r,c)-(r+1,c+1) diagonal into two triangles.
    """
    nx: int
    ny: int
    lx: float = 1.0
    ly: float = 1.0
    xy: np.ndarray = field(init=False)
    tri: np.ndarray = field(init=False)

    def __post_init__(self):
        xs = np.linspace(0.0, self.lx, self.nx)
        ys = np.linspace(0.0, self.ly, self.ny)
        X, Y = np.meshgrid(xs, ys)
        self.xy = np.stack([X.ravel(), Y.ravel()], axis=1)
        r, c = np.meshgrid(np.arange(self.ny - 1), np.arange(self.nx - 1), indexing="ij")
        n00 = (r * self.nx + c).ravel()
        n01 = n00 + 1
        n10 = n00 + self.nx
        n11 = n10 + 1
        t1 = np.stack([n00, n01, n11], axis=1)
        t2 = np.stack([n00, n11, n10], axis=1)
        self.tri = np.concatenate([t1, t2], axis=0).astype(np.int64)

    @property
    def n_nodes(self) -> int:
        return self.nx * self.ny

    def boundary(self, side: str) -> np.ndarray:
        nx, ny = self.nx, self.ny
        if side == "bottom":
            return np.arange(nx)
        if side == "top":
            return (ny - 1) * nx + np.arange(nx)
        if side == "left":
            return np.arange(ny) * nx
        if side == "right":
            return np.arange(ny) * nx + nx - 1
        raise ValueError(side)

    def all_boundary(self) -> np.ndarray:
        return np.unique(np.concatenate([self.boundary(s) for s in ("bottom", "top", "left", "right")]))

    def centroids(self) -> np.ndarray:
        return self.xy[self.tri].mean(axis=1)

    def grads(self):
        """Per-element P1 basis gradients (ne,3,2) and areas (ne,)."""
        p = self.xy[self.tri]
        d1 = p[:, 1] - p[:, 0]
        d2 = p[:, 2] - p[:, 0]
        det = d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0]
        area = 0.5 * np.abs(det)
        # gradients of barycentric coordinates
        g = np.empty((len(self.tri), 3, 2))
        g[:, 1, 0] = d2[:, 1] / det
        g[:, 1, 1] = -d2[:, 0] / det
        g[:, 2, 0] = -d1[:, 1] / det
        g[:, 2, 1] = d1[:, 0] / det
        g[:, 0] = -(g[:, 1] + g[:, 2])
        return g, area


@dataclass
class System:
    """CSR system (int32 indices, sorted rows) with right-hand side."""
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    b: np.ndarray
    dof: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.rowptr) - 1

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))


def _assemble(rows, cols, vals, n) -> sp.csr_matrix:
    A = sp.coo_matrix((vals.ravel(), (rows.ravel(), cols.ravel())), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def _scalar_pattern_assembly(mesh: Mesh, ke: np.ndarray) -> sp.csr_matrix:
    """Assemble per-element 3x3 blocks ke (ne,3,3)."""
    t = mesh.tri
    rows = np.repeat(t[:, :, None], 3, axis=2)
    cols = np.repeat(t[:, None, :], 3, axis=1)
    return _assemble(rows, cols, ke, mesh.n_nodes)


def p1_stiffness(mesh: Mesh, coef: np.ndarray | None = None) -> sp.csr_matrix:
    g, area = mesh.grads()
    ke = np.einsum("eik,ejk->eij", g, g) * area[:, None, None]
    if coef is not None:
        ke = ke * coef[:, None, None]
    return _scalar_pattern_assembly(mesh, ke)


def p1_mass(mesh: Mesh, coef: np.ndarray | None = None) -> sp.csr_matrix:
    _, area = mesh.grads()
    base = (np.ones((3, 3)) + np.eye(3)) / 12.0
    ke = base[None] * area[:, None, None]
    if coef is not None:
        ke = ke * coef[:, None, None]
    return _scalar_pattern_assembly(mesh, ke)


def plane_strain_stiffness(mesh: Mesh, E: np.ndarray | float, nu: float) -> sp.csr_matrix:
    """P1 plane-strain elasticity, blocked [ux; uy] ordering, per-element E."""
    g, area = mesh.grads()
    ne = len(mesh.tri)
    E = np.broadcast_to(np.asarray(E, dtype=np.float64), (ne,))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    gx, gy = g[:, :, 0], g[:, :, 1]
    # blocks of the 6x6 element matrix indexed [comp_a, comp_b, i, j]
    kxx = (lam + 2 * mu)[:, None, None] * gx[:, :, None] * gx[:, None, :] + mu[:, None, None] * gy[:, :, None] * gy[:, None, :]
    kyy = (lam + 2 * mu)[:, None, None] * gy[:, :, None] * gy[:, None, :] + mu[:, None, None] * gx[:, :, None] * gx[:, None, :]
    kxy = lam[:, None, None] * gx[:, :, None] * gy[:, None, :] + mu[:, None, None] * gy[:, :, None] * gx[:, None, :]
    kyx = np.transpose(kxy, (0, 2, 1))
    nn = mesh.n_nodes
    t = mesh.tri
    R = np.repeat(t[:, :, None], 3, axis=2)
    C = np.repeat(t[:, None, :], 3, axis=1)
    a = area[:, None, None]
    rows = np.concatenate([R, R, R + nn, R + nn])
    cols = np.concatenate([C, C + nn, C, C + nn])
    vals = np.concatenate([kxx * a, kxy * a, kyx * a, kyy * a])
    return _assemble(rows, cols, vals, 2 * nn)


def apply_dirichlet(A: sp.csr_matrix, b: np.ndarray, dofs: np.ndarray, values: np.ndarray | float = 0.0):
    """Symmetric elimination keeping the pattern: b -= A[:,D] u_D, zero rows and
    columns of D, unit diagonal, b_D = u_D."""
    n = A.shape[0]
    uD = np.zeros(n)
    uD[dofs] = values
    b = b - A @ uD
    isD = np.zeros(n, dtype=bool)
    isD[dofs] = True
    indptr, indices, data = A.indptr, A.indices, A.data.copy()
    row_of = np.repeat(np.arange(n), np.diff(indptr))
    mask = isD[row_of] | isD[indices]
    data[mask] = 0.0
    diag = mask & (row_of == indices) & isD[row_of]
    data[diag] = 1.0
    A2 = sp.csr_matrix((data, indices.copy(), indptr.copy()), shape=A.shape)
    b = b.copy()
    b[dofs] = uD[dofs]
    return A2, b


def strain_gradient_tensor(mesh: Mesh):
    """The third-order tensor of Crack's elastic Hessian for P1 elements
    (Strain_grad, FEM2D.c:6068-6112): row i (node), triples
    (n b + j, n d + k, sum over the elements e holding i, j and k of
    grad_b phi_j grad_d phi_k |e| / 3), b, d in {0, 1}.  Contracted with nodal
    moduli C^i_abcd it gives K_(a j),(c k) = sum_i sum_bd C^i_abcd B^i_(b j),(d k)
    -- with the same C0 at every node, the P1 elasticity stiffness.
    Returns (bptr, bi1, bi2, bval): the triples of a row sorted by (bi1, bi2)."""
    g, area = mesh.grads()
    t = mesh.tri
    n = mesh.n_nodes
    ne = len(t)
    # per element: (i, j, k) over its 3 x 3 x 3 node triples, (b, d) over 2 x 2
    I = np.broadcast_to(t[:, :, None, None, None, None], (ne, 3, 3, 3, 2, 2))
    J = np.broadcast_to(t[:, None, :, None, None, None], (ne, 3, 3, 3, 2, 2))
    K = np.broadcast_to(t[:, None, None, :, None, None], (ne, 3, 3, 3, 2, 2))
    Bd = np.broadcast_to(np.arange(2)[None, None, None, None, :, None], (ne, 3, 3, 3, 2, 2))
    Dd = np.broadcast_to(np.arange(2)[None, None, None, None, None, :], (ne, 3, 3, 3, 2, 2))
    val = (g[:, None, :, None, :, None] * g[:, None, None, :, None, :] *
           (area / 3.0)[:, None, None, None, None, None])
    val = np.broadcast_to(val, (ne, 3, 3, 3, 2, 2))
    row = I.ravel().astype(np.int64)
    c1 = (Bd * n + J).ravel().astype(np.int64)
    c2 = (Dd * n + K).ravel().astype(np.int64)
    key = (row * (2 * n) + c1) * (2 * n) + c2
    uk, inv = np.unique(key, return_inverse=True)
    acc = np.zeros(len(uk))
    np.add.at(acc, inv, val.ravel())
    rows = uk // (4 * n * n)
    rest = uk % (4 * n * n)
    bi1 = (rest // (2 * n)).astype(np.int32)
    bi2 = (rest % (2 * n)).astype(np.int32)
    bptr = np.zeros(n + 1, np.int32)
    np.add.at(bptr, rows + 1, 1)
    return np.cumsum(bptr).astype(np.int32), bi1, bi2, acc


def nodal_moduli(n: int, scale: np.ndarray | float = 1.0, E: float = 210e3, nu: float = 0.3) -> np.ndarray:
    """Plane-strain isotropic moduli per node, C^i_abcd = scale_i (lambda d_ab d_cd +
    mu (d_ac d_bd + d_ad d_bc)), component (a,b,c,d) of node i at
    n (8a + 4b + 2c + d) + i (Crack's Rank4_modulus layout)."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    sc = np.broadcast_to(np.asarray(scale, dtype=np.float64), (n,))
    C = np.empty(16 * n)
    for a in range(2):
        for b in range(2):
            for c in range(2):
                for d in range(2):
                    v = lam * (a == b) * (c == d) + mu * ((a == c) * (b == d) + (a == d) * (b == c))
                    C[n * (8 * a + 4 * b + 2 * c + d): n * (8 * a + 4 * b + 2 * c + d + 1)] = v * sc
    return C


def elastic_pattern(mesh: Mesh):
    """(rowptr, col) of the blocked [ux; uy] P1 elasticity matrix"""
    K = plane_strain_stiffness(mesh, 1.0, 0.3)
    K.sort_indices()
    return K.indptr.astype(np.int32), K.indices.astype(np.int32)


def _to_system(A: sp.csr_matrix, b: np.ndarray, dof: int, meta: dict) -> System:
    A.sort_indices()
    return System(A.indptr.astype(np.int32), A.indices.astype(np.int32),
                  np.ascontiguousarray(A.data, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64),
                  dof, meta)


# ---------------------------------------------------------------- configs

SEED_POISSON = 0x5EED0001
SEED_NOTCH = 0x5EED0002
SEED_DAMAGE_H = 0x5EED0003
SEED_DAMAGE_DOLD = 0x5EED0004
SEED_BATCH = 0x5EED0005
SEED_GRF = 0x5EED0006


def poisson(m: int, seed: int = SEED_POISSON) -> System:
    """P1 Poisson on the unit square, m x m nodes, homogeneous Dirichlet on the
    whole boundary, load b = M f with f_i ~ U[-1,1] per node (seeded)."""
    mesh = Mesh(m, m)
    K = p1_stiffness(mesh)
    M = p1_mass(mesh)
    f = uniform(seed, mesh.n_nodes, -1.0, 1.0)
    b = M @ f
    A, b = apply_dirichlet(K, b, mesh.all_boundary(), 0.0)
    return _to_system(A, b, 1, {"kind": "poisson", "m": m, "seed": seed})


def notch_segment(seed: int, length: float = 0.3, lx: float = 1.0, ly: float = 1.0):
    """Straight notch from the left edge: start height U[0.3,0.7]*ly, angle U[-15,15] deg."""
    u = uniform(seed, 2)
    y0 = (0.3 + 0.4 * u[0]) * ly
    th = math.radians(-15.0 + 30.0 * u[1])
    return np.array([0.0, y0]), np.array([length * math.cos(th), y0 + length * math.sin(th)])


def _dist_to_segment(p: np.ndarray, a: np.ndarray, bpt: np.ndarray) -> np.ndarray:
    ab = bpt - a
    t = np.clip(((p - a) @ ab) / (ab @ ab), 0.0, 1.0)
    proj = a[None, :] + t[:, None] * ab[None, :]
    return np.linalg.norm(p - proj, axis=1)


def degraded_elasticity(m: int, seed: int = SEED_NOTCH, E: float = 210e3, nu: float = 0.3, l0: float = 0.01,
                        top_disp: float = 1e-3, nx: int | None = None, lx: float = 1.0, ly: float = 1.0,
                        E_elem: np.ndarray | None = None, notch: tuple | None = None) -> System:
    """Phase-field degraded plane-strain elasticity: stiffness scaled per element
    by (1-d)^2+1e-6, d = exp(-dist/l0) to a seeded notch; bottom clamped, top
    u_y = top_disp.  Blocked [ux; uy] ordering, dof = 2."""
    mesh = Mesh(nx if nx is not None else m, m, lx, ly)
    a, bb = notch if notch is not None else notch_segment(seed, 0.3 * lx, lx, ly)
    d = np.exp(-_dist_to_segment(mesh.centroids(), a, bb) / l0)
    g = (1.0 - d) ** 2 + 1e-6
    Ee = (E if E_elem is None else E_elem) * g
    K = plane_strain_stiffness(mesh, Ee, nu)
    nn = mesh.n_nodes
    bot = mesh.boundary("bottom")
    top = mesh.boundary("top")
    dofs = np.concatenate([bot, bot + nn, top + nn])
    vals = np.concatenate([np.zeros(2 * len(bot)), np.full(len(top), top_disp)])
    A, b = apply_dirichlet(K, np.zeros(2 * nn), dofs, vals)
    return _to_system(A, b, 2, {"kind": "elasticity", "m": m, "nx": mesh.nx, "seed": seed})


def damage_system(m: int, seed_h: int = SEED_DAMAGE_H, seed_d: int = SEED_DAMAGE_DOLD, Gc: float = 2.7,
                  l0: float = 0.01, h_scale: float = 1.0):
    """Damage subproblem Gc*l0*K + (Gc/l0 + 2H)*M with H ~ h_scale *
    lognormal(0,1) per element and bounds d_old <= d <= 1, d_old ~ U[0,0.2]."""
    mesh = Mesh(m, m)
    H = h_scale * np.exp(normal(seed_h, len(mesh.tri)))
    K = p1_stiffness(mesh)
    M = p1_mass(mesh, Gc / l0 + 2.0 * H)
    A = (Gc * l0) * K + M
    A = sp.csr_matrix(A)
    A.sort_indices()
    # driving force of the AT2 functional: b = -2 * M_H * 1 (psi+ pushes d up)
    MH = p1_mass(mesh, 2.0 * H)
    b = -(MH @ np.ones(mesh.n_nodes))
    d_old = uniform(seed_d, mesh.n_nodes, 0.0, 0.2)
    sys_ = _to_system(A, b, 1, {"kind": "damage", "m": m})
    return sys_, d_old, np.ones(mesh.n_nodes)


# ------------------------------------------------- adjoint (config 5) inputs

SEED_GRF_OBS = 0x5EED0106
SEED_OBS_NODES = 0x5EED0206


def grf_cells(nc: int, seed: int, length: float = 0.1, var: float = 0.25) -> np.ndarray:
    """Stationary Gaussian field with covariance var * exp(-r^2 / (2 length^2))
    at the centres of the nc x nc cells of the unit square, by circulant
    embedding on a (2 nc) x (2 nc) periodic grid (negative eigenvalues of the
    embedding, if any, are set to zero).  Returns g[r, c] (r along y)."""
    M = 2 * nc
    h = 1.0 / nc
    k = np.arange(M)
    d = np.minimum(k, M - k) * h
    c = var * np.exp(-(d[:, None] ** 2 + d[None, :] ** 2) / (2.0 * length ** 2))
    lam = np.maximum(np.fft.fft2(c).real, 0.0)
    z = normal(seed, 2 * M * M)
    w = (z[: M * M] + 1j * z[M * M:]).reshape(M, M)
    f = np.fft.fft2(np.sqrt(lam / (M * M)) * w)
    return np.ascontiguousarray(f.real[:nc, :nc])


def element_field(mesh: "Mesh", cells: np.ndarray) -> np.ndarray:
    """Per-element values from per-cell values: both triangles of a cell (the
    element order of Mesh.tri is all first triangles, then all second ones)."""
    v = cells.ravel()
    return np.concatenate([v, v])


def elasticity_theta(mesh: "Mesh", theta: np.ndarray, nu: float = 0.3, top_disp: float = 1e-3):
    """P1 plane strain with Young's modulus theta per element, bottom clamped,
    top u_y = top_disp; returns the System and the constrained dofs."""
    K = plane_strain_stiffness(mesh, theta, nu)
    nn = mesh.n_nodes
    bot = mesh.boundary("bottom")
    top = mesh.boundary("top")
    dofs = np.concatenate([bot, bot + nn, top + nn])
    vals = np.concatenate([np.zeros(2 * len(bot)), np.full(len(top), top_disp)])
    A, b = apply_dirichlet(K, np.zeros(2 * nn), dofs, vals)
    return _to_system(A, b, 2, {"kind": "elasticity_theta", "m": mesh.nx}), dofs


@dataclass
class AdjointCase:
    """Config-5 inputs: theta = exp(g) with g a seeded Gaussian random field,
    observations of both displacement components at a seeded 5 % of the nodes,
    taken from the solution for a second field theta_obs."""
    mesh: "Mesh"
    theta: np.ndarray
    theta_obs: np.ndarray
    obs_nodes: np.ndarray
    nu: float = 0.3

    @property
    def obs_dofs(self) -> np.ndarray:
        nn = self.mesh.n_nodes
        return np.concatenate([self.obs_nodes, self.obs_nodes + nn])

    def system(self, theta: np.ndarray):
        return elasticity_theta(self.mesh, theta, self.nu)

    def loss(self, u: np.ndarray, u_obs: np.ndarray) -> float:
        r = u[self.obs_dofs] - u_obs[self.obs_dofs]
        return 0.5 * float(r @ r)

    def dloss(self, u: np.ndarray, u_obs: np.ndarray, constrained: np.ndarray) -> np.ndarray:
        g = np.zeros_like(u)
        d = self.obs_dofs
        g[d] = u[d] - u_obs[d]
        g[constrained] = 0.0
        return g


def adjoint_case(m: int, seed: int = SEED_GRF, seed_obs: int = SEED_GRF_OBS, seed_nodes: int = SEED_OBS_NODES,
                 frac: float = 0.05, nu: float = 0.3) -> AdjointCase:
    mesh = Mesh(m, m)
    theta = np.exp(element_field(mesh, grf_cells(m - 1, seed)))
    theta_obs = np.exp(element_field(mesh, grf_cells(m - 1, seed_obs)))
    keys = uniform(seed_nodes, mesh.n_nodes)
    k = max(1, int(round(frac * mesh.n_nodes)))
    obs = np.sort(np.argsort(keys, kind="stable")[:k])
    return AdjointCase(mesh, theta, theta_obs, obs, nu)


# ------------------------------------------------- batched specimens (config 4)

def batch_specimens(count: int = 512, seed: int = SEED_BATCH, side_min: int = 32, side_max: int = 180,
                    E: float = 210e3, nu: float = 0.3) -> list:
    """Independent degraded-elasticity specimens: structured triangulated
    rectangles with nx, ny ~ U{side_min..side_max} nodes (spacing 1/(nx-1),
    width 1), each with its own seeded notch (height and angle) and
    elementwise E = E * lognormal(0, 0.5).  Specimen k uses the streams of
    seed + 16 k."""
    out = []
    for k in range(count):
        sk = seed + 16 * k
        u = uniform(sk, 2)
        span = side_max - side_min + 1
        nx = side_min + min(int(u[0] * span), span - 1)
        ny = side_min + min(int(u[1] * span), span - 1)
        lx = 1.0
        ly = (ny - 1) / (nx - 1)
        notch = notch_segment(sk + 1, 0.3 * lx, lx, ly)
        ne = 2 * (nx - 1) * (ny - 1)
        E_elem = E * np.exp(0.5 * normal(sk + 2, ne))
        s = degraded_elasticity(ny, seed=sk + 1, nx=nx, lx=lx, ly=ly, E_elem=E_elem, notch=notch, nu=nu)
        s.meta.update({"nx": nx, "ny": ny, "specimen": k})
        out.append(s)
    return out


def block_diagonal(systems: list):
    """The union CSR of independent systems: (row_off, rowptr, col, val, b)."""
    row_off = np.zeros(len(systems) + 1, dtype=np.int32)
    for k, s in enumerate(systems):
        row_off[k + 1] = row_off[k] + s.n
    nnz_off = np.cumsum([0] + [s.nnz for s in systems])
    rowptr = np.zeros(row_off[-1] + 1, dtype=np.int32)
    col = np.empty(nnz_off[-1], dtype=np.int32)
    val = np.empty(nnz_off[-1])
    for k, s in enumerate(systems):
        rowptr[row_off[k] + 1:row_off[k + 1] + 1] = s.rowptr[1:] + nnz_off[k]
        col[nnz_off[k]:nnz_off[k + 1]] = s.col + row_off[k]
        val[nnz_off[k]:nnz_off[k + 1]] = s.val
    b = np.concatenate([s.b for s in systems])
    return row_off, rowptr, col, val, b


# ---------------------------------------------------------------- sketches

SEED_SKETCH = 0x5EED00A0


def sketch_indices(n: int, count: int, seed: int = SEED_SKETCH) -> np.ndarray:
    """`count` distinct indices of [0, n), sorted: the first `count` distinct
    values of floor(u_k * n) along the SplitMix64 stream."""
    draws = np.minimum((uniform(seed, 4 * count) * n).astype(np.int64), n - 1)
    _, first = np.unique(draws, return_index=True)
    if len(first) < count:
        raise ValueError("sketch_indices: too few distinct draws")
    return np.sort(draws[np.sort(first)[:count]])


def gaussian_sketch(x: np.ndarray, rows: int = 64, seed: int = SEED_SKETCH) -> np.ndarray:
    """S x for the seeded Gaussian S of `rows` x len(x) entries N(0, 1/rows),
    S's row r being normal(seed + 1 + r, len(x)) / sqrt(rows): E|S e|^2 =
    |e|^2, so |S x - S y| estimates |x - y| (relative spread ~ sqrt(2/rows)).
    x may hold several vectors as columns (n, k): S x is then (rows, k)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty((rows,) + x.shape[1:])
    for r in range(rows):
        out[r] = normal(seed + 1 + r, x.shape[0]) @ x
    return out / math.sqrt(rows)
