"""Config-5 inputs and the CPU reference gradient, without a GPU: the
Gaussian random field has the requested covariance, and the adjoint gradient
of the reference (tests/adjoint_ref.py) matches central finite differences of
the loss with direct solves -- which pins the sign, the Dirichlet handling
and the element matrices the device kernel is compared with.  At h = 1e-6 the
difference quotient is evaluated without cancellation (tests/adjoint_ref.py
fd_pieces): naively, the rounding of the assembled perturbed matrices alone
leaves ~1e-5 relative noise at this step."""
import numpy as np

from fem_amd import problems as P
from tests import adjoint_ref as R


def test_grf_covariance():
    a = np.array([P.grf_cells(32, s) for s in range(300)])
    assert abs(a.var(axis=0).mean() - 0.25) < 0.02
    lag = 3  # cells of 1/32
    cov = np.mean(a[:, :, :-lag] * a[:, :, lag:])
    expected = 0.25 * np.exp(-(lag / 32) ** 2 / (2 * 0.1 ** 2))
    assert abs(cov - expected) < 0.02


def test_element_matrices_assemble_the_stiffness():
    case = P.adjoint_case(9)
    Ke = R.element_matrices(case.mesh, case.nu)
    ed = R.element_dofs(case.mesh)
    n = 2 * case.mesh.n_nodes
    K = np.zeros((n, n))
    for e in range(len(Ke)):
        K[np.ix_(ed[e], ed[e])] += case.theta[e] * Ke[e]
    K2 = P.plane_strain_stiffness(case.mesh, case.theta, case.nu).toarray()
    assert np.allclose(K, K2, rtol=1e-12, atol=1e-12 * np.abs(K2).max())


def test_reference_gradient_matches_finite_differences():
    case = P.adjoint_case(17)
    L, grad, u, u_obs, lam = R.loss_and_gradient(case)
    assert L > 0
    # plain central differences, h = 1e-4 (truncation and noise both < 1e-7)
    h = 1e-4
    for v in R.directions(len(case.theta), 3):
        fd = (R.loss_at(case, case.theta + h * v, u_obs) - R.loss_at(case, case.theta - h * v, u_obs)) / (2 * h)
        assert abs(fd - grad @ v) <= 1e-6 * abs(grad @ v), (fd, grad @ v)
    # h = 1e-6 with the cancellation-free evaluation
    h = 1e-6
    from scipy.sparse.linalg import splu
    for v in R.directions(len(case.theta), 5):
        r0, rv = R.fd_pieces(case, u, v)
        d = [splu(case.system(case.theta + sg * h * v)[0].scipy().tocsc()).solve(r0 + sg * h * rv) for sg in (1, -1)]
        fd = R.fd_quotient(case, d[0], d[1], u, u_obs, h)
        assert abs(fd - grad @ v) <= 1e-9 * abs(grad @ v), (fd, grad @ v)
