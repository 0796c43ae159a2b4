"""Batched AMG-PCG (lina_batch) on independent degraded-elasticity specimens
of different sizes, notches and elementwise moduli:

* every specimen reaches |b - A x| <= 1e-8 |b| (checked on the host) and its
  solution agrees with a direct solve;
* iteration counts match the single-system pcg of the same hierarchy with
  the same stopping rule (within rounding of the dot products);
* the hipGraph replay gives bit-identical iterates and counts to the eager
  loop, on two consecutive solves;
* the autotuned row partitions give the same convergence."""
import numpy as np
import pytest
from scipy.sparse.linalg import spsolve

from fem_amd import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def specimens():
    return P.batch_specimens(10, side_min=17, side_max=40)


def _params(gpu, autotune=0):
    return gpu.params("vcycle", 2, 4, precision="mixed", autotune=autotune)


def test_batch_converges_and_matches_single_solves(gpu, specimens):
    row_off, rp, ci, v, b = P.block_diagonal(specimens)
    bt = gpu.Batch(row_off, rp, ci, v, _params(gpu))
    out = bt.pcg(b, tol=1e-8, max_iter=2000)
    say("batch iters", out["iters"].tolist(), "launched", out["launched"], "ms", out["t_solve_ms"])
    assert out["active"] == 0
    for k, s in enumerate(specimens):
        x = out["x"][row_off[k]:row_off[k + 1]]
        A = s.scipy()
        rr = np.linalg.norm(s.b - A @ x) / np.linalg.norm(s.b)
        xd = spsolve(A.tocsc(), s.b)
        err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
        single = gpu.AMG(s.rowptr, s.col, s.val, _params(gpu)).pcg(s.b, tol=1e-8, max_iter=2000, stop="relres")
        say("  specimen", k, s.n, "iters", out["iters"][k], "single", single["iters"], "relres", rr, "err", err)
        assert rr <= 1.01e-8
        assert err <= 1e-5
        assert abs(out["iters"][k] - single["iters"]) <= max(2, 0.05 * single["iters"])


def test_graph_replay_is_identical_to_eager(gpu, specimens):
    row_off, rp, ci, v, b = P.block_diagonal(specimens)
    bt = gpu.Batch(row_off, rp, ci, v, _params(gpu))
    eager = bt.pcg(b, tol=1e-8, max_iter=2000, graph_iters=0)
    for rep in range(2):
        g = bt.pcg(b, tol=1e-8, max_iter=2000, graph_iters=7)
        say("graph replay", rep, "launched", g["launched"], "ms", g["t_solve_ms"], "eager ms", eager["t_solve_ms"])
        assert np.array_equal(g["iters"], eager["iters"])
        assert np.array_equal(g["x"], eager["x"])


def test_autotuned_partitions_converge(gpu, specimens):
    row_off, rp, ci, v, b = P.block_diagonal(specimens)
    bt = gpu.Batch(row_off, rp, ci, v, _params(gpu, autotune=1))
    tuned = bt.tuned()
    say("tuned", tuned)
    assert len(tuned) > 0 and all(t[2] in (1, 4, 8, 16, 32) and t[3] in (1, 4, 8, 16, 32) for t in tuned)
    out = bt.pcg(b, tol=1e-8, max_iter=2000, graph_iters=5)
    assert out["active"] == 0
    for k, s in enumerate(specimens):
        x = out["x"][row_off[k]:row_off[k + 1]]
        rr = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
        assert rr <= 1.01e-8
