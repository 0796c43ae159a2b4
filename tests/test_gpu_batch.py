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


def test_converged_specimens_stay_frozen_in_shared_row_blocks(gpu):
    """Small specimens (several share one row block of the union's kernels)
    interleaved with larger, slower ones, so row blocks hold converged and
    active systems at once.  A converged specimen's iterate must not move
    while its neighbours iterate on: the specimen's X after the full solve
    equals, bit for bit, its X from a solve stopped at exactly its own
    iteration count.  Each specimen also matches a single-system pcg of the
    same parameters (iterations within 1, solutions to the tolerance)."""
    small = P.batch_specimens(9, seed=P.SEED_BATCH + 7001, side_min=9, side_max=12)
    large = P.batch_specimens(3, seed=P.SEED_BATCH + 9001, side_min=36, side_max=44)
    sp = []
    for k in range(9):
        sp.append(small[k])
        if k % 3 == 2:
            sp.append(large[k // 3])
    row_off, rp, ci, v, b = P.block_diagonal(sp)
    prm = gpu.params("vcycle", 2, 4, precision="fp64")
    bt = gpu.Batch(row_off, rp, ci, v, prm)
    full = bt.pcg(b, tol=1e-10, max_iter=3000)
    say("iterations", full["iters"].tolist(), "sizes", [s.n for s in sp])
    assert full["active"] == 0
    assert full["iters"].max() > 2 * full["iters"].min()
    early = sorted(set(int(i) for i in full["iters"]))[:3]
    for k_stop in early:
        part = bt.pcg(b, tol=1e-10, max_iter=k_stop)
        for k in np.flatnonzero(full["iters"] == k_stop):
            lo, hi = row_off[k], row_off[k + 1]
            say("  specimen", int(k), "stopped at", k_stop, "frozen:", np.array_equal(part["x"][lo:hi], full["x"][lo:hi]))
            assert part["iters"][k] == k_stop
            assert np.array_equal(part["x"][lo:hi], full["x"][lo:hi])
    for k, s in enumerate(sp):
        single = gpu.AMG(s.rowptr, s.col, s.val, prm).pcg(s.b, tol=1e-10, max_iter=3000, stop="relres")
        x = full["x"][row_off[k]:row_off[k + 1]]
        d = np.linalg.norm(x - single["x"]) / np.linalg.norm(single["x"])
        say("  specimen", k, s.n, "iters", int(full["iters"][k]), "single", single["iters"], "rel", d)
        assert abs(int(full["iters"][k]) - single["iters"]) <= 1
        assert d <= 1e-8
