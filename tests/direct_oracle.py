"""The FP64 direct-solve oracle of the 2.1M-DOF degraded-elasticity system,
read from its committed sketch (tests/golden/direct_elasticity1025.npz,
written by tests/golden/make_direct_sketch.py with SuperLU in FP64).

    errors(s, [x1, x2, ...]) -> one dict per iterate:
      sketch  |S x - S x_direct| / |x_direct|  (64-row Gaussian sketch: an
              estimate of the relative 2-norm error, +-18 % at one sigma)
      sample  |x - x_direct| / |x_direct| over the 20,000 sampled indices
      sample_max  max |x - x_direct| / max |x_direct| over those indices

Before it answers it checks that `s` is the system the sketch was made from
(size, nonzeros, sum of values, |b|)."""
import os

import numpy as np

from fem_amd import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture(m: int = 1025):
    path = os.path.join(HERE, "golden", f"direct_elasticity{m}.npz")
    return np.load(path) if os.path.exists(path) else None


def make(s, x, relres: float, solver: str) -> dict:
    """the sketch of the direct solution x of system s (what the fixture holds)"""
    idx = P.sketch_indices(s.n, min(20000, s.n // 4))
    return {"n": s.n, "nnz": s.nnz, "val_sum": float(np.sum(s.val)), "b_norm": float(np.linalg.norm(s.b)),
            "idx": idx.astype(np.int32), "x_idx": x[idx], "x_norm": float(np.linalg.norm(x)),
            "sketch": P.gaussian_sketch(x), "sketch_rows": 64, "sketch_seed": P.SEED_SKETCH, "relres": relres,
            "solver": solver}


def errors(s, xs, m: int = 1025, g=None):
    g = fixture(m) if g is None else g
    if g is None:
        raise FileNotFoundError(f"no direct-solve sketch for m={m}")
    if (int(g["n"]) != s.n or int(g["nnz"]) != s.nnz or float(g["b_norm"]) != float(np.linalg.norm(s.b))
            or abs(float(g["val_sum"]) - float(np.sum(s.val))) > 1e-12 * abs(float(g["val_sum"]))):
        raise ValueError("the direct-solve sketch belongs to a different system")
    X = np.stack([np.asarray(x, dtype=np.float64) for x in xs], axis=1)
    SX = P.gaussian_sketch(X, int(g["sketch_rows"]), int(g["sketch_seed"]))
    idx = g["idx"]
    xd = g["x_idx"]
    out = []
    for k in range(X.shape[1]):
        e = X[idx, k] - xd
        out.append({"sketch": float(np.linalg.norm(SX[:, k] - g["sketch"]) / float(g["x_norm"])),
                    "sample": float(np.linalg.norm(e) / np.linalg.norm(xd)),
                    "sample_max": float(np.abs(e).max() / np.abs(xd).max())})
    return out, {"kind": str(g["solver"]), "relres": float(g["relres"]), "x_norm": float(g["x_norm"])}
