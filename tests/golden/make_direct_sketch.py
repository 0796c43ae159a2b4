"""Write a seeded sketch of the FP64 direct solution of the 2.1M-DOF
degraded-elasticity system (1025x1025 nodes) as the independent oracle of the
headline solve's error.

    python tests/golden/make_direct_sketch.py [m]     (default m = 1025)

SuperLU (scipy.sparse.linalg.splu, COLAMD ordering, FP64) factors the system
in this container (several minutes, ~20 GB); the full solution (16.8 MB) is
not stored, only what measures an iterate's error against it:

* x_direct at 20,000 seeded indices (problems.sketch_indices);
* |x_direct|_2 and the true relative residual |b - A x_direct| / |b|;
* S x_direct for the seeded 64-row Gaussian S of problems.gaussian_sketch,
  so that |S x - S x_direct| / |x_direct| estimates |x - x_direct| / |x_direct|
  within ~ +-18 % (chi-square with 64 degrees of freedom);
* checksums of the assembled system (n, nnz, sum of values, |b|), which the
  readers compare before trusting the sketch.

Output: tests/golden/direct_elasticity{m}.npz
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fem_amd import problems as P  # noqa: E402
from tests import direct_oracle as D  # noqa: E402


def main(m: int = 1025) -> None:
    from scipy.sparse.linalg import splu
    t0 = time.time()
    s = P.degraded_elasticity(m)
    A = s.scipy()
    print(f"system n={s.n} nnz={s.nnz} ({time.time() - t0:.1f}s)", flush=True)
    t0 = time.time()
    lu = splu(A.tocsc())
    print(f"splu factor {time.time() - t0:.1f}s", flush=True)
    x = lu.solve(s.b)
    del lu
    r = s.b - A @ x
    relres = float(np.linalg.norm(r) / np.linalg.norm(s.b))
    print(f"relres {relres:.3e}", flush=True)
    np.savez_compressed(os.path.join(HERE, f"direct_elasticity{m}.npz"), m=m,
                        **D.make(s, x, relres, "scipy splu (SuperLU, COLAMD), FP64"))
    print("written", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1025)
