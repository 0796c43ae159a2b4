"""Write the golden fixtures of the LinA solver path from the compiled reference.

    make -C oracle ref && python tests/golden/make_golden.py

Runs the reference LinA (oracle/_ref/libreflina.so, built from
/root/reference/LinA by oracle/Makefile) on small seeded systems and stores
inputs and outputs as compressed .npz next to this script:

* the full AMG hierarchy of AMG_setup (labels, A/P/R of every level);
* AMG_solve results for each smoother from a zero initial guess;
* pcg with the reference preconditioner (PCG_set_AMG_preconditioner) and
  with the symmetric multicolour V-cycle, including the in-place rounding of
  B and get_pcg_residuum;
* krylov.npz: ILU(0) factors (incomplete_LU_factorization), pcg with
  PCG_default's ILU(0) preconditioner, FEM.c's
  GMRES(30)+ILU(0) path, pcg with one V-cycle of the level-dependent
  smoother (MC-SSOR finest, l1-Jacobi below) and BiCGStab with the
  BI_set_AMG_preconditioner hierarchy, on three small systems.

    python tests/golden/make_golden.py krylov    # krylov.npz only

The reference is only run here; the fixtures are data (inputs and outputs).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fem_amd import problems as P  # noqa: E402
from oracle.bind import Ref  # noqa: E402

CASES = [
    # name, system builder, dof, depth, coarsening
    ("poisson17", lambda: P.poisson(17), 4, 1),
    ("poisson33", lambda: P.poisson(33), 4, 1),
    ("poisson33_std", lambda: P.poisson(33), 5, 2),
    ("elast17", lambda: P.degraded_elasticity(17), 4, 1),
    ("elast33", lambda: P.degraded_elasticity(33), 4, 1),
]

SMOOTHERS = [("sor", 1.6), ("jacobi", 1.0), ("mc_sor", 1.2), ("l1_jacobi", 1.0), ("mc_ssor", 1.0), ("ssor", 1.0)]


KRYLOV = [
    # name, system builder, right-hand side seed
    ("poisson17", lambda: P.poisson(17), 1),
    ("damage17", lambda: P.damage_system(17)[0], 2),
    ("elast17", lambda: P.degraded_elasticity(17), 3),
]


def krylov(ref):
    out = {}
    for name, build, seed in KRYLOV:
        s = build()
        b = np.random.default_rng(seed).standard_normal(s.n)
        out.update({f"{name}_rowptr": s.rowptr, f"{name}_col": s.col, f"{name}_val": s.val, f"{name}_b": b,
                    f"{name}_dof": s.dof})
        M = ref.matrix(s.rowptr, s.col, s.val)
        ok, L, U = ref.ilu(M)
        assert ok
        for w, F in (("L", L), ("U", U)):
            out.update({f"{name}_ilu{w}_rp": F.rowptr, f"{name}_ilu{w}_ci": F.col, f"{name}_ilu{w}_v": F.val})
        c = ref.pcg_ilu(M, b, tol=1e-10, max_iter=2000)
        out.update({f"{name}_pcgilu_iters": c["iters"], f"{name}_pcgilu_x": c["x"], f"{name}_pcgilu_res": c["res"]})
        g = ref.gmres_ilu(M, b, m=30, tol=1e-8, max_iter=600)
        out.update({f"{name}_gmres_iters": g["iters"], f"{name}_gmres_x": g["x"], f"{name}_gmres_eta": g["eta"]})
        h = ref.pcg_amg(M, b, dof=s.dof, depth=4, smiter=1, aggr_smiter=1, smoother="hybrid", alpha=1.0,
                        amg_eps=1e-2, amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=2000, coarse_sweeps=1)
        out.update({f"{name}_pcghyb_iters": h["iters"], f"{name}_pcghyb_x": h["x"], f"{name}_pcghyb_res": h["res"]})
        r = ref.bicgstab_amg(M, b, dof=s.dof, depth=4, smoother="sor", alpha=1.0, eps=1e-8, max_iter=300)
        out.update({f"{name}_bicg_iters": r["iters"], f"{name}_bicg_x": r["x"]})
        ref.free(M)
    path = os.path.join(HERE, "krylov.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")


def main():
    ref = Ref()
    if sys.argv[1:] == ["krylov"]:
        krylov(ref)
        return
    for name, build, depth, coarsening in CASES:
        s = build()
        out = {"rowptr": s.rowptr, "col": s.col, "val": s.val, "b": s.b, "dof": s.dof, "depth": depth,
               "coarsening": coarsening}
        M = ref.matrix(s.rowptr, s.col, s.val)
        I = ref.amg_setup(M, dof=s.dof, depth=depth, coarsening=coarsening)
        for l, lv in enumerate(ref.amg_levels(I, depth)):
            for w in ("A", "P", "R"):
                if w in lv:
                    out[f"L{l}_{w}_rp"] = lv[w].rowptr
                    out[f"L{l}_{w}_ci"] = lv[w].col
                    out[f"L{l}_{w}_v"] = lv[w].val
            if "labels" in lv:
                out[f"L{l}_labels"] = lv["labels"]
        for sm, alpha in SMOOTHERS:
            k, x = ref.amg_solve(I, s.b, np.zeros(s.n), smoother=sm, alpha=alpha, eps=1e-6, max_iter=6)
            out[f"amgsolve_{sm}_k"] = k
            out[f"amgsolve_{sm}_x"] = x
        ref.amg_free(I)
        if coarsening == 1:
            r = ref.pcg_amg(M, s.b, dof=s.dof, depth=depth, tol=1e-10, max_iter=500)
            out.update({"pcgref_iters": r["iters"], "pcgref_x": r["x"], "pcgref_b": r["b_out"],
                        "pcgref_res": r["res"], "pcgref_vc": r["vcycles"]})
            r = ref.pcg_amg(M, s.b, dof=s.dof, depth=depth, smiter=1, aggr_smiter=1, smoother="mc_ssor", alpha=1.0,
                            amg_eps=1e-2, amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=2000)
            out.update({"pcgvc_iters": r["iters"], "pcgvc_x": r["x"], "pcgvc_b": r["b_out"], "pcgvc_res": r["res"]})
        ref.free(M)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")
    krylov(ref)


if __name__ == "__main__":
    main()
