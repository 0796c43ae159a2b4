#!/usr/bin/env python3
"""Headline benchmark: AMG-PCG solve time on the 2.1M-DOF degraded-elasticity
system, mixed precision vs FP64, with the finest-level SpMV and the whole CG
iteration (algorithmic bytes summed over the levels) against the HBM
roofline, the error against a committed SuperLU solution sketch, and the
reference LinA timed on the host cores beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 1025]

Workload: P1 plane-strain elasticity on a 1025x1025-node triangulated unit
square (2,101,250 DOF, blocked [ux; uy]), E = 210e3, nu = 0.3, stiffness scaled
per element by (1-d)^2 + 1e-6 with d = exp(-dist/0.01) to a seeded notch of
length 0.3 from the left edge; bottom clamped, top u_y = 1e-3.  Preconditioner:
the LinA AMG hierarchy (aggressive first coarsening, bit-identical to the
reference's AMG_setup) with one symmetric V-cycle per CG iteration:
multicolour SOR forward and backward on the finest level, one l1-Jacobi
sweep on the levels below, levels 1, 2 and 4 (0 = coarsest) passed through
unsmoothed (--skip-levels; the full cycle's mixed solve is reported beside
it).  A "step" is one complete AMG-PCG solve from x0 = 0 to
|b-Ax|/|b| <= tol, with A, b and the hierarchy resident in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class _CStdoutToStderr:
    """C-level writes to stdout (the reference's printf warnings) go to
    stderr while inside, so stdout carries only the JSON line"""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.libc.fflush(None)
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        self.libc.fflush(None)
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def cpu_baseline(s, amg, gpu_iters: int) -> dict:
    """The reference LinA (oracle/_ref, -O3 for x86-64-v3 where built) timed
    on the full 2.1M-DOF system on this machine's host cores: its pcg with one
    symmetric lexicographic-SOR V(1,1) per iteration (the reference's SOR
    arithmetic, AMG_set_smoother) on the same hierarchy (handed over from the
    library's host setup, bit-identical to AMG_setup; the reference's own
    AMG_setup takes ~14 min here), the coarsest level solved with the dense
    inverse of its dense rows.  Timed: pcg with 2 and with 5 iterations from
    x = 0; the difference / 3 is the cost of one iteration, at 1 thread and at
    the host's core share.  value = that cost x the iteration count of the
    GPU solve with every level smoothed (the reference's cycle shape)."""
    from oracle.bind import REF_FAST_SO, REF_SO, Ref
    path = REF_FAST_SO if os.path.exists(REF_FAST_SO) else REF_SO
    ref = Ref(path)
    ref.L.refh_init(2)
    t0 = time.time()
    lv = [amg.level(l) for l in range(amg.depth)]
    h, keep = ref.assemble(lv, 2, amg.signs())
    M = ref.matrix(s.rowptr, s.col, s.val)
    t_asm = time.time() - t0
    log(f"cpu baseline: reference hierarchy assembled in {t_asm:.1f}s ({os.path.basename(path)})")
    per = {}
    cores = min(16, os.cpu_count() or 1)
    for th in sorted({1, cores}):
        ref.L.refh_set_threads(th)
        r2 = ref.pcg_assembled(h, M, s.b, max_iter=2)
        r5 = ref.pcg_assembled(h, M, s.b, max_iter=5)
        per[th] = (r5["t_solve"] - r2["t_solve"]) / 3
        log(f"cpu baseline: {th} thread(s): pcg 2 its {r2['t_solve']:.2f}s, 5 its {r5['t_solve']:.2f}s -> "
            f"{per[th]:.3f}s per iteration")
    ref.free(M)
    best = min(per, key=per.get)
    return {"value": per[best] * gpu_iters, "unit": "s per solve (measured per-iteration cost x GPU iterations)",
            "cores": best, "kind": "reference",
            "sample": (f"reference pcg + AMG_solve (PCGS.c, Multigrid.c; {os.path.basename(path)}) on the full "
                       f"{s.n}-DOF system and hierarchy, one symmetric lexicographic-SOR V(1,1) per iteration: "
                       f"{per[1]:.3f}s per iteration at 1 thread, {per[cores]:.3f}s at {cores}; x {gpu_iters} "
                       f"iterations (the GPU solve's count with every level smoothed); AMG_setup excluded (the "
                       f"reference's takes ~14 min on one core of the build container)"),
            "per_iteration_s": {str(k): v for k, v in per.items()}, "assemble_s": t_asm}


def spmv_alg_bytes(nnz: int, rows: int, cols: int, prec: str) -> int:
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + row
    pointers + x gathered once + y written once (FP64: 12 nnz + 4 (rows+1) +
    8 cols + 8 rows; FP32 values and vectors: 8 nnz + 4 (rows+1) + 4 cols +
    4 rows).  The library stores 16-bit column codes where they fit, so it
    moves fewer bytes than this."""
    if prec == "fp64":
        return 12 * nnz + 4 * (rows + 1) + 8 * cols + 8 * rows
    return 8 * nnz + 4 * (rows + 1) + 4 * cols + 4 * rows


def iteration_bytes(amg, precision: str, coarse_sweeps: int = 1, skip_mask: int = 0) -> dict:
    """Algorithmic bytes of one AMG-PCG iteration from the actual level sizes:
    the outer FP64 SpMV, two dots (16 n each), the fused x/r update with the
    residual norm (48 n) and p = z + beta p (24 n); per level above the
    coarsest its smoothing sweeps (a sweep = SpMV + b and the diagonal
    scaling read; two on the finest level -- the forward and the reverse
    multicolour SOR pass -- and 2 x coarse_sweeps l1-Jacobi sweeps below, of
    which the first starts from the zero guess and reads no matrix entry:
    f, the l1 diagonal and x, 3 words a row), one residual (SpMV + b), the
    restriction R and the prolongation P; the coarsest level's dense inverse
    applied once (its FP64 upper triangle, 4 nd^2, + 16 nd, nd its dense
    rows).  A pass-through level (bit l of skip_mask) moves only R and P.
    The hierarchy's arithmetic is FP32 in the mixed mode."""
    top = amg.depth - 1
    L = [amg.level_sizes(l) for l in range(amg.depth)]
    n0, nnz0 = L[top]["fine"], L[top]["nnz_A"]
    outer = spmv_alg_bytes(nnz0, n0, n0, "fp64") + 2 * 16 * n0 + 48 * n0 + 24 * n0
    vp = "fp64" if precision == "fp64" else "fp32"
    vb = 8 if vp == "fp64" else 4
    per_level = []
    for l in range(top, 0, -1):
        d = L[l]
        n, nc = d["fine"], d["coarse"]
        a = spmv_alg_bytes(d["nnz_A"], n, n, vp)
        if l != top and (skip_mask >> l) & 1:
            per_level.append(int(spmv_alg_bytes(d["nnz_R"], nc, n, vp) + spmv_alg_bytes(d["nnz_P"], n, nc, vp)))
            continue
        if l == top:
            smooth = 2 * (a + 2 * vb * n)
        else:
            smooth = 3 * vb * n + (2 * coarse_sweeps - 1) * (a + 2 * vb * n)
        per_level.append(int(smooth + (a + vb * n) + spmv_alg_bytes(d["nnz_R"], nc, n, vp) +
                             spmv_alg_bytes(d["nnz_P"], n, nc, vp)))
    nd = amg.coarse_rows()[0]
    coarse = 4 * nd * nd + 16 * nd
    return {"outer": int(outer), "levels": per_level, "coarse": int(coarse),
            "total": int(outer + sum(per_level) + coarse)}


def copy_gbs(lina, n=1 << 27, reps=5):
    """hipMemcpy device-to-device of n doubles (read n*8 + write n*8 bytes):
    the attainable streaming rate beside the 8 TB/s peak."""
    a = lina.DeviceVector(n)
    b = lina.DeviceVector(n)
    a.fill(1.0)
    b.copy_from(a)
    t = time.perf_counter()
    for _ in range(reps):
        b.copy_from(a)
    dt = (time.perf_counter() - t) / reps
    return 16.0 * n / dt / 1e9


def kernels_only(s, lina):
    """The launches the counter runs read: one copy of known size (the
    calibration of FETCH_SIZE / WRITE_SIZE) and the finest-level SpMVs, on a
    device matrix handle (the same kernel and row blocks as the solver's)."""
    gbs = copy_gbs(lina, reps=3)
    m = lina.Matrix(s.rowptr, s.col, s.val)
    ms64 = m.time_spmv("fp64", reps=5)
    ms32 = m.time_spmv("mixed", reps=5)
    print(json.dumps({"copy_bytes": 16 * (1 << 27), "copy_gbs": gbs, "spmv_fp64_bytes": m.spmv_bytes("fp64"),
                      "spmv_fp64_ms": ms64, "spmv_fp32_bytes": m.spmv_bytes("mixed"), "spmv_fp32_ms": ms32}))


def measured_traffic(path=os.path.join(ROOT, "profiles", "traffic_spmv_outer_elasticity1025.md")):
    """HBM bytes per launch of the finest-level SpMV from the committed
    counter summary (the line `traffic_bytes_fp64: <bytes>`), or None."""
    try:
        for line in open(path):
            if line.strip().startswith("traffic_bytes_fp64:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def _cpu_run(code: str, timeout: int = 900) -> dict:
    """run `code` in a fresh single-threaded interpreter, last JSON line out"""
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=timeout)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        raise RuntimeError(f"cpu run failed ({r.returncode}): {r.stderr[-1500:]}")
    return json.loads(lines[-1])


def workload_poisson(args, lina, problems):
    """The reference's own CPU case: 257^2 P1 Poisson (66,049 DOF), AMG-PCG to
    relative residual 1e-10 with one symmetric V(1,1) per CG iteration.  The
    GPU runs it with the exact lexicographic SSOR V-cycle (the reference's
    arithmetic, iterates equal to 1e-12) and with its multicolour form; the
    compiled reference runs the lexicographic one on the same system on one
    host core."""
    m, depth = 257, 5
    s = problems.poisson(m)
    log(f"poisson {m}^2: n={s.n} nnz={s.nnz}")
    res = {}
    A = s.scipy()
    for sm in ("ssor", "mc_ssor"):
        prm = lina.params("reference", 1, depth, smoothing_iterations=1, aggr_smoothing_iterations=1,
                          amg_max_iter=1, zero_init=1, smoother=sm, omega=1.0, autotune=1)
        t0 = time.perf_counter()
        amg = lina.AMG(s.rowptr, s.col, s.val, prm)
        t_setup = time.perf_counter() - t0
        for _ in range(args.warmup):
            amg.pcg(s.b, tol=1e-10, max_iter=500)
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            r = amg.pcg(s.b, tol=1e-10, max_iter=500)
            times.append(time.perf_counter() - t0)
        amg.close()
        rr = float(np.linalg.norm(s.b - A @ r["x"]) / np.linalg.norm(s.b))
        res[sm] = {"setup_s": t_setup, "solve_ms": 1e3 * float(np.mean(times)), "iters": r["iters"],
                   "true_relres": rr}
        log(sm, res[sm])
    cpu = None
    try:
        code = f"""
import json, sys, time
sys.path.insert(0, {ROOT!r})
import numpy as np
from fem_amd import problems as P
from oracle.bind import Ref
ref = Ref()
s = P.poisson({m})
M = ref.matrix(s.rowptr, s.col, s.val)
r = ref.pcg_amg(M, s.b, dof=1, depth={depth}, smiter=1, aggr_smiter=1, smoother="ssor", alpha=1.0, amg_eps=1e-2,
                amg_max_iter=1, zero_init=True, tol=1e-10, max_iter=500)
print(json.dumps({{"iters": r["iters"], "t_setup": r["t_setup"], "t_solve": r["t_solve"]}}))
"""
        c = _cpu_run(code)
        cpu = {"value": 1e3 * c["t_solve"], "unit": "ms per solve", "cores": 1, "kind": "reference",
               "sample": f"reference pcg (PCGS.c) with one lexicographic SSOR V(1,1) per iteration on the same "
                         f"257x257 system: {c['iters']} iterations, AMG_setup {c['t_setup']:.2f}s"}
        log(f"cpu reference: {cpu}")
    except Exception as e:
        log(f"cpu baseline skipped: {e}")
    return {"metric": "AMG-PCG solve time @66k DOF (257^2 Poisson, relres 1e-10, one V(1,1) per iteration)",
            "value": res["mc_ssor"]["solve_ms"], "unit": "ms", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": False, "dtype": "f64",
            "data": "synthetic (seeded uniform[-1,1] load)",
            "config": {"workload": "P1 Poisson 257x257, homogeneous Dirichlet, AMG depth 5; value = the multicolour "
                                   "SSOR V(1,1); results.ssor = the exact lexicographic SSOR V(1,1) (the "
                                   "reference's arithmetic, iterates equal to 1e-12)"},
            "results": res, "cpu_baseline": cpu}


def workload_damage(args, lina, problems):
    """Reflective Newton on the 1025^2 damage subproblem (1.05M DOF, d_old <= d
    <= 1) with each inner solver; the reference's RN (ILU-PCG inner) on a
    129^2 member of the family on one host core beside it."""
    m = 1025
    res = {}
    os.environ.setdefault("LINA_RN_TIMING", "1")  # one progress line per Newton step
    for h_scale in (1.0,):
        s, dold, ones = problems.damage_system(m, h_scale=h_scale)
        log(f"damage h_scale={h_scale}: n={s.n} nnz={s.nnz}")
        # the BiCGStab variant's AMG: the reference's AMG_solve settings (eps
        # 1e-4, at most 20 V-cycles from the warm start) in FP32 with
        # multicolour SOR sweeps (one V-cycle from the warm start instead made
        # the inner solves fail and the Newton iteration stall; the exact
        # lexicographic SOR's 2,049-level sweeps took 209 s in all)
        # the ILU inner solvers also in multicolour order (the factors of
        # P A P^T: sweeps in as many levels as colours, on the whole GPU)
        for solver, prec, order in (("pcg", "fp64", "natural"), ("gmres", "fp64", "natural"),
                                    ("bicgstab", "mixed", "natural"), ("pcg", "fp64", "multicolour"),
                                    ("gmres", "fp64", "multicolour")):
            amg = {"precision": 1 if prec == "mixed" else 0}
            if solver == "bicgstab":
                amg["smoother"] = 2  # MC-SOR
            prm = lina.rn_params(solver, 1, 6, delta=float(np.sqrt(s.n) / 2), amg=amg, ilu_ordering=order)
            t0 = time.perf_counter()
            rn = lina.ReflectiveNewton(s.rowptr, s.col, s.val, s.b, dold, ones, prm)
            t1 = time.perf_counter()
            g = rn.solve(dold)
            t2 = time.perf_counter()
            rn.close()
            key = f"h{h_scale:g}_{solver}_{prec}" + ("_multicolour" if order == "multicolour" else "")
            res[key] = {"setup_s": t1 - t0, "solve_s": t2 - t1, "newton_iters": g["iters"],
                        "inner_iters": g["inner_iterations"], "line_searches": g["line_searches"],
                        "neg_curvature": g["neg_curvature"], "q": g["q"], "dq": g["dq"],
                        "converged": bool(g["dq"] <= 1e-8)}
            log(key, res[key])
    cpu = None
    try:
        code = f"""
import json, sys, time
sys.path.insert(0, {ROOT!r})
import numpy as np
from fem_amd import problems as P
from oracle.bind import Ref
ref = Ref(); ref.L.refh_init(2)
s, dold, ones = P.damage_system(129, h_scale=1.0)
H = ref.matrix(s.rowptr, s.col, s.val)
t = time.time()
r = ref.reflective_newton(H, s.b, dold, ones, dold.copy(), solver=3, dof=1, depth=6, max_iter=50, eps=1e-8,
                          delta=float(np.sqrt(s.n) / 2))
print(json.dumps({{"n": s.n, "iters": r["iters"], "t": time.time() - t}}))
"""
        c = _cpu_run(code)
        gi = res["h1_pcg_fp64"]
        per = c["t"] / max(c["iters"], 1) / c["n"]
        cpu = {"value": per * m * m * gi["newton_iters"], "unit": "s per RN solve (scaled)", "cores": 1,
               "kind": "reference",
               "sample": f"reference reflective_newton with RN_PCG (ILU(0)-PCG inner), 129x129 "
                         f"({c['n']} DOF): {c['iters']} Newton steps in {c['t']:.2f}s; scaled per Newton step "
                         f"and DOF to the 1025x1025 solve's step count"}
    except Exception as e:
        log(f"cpu baseline skipped: {e}")
    best = res["h1_pcg_fp64"]["solve_s"]
    return {"metric": "reflective-Newton damage solve time @1.05M DOF", "value": best, "unit": "s",
            "value_multicolour_ilu": res["h1_pcg_fp64_multicolour"]["solve_s"],
            "n_gpus": 1, "higher_is_better": False, "dtype": "f64 (mixed variant: f32 AMG)",
            "data": "synthetic (seeded lognormal H, uniform d_old)",
            "config": {"workload": "damage subproblem P1 1025x1025, H ~ lognormal(0,1), d_old <= d <= 1, RN eps "
                                   "1e-8; value = the ILU(0)-PCG inner solver (natural order, the reference's "
                                   "factors); value_multicolour_ilu = the same in multicolour order"},
            "results": res, "cpu_baseline": cpu}


def workload_batch(args, lina, problems):
    """512 batched degraded-elasticity specimens (sides U{32..180}), FP32
    hierarchy / FP64 pcg to relative residual 1e-8, hipGraph replay."""
    from scipy.sparse.linalg import splu
    count = 512
    t0 = time.time()
    sp = problems.batch_specimens(count)
    row_off, rp, ci, v, b = problems.block_diagonal(sp)
    log(f"{count} specimens, {row_off[-1]} DOF, {len(ci)} nnz, built in {time.time() - t0:.1f}s")
    t0 = time.time()
    bt = lina.Batch(row_off, rp, ci, v, lina.params("vcycle", 2, 5, precision="mixed", autotune=1))
    t_setup = time.time() - t0
    log(f"setup {t_setup:.1f}s, tuned {bt.tuned()}")
    runs = []
    for k in range(args.warmup + args.steps):
        out = bt.pcg(b, tol=1e-8, max_iter=3000, graph_iters=10)
        if k >= args.warmup:
            runs.append(out)
        log(f"solve {k}: {out['t_solve_ms']:.1f} ms, launched {out['launched']}, active {out['active']}")
    ms = float(np.mean([r["t_solve_ms"] for r in runs]))
    out = runs[-1]
    # oracle: direct FP64 solves of every 32nd specimen and the largest
    idx = sorted(set(list(range(0, count, 32)) + [int(np.argmax([s.n for s in sp]))]))
    errs = []
    for k in idx:
        s = sp[k]
        xd = splu(s.scipy().tocsc()).solve(s.b)
        x = out["x"][row_off[k]:row_off[k + 1]]
        errs.append(float(np.linalg.norm(x - xd) / np.linalg.norm(xd)))
    cpu = None
    try:
        ks = idx[:4]
        code = f"""
import json, sys, time
sys.path.insert(0, {ROOT!r})
from fem_amd import problems as P
from oracle.bind import Ref
ref = Ref(); ref.L.refh_init(2)
sp = P.batch_specimens({count})
tot = 0.0; n = 0
for k in {ks!r}:
    s = sp[k]
    M = ref.matrix(s.rowptr, s.col, s.val)
    o = ref.pcg_amg(M, s.b, dof=2, depth=5, smiter=1, aggr_smiter=1, smoother="ssor", alpha=1.0, amg_eps=1e-2,
                    amg_max_iter=1, zero_init=True, tol=1e-8, max_iter=5000)
    tot += o["t_solve"]; n += s.n
print(json.dumps({{"t": tot, "n": n}}))
"""
        c = _cpu_run(code)
        per_dof = c["t"] / c["n"]
        cpu = {"value": 1.0 / (per_dof * row_off[-1] / count), "unit": "solves/s", "cores": 1, "kind": "reference",
               "sample": f"reference pcg + AMG_solve (one symmetric SOR V(1,1), relres 1e-8) on specimens {ks} "
                         f"({c['n']} DOF, {c['t']:.2f}s); solves/s of the whole batch at that time per DOF"}
    except Exception as e:
        log(f"cpu baseline skipped: {e}")
    return {"metric": "batched AMG-PCG solves/s (512 specimens, relres 1e-8)", "value": count / (ms * 1e-3),
            "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "dtype": "f32 hierarchy / f64 outer", "data": "synthetic (seeded specimens)",
            "config": {"workload": "512 degraded-elasticity specimens, sides U{32..180}, lognormal(0,0.5) E",
                       "dof": int(row_off[-1]), "nnz": int(len(ci)), "depth": 5, "graph_iters": 10},
            "setup_s": t_setup, "iters": {"min": int(out["iters"].min()), "mean": float(out["iters"].mean()),
                                          "max": int(out["iters"].max())},
            "max_relres": float(out["relres"].max()),
            "oracle_error": {"specimens": idx, "max": max(errs), "median": float(np.median(errs))},
            "cpu_baseline": cpu}


def workload_staggered(args, lina, problems):
    """The staggered loop's elastic step on the device: for a sequence of
    damage states (a notch growing by 0.01 per step, nodal factor
    (1-d)^2 + 1e-6, d = exp(-dist/0.01)) on the size x size mesh, the nodal
    moduli are in HBM when a step starts; a step re-assembles the stiffness
    and the Dirichlet-treated right-hand side from Crack's strain-gradient
    tensor (lina_assembly: the contraction of FEM2D's use_map, bit-identical),
    recomputes the AMG hierarchy from the device values (device Galerkin
    products) and solves with AMG-PCG (FP32 hierarchy, FP64 outer) to relres
    tol from the previous step's solution.  Beside it: the same recompute
    from host values, and the error of the last step against SuperLU."""
    from scipy.sparse.linalg import splu
    m = args.size
    mesh = problems.Mesh(m, m)
    n = mesh.n_nodes
    t0 = time.time()
    B = problems.strain_gradient_tensor(mesh)
    rp, ci = problems.elastic_pattern(mesh)
    log(f"mesh {m}x{m}: {2 * n} DOF, {len(ci)} nnz, tensor {int(B[0][-1])} triples, built in {time.time() - t0:.1f}s")
    t0 = time.time()
    asm = lina.Assembly(n, *B, rp, ci)
    bot, top = mesh.boundary("bottom"), mesh.boundary("top")
    flags = np.zeros(2 * n, np.uint8)
    flags[np.concatenate([bot, bot + n, top + n])] = 1
    u = np.zeros(2 * n)
    u[top + n] = 1e-3
    asm.set_dirichlet(flags)
    log(f"assembly plan: {asm.terms} products into {asm.nnz} entries, {time.time() - t0:.1f}s")
    dC, dv, du, db, dx = (lina.DeviceVector(k) for k in (16 * n, asm.nnz, 2 * n, 2 * n, 2 * n))
    du.upload(u)
    dx.fill(0.0)

    def moduli(step):
        a, bb = problems.notch_segment(problems.SEED_NOTCH, 0.3 + 0.01 * step)
        d = np.exp(-problems._dist_to_segment(mesh.xy, a, bb) / 0.01)
        return problems.nodal_moduli(n, (1.0 - d) ** 2 + 1e-6)

    dC.upload(moduli(0))
    db.fill(0.0)
    asm.values(dC.ptr, dv.ptr, du.ptr, db.ptr)
    t0 = time.time()
    prm = lina.params("vcycle", 2, args.depth, precision="mixed", autotune=1)
    amg = lina.AMG(rp, ci, dv.download(), prm)
    t_setup = time.time() - t0
    log(f"AMG setup + upload {t_setup:.1f}s")
    rows = []
    for k in range(1, 1 + args.warmup + args.steps):
        dC.upload(moduli(k))                    # the damage solve's output, in HBM before the step
        db.fill(0.0)
        t = time.perf_counter()
        asm.values(dC.ptr, dv.ptr, du.ptr, db.ptr)
        t1 = time.perf_counter()
        amg.recompute_device(dv.ptr)
        t2 = time.perf_counter()
        st = amg.pcg_device(dx.ptr, db.ptr, tol=args.tol, max_iter=20000, stop="relres")
        t3 = time.perf_counter()
        r = {"assembly_ms": (t1 - t) * 1e3, "recompute_ms": (t2 - t1) * 1e3, "solve_ms": (t3 - t2) * 1e3,
             "step_ms": (t3 - t) * 1e3, "iters": st["iters"]}
        log(f"step {k}: {r}")
        if k > args.warmup:
            rows.append(r)
    # the host-value recompute of the same values, for comparison
    hv = dv.download()
    t = time.perf_counter()
    amg.recompute(hv)
    t_host_rc = (time.perf_counter() - t) * 1e3
    # the last step's solution against a direct solve of the same system
    bh = db.download()
    import scipy.sparse as sps
    A = sps.csr_matrix((hv, ci, rp), shape=(2 * n, 2 * n))
    err = None
    if 2 * n <= 600_000:
        xd = splu(A.tocsc()).solve(bh)
        err = float(np.linalg.norm(dx.download() - xd) / np.linalg.norm(xd))
    mean = {k: float(np.mean([r[k] for r in rows])) for k in rows[0]}
    asm_bytes = 8 * (asm.nnz + 1) + asm.terms * (4 + 4 + 8 + 8) + 8 * asm.nnz
    return {"metric": "staggered elastic step on the device (re-assembly + AMG recompute + AMG-PCG)",
            "value": mean["step_ms"] / 1e3, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": mean["step_ms"], "higher_is_better": False, "dtype": "f32 hierarchy / f64 outer",
            "data": "synthetic (seeded notch growing by 0.01 a step)",
            "config": {"workload": f"degraded plane-strain elasticity, Crack's strain-gradient tensor x nodal "
                                   f"moduli, {m}x{m} nodes, AMG-PCG to relres {args.tol:g}",
                       "dof": 2 * n, "nnz": int(len(ci)), "tensor_triples": int(B[0][-1]), "depth": args.depth},
            "mean": mean, "host_value_recompute_ms": t_host_rc, "setup_s": t_setup, "err_vs_direct": err,
            "assembly_roofline": {"bytes_moved": asm_bytes, "achieved_gbs": asm_bytes / (mean["assembly_ms"] * 1e-3)
                                  / 1e9, "peak": HBM_PEAK_GBS,
                                  "bytes": "plan offsets + per product (moduli index, tensor index, gathered "
                                           "modulus, tensor value) + values written"},
            "cpu_baseline": None}


def workload_adjoint(args, lina, problems):
    """Adjoint gradient on 513^2 plane strain with theta = exp(GRF): forward
    AMG-PCG, adjoint solve, element gradient kernel; 20 central-difference
    directions (h = 1e-6) with device solves; FP64 CPU reference gradient."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from scipy.sparse.linalg import splu
    m = 513
    case = problems.adjoint_case(m)
    s, dofs = case.system(case.theta)
    so, _ = case.system(case.theta_obs)
    log(f"adjoint case n={s.n} elements={len(case.theta)}")
    t0 = time.time()
    amg = lina.AMG(s.rowptr, s.col, s.val, lina.params("vcycle", 2, 7, precision="mixed"))
    t_setup = time.time() - t0
    op_obs = lina.Matrix(so.rowptr, so.col, so.val)
    u_obs = amg.pcg_operator(op_obs, so.b, tol=1e-12, max_iter=20000)["x"]
    op_obs.close()
    times = []
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        fw = amg.pcg(s.b, tol=1e-12, max_iter=20000, stop="relres")
        u = fw["x"]
        bw = amg.elasticity_adjoint(case.mesh.xy, case.mesh.tri, case.nu, u, case.dloss(u, u_obs, dofs), tol=1e-12,
                                    max_iter=20000)
        if k >= args.warmup:
            times.append(time.perf_counter() - t0)
    grad = bw["grad"]
    # central differences with device solves (cancellation-free form)
    import adjoint_ref as R
    fd_err = []
    h = 1e-6
    for v in R.directions(len(case.theta), 20):
        r0, rv = R.fd_pieces(case, u, v)
        d = []
        for sg in (1, -1):
            sy, _ = case.system(case.theta + sg * h * v)
            op = lina.Matrix(sy.rowptr, sy.col, sy.val)
            d.append(amg.pcg_operator(op, r0 + sg * h * rv, tol=1e-12, max_iter=20000)["x"])
            op.close()
        fd = R.fd_quotient(case, d[0], d[1], u, u_obs, h)
        fd_err.append(abs(fd - grad @ v) / abs(grad @ v))
        log(f"FD direction {len(fd_err)}: rel err {fd_err[-1]:.2e}")
    log(f"FD max rel err {max(fd_err):.2e}; CPU reference gradient (SuperLU) ...")
    # FP64 CPU reference gradient (direct solves)
    t0 = time.time()
    lu = splu(s.scipy().tocsc())
    u_ref = lu.solve(s.b)
    uo_ref = splu(so.scipy().tocsc()).solve(so.b)
    lam_ref = lu.solve(case.dloss(u_ref, uo_ref, dofs))
    Ke = R.element_matrices(case.mesh, case.nu)
    ed = R.element_dofs(case.mesh)
    g_ref = -np.einsum("ei,eij,ej->e", lam_ref[ed], Ke, u_ref[ed])
    t_cpu = time.time() - t0
    gerr = float(np.linalg.norm(grad - g_ref) / np.linalg.norm(g_ref))
    ms = float(np.mean(times)) * 1e3
    return {"metric": "adjoint gradient time (forward + adjoint + dL/dtheta) @526k DOF", "value": ms / 1e3,
            "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": False, "dtype": "f32 hierarchy / f64 outer",
            "data": "synthetic (seeded GRF moduli and observations)",
            "config": {"workload": "P1 plane strain 513x513, theta = exp(GRF length 0.1 var 0.25), 5% observed nodes",
                       "dof": s.n, "elements": len(case.theta), "depth": 7, "tol": 1e-12},
            "setup_s": t_setup, "forward_iters": fw["iters"], "adjoint_iters": bw["iters"],
            "fd_check": {"directions": 20, "h": h, "max_rel_err": max(fd_err)},
            "cpu_reference_gradient": {"rel_err": gerr, "time_s": t_cpu,
                                       "kind": "scipy splu FP64 forward/adjoint + B^T C B element matrices"},
            "cpu_baseline": {"value": t_cpu, "unit": "s", "cores": 1, "kind": "port",
                             "sample": "FP64 CPU reference gradient (SuperLU direct solves) at full size"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=1025, help="nodes per side")
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-10, help="relative residual target of the timed solves")
    ap.add_argument("--no-cpu", action="store_true", help="skip the reference CPU baseline")
    ap.add_argument("--coarse-sweeps", type=int, default=1, help="l1-Jacobi sweeps on the levels below the finest")
    ap.add_argument("--skip-levels", default="1,2,4",
                    help="levels (0 = coarsest) the V-cycle passes through without smoothing; '' for none")
    ap.add_argument("--kernels-only", action="store_true",
                    help="no solves: a 1 GiB device copy and a few finest-level SpMVs, for counter runs")
    ap.add_argument("--workload", default="elasticity",
                    choices=["elasticity", "poisson", "damage", "batch", "adjoint", "staggered"],
                    help="elasticity: the headline 2.1M-DOF AMG-PCG (default); poisson: 257^2 Poisson AMG-PCG "
                         "beside the reference on the same system; damage: reflective Newton on the 1025^2 damage "
                         "subproblem; batch: 512 batched specimens; adjoint: 513^2 adjoint gradient; staggered: "
                         "the elastic step re-assembled, recomputed and solved on the device")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); a row-block sharded solve "
              f"is not built", file=sys.stderr)
        sys.exit(2)

    from fem_amd import lina, problems

    if args.workload != "elasticity":
        out = {"poisson": workload_poisson, "damage": workload_damage, "batch": workload_batch,
               "adjoint": workload_adjoint, "staggered": workload_staggered}[args.workload](
            args, lina, problems)
        print(json.dumps(out))
        return

    t0 = time.time()
    s = problems.degraded_elasticity(args.size)
    log(f"system n={s.n} nnz={s.nnz} built in {time.time() - t0:.1f}s")
    if args.kernels_only:
        kernels_only(s, lina)
        return
    skip = [int(v) for v in args.skip_levels.split(",") if v.strip()]
    if any(not 0 < l < args.depth - 1 for l in skip):
        sys.exit(f"--skip-levels: levels must lie in 1..{args.depth - 2}")
    mask = sum(1 << l for l in set(skip))
    t0 = time.time()
    prm64 = lina.params("vcycle", 2, args.depth, coarse_sweeps=args.coarse_sweeps, autotune=1, skip_levels=mask)
    amg = lina.AMG(s.rowptr, s.col, s.val, prm64)
    t_setup = time.time() - t0
    sizes = [amg.level(l)["fine"] for l in range(amg.depth)] if amg.depth < 12 else []
    tuned = amg.tuned()
    lpr_top = [t for t in tuned if t[0] == amg.depth - 1 and t[1] == "A"]
    lpr64 = lpr_top[0][2] if lpr_top else 1
    log(f"AMG setup + upload {t_setup:.1f}s, level sizes {sizes}, autotuned row partitions {tuned}")
    prm32 = lina.params("vcycle", 2, args.depth, precision="mixed", coarse_sweeps=args.coarse_sweeps, autotune=1,
                        skip_levels=mask)

    dx = lina.DeviceVector(s.n)
    db = lina.DeviceVector(s.n)
    db0 = lina.DeviceVector(s.n)
    db0.upload(s.b)

    def solve(prm, tol, max_iter=20000):
        amg.set_params(prm)
        dx.fill(0.0)
        db.copy_from(db0)
        t = time.perf_counter()
        st = amg.pcg_device(dx.ptr, db.ptr, tol=tol, max_iter=max_iter, stop="relres")
        st["wall_ms"] = (time.perf_counter() - t) * 1e3
        return st

    A = s.scipy()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import direct_oracle
    have_direct = direct_oracle.fixture(args.size) is not None

    res = {}
    xs = {}
    amg.kernel_timing(True)
    for name, prm in (("mixed", prm32), ("fp64", prm64)):
        for _ in range(args.warmup):
            solve(prm, args.tol)
        runs = [solve(prm, args.tol) for _ in range(args.steps)]
        x = dx.download()
        xs[name] = x
        rr = float(np.linalg.norm(s.b - A @ x) / np.linalg.norm(s.b))
        nk = sum(r["spmv_launches"] for r in runs)
        res[name] = {"ms": float(np.mean([r["wall_ms"] for r in runs])),
                     "device_ms": float(np.mean([r["t_solve_ms"] for r in runs])),
                     "spmv_ms": float(sum(r["spmv_ms"] * r["spmv_launches"] for r in runs) / max(nk, 1)),
                     "spmv_launches": nk,
                     "iters": runs[-1]["iters"], "vcycles": runs[-1]["vcycles"], "relres": rr}
        log(f"{name}: {res[name]}")
    # the same solve with every level smoothed, beside the pass-through cycle
    full = None
    if mask:
        prm_full = lina.params("vcycle", 2, args.depth, precision="mixed", coarse_sweeps=args.coarse_sweeps,
                               autotune=1)
        r = solve(prm_full, args.tol)
        full = {"ms": r["wall_ms"], "iters": r["iters"]}
        log(f"mixed with every level smoothed: {full}")
    # error against SuperLU's FP64 direct solution (committed sketch)
    oracle = None
    if have_direct:
        errs, oracle = direct_oracle.errors(s, [xs["mixed"], xs["fp64"]], args.size)
        for name, e in zip(("mixed", "fp64"), errs):
            res[name]["err"] = e["sketch"]
            res[name]["err_sampled"] = e["sample"]
            res[name]["err_sampled_max"] = e["sample_max"]
            log(f"{name} error vs direct solve: {e}")
        oracle = {"kind": oracle["kind"] + " of the same system (tests/golden/direct_elasticity%d.npz)" % args.size,
                  "true_relres": oracle["relres"],
                  "err": "|S x - S x_direct| / |x_direct| with a seeded 64-row Gaussian S (estimate, +-18 %); "
                         "err_sampled: the relative error over 20,000 seeded entries"}

    # finest-level SpMV against the HBM roofline: the outer-loop A*p launches
    # of the timed mixed solves (every fourth), timed by HIP events that the
    # launch records at the kernel's own start and end on the library stream;
    # beside it the same kernel back to back.  Algorithmic bytes:
    # 12 nnz + 4 (n+1) + 16 n (what the kernel stores and moves instead,
    # with 16-bit column codes, is stored_bytes_per_launch)
    lvl = amg.depth - 1
    stored_bytes = amg.spmv_bytes(lvl, "fp64")
    spmv_bytes = spmv_alg_bytes(s.nnz, s.n, s.n, "fp64")
    spmv_ms = res["mixed"]["spmv_ms"]
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    # the whole iteration against the same roofline
    for name in ("mixed", "fp64"):
        ib = iteration_bytes(amg, name, args.coarse_sweeps, mask)
        ms_it = res[name]["device_ms"] / max(res[name]["iters"], 1)
        res[name]["iteration"] = {"bytes": ib["total"], "ms": ms_it,
                                  "achieved_gbs": ib["total"] / (ms_it * 1e-3) / 1e9,
                                  "frac": ib["total"] / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "outer_bytes": ib["outer"], "level_bytes": ib["levels"],
                                  "coarse_bytes": ib["coarse"]}
        log(f"{name} iteration: {ib['total'] / 1e9:.3f} GB algorithmic, {ms_it:.3f} ms -> "
            f"{res[name]['iteration']['achieved_gbs']:.0f} GB/s ({res[name]['iteration']['frac']:.2f} of peak); "
            f"levels {[round(b / 1e6, 1) for b in ib['levels']]} MB, coarse {ib['coarse'] / 1e6:.1f} MB")
    b2b_ms = amg.time_spmv(lvl, "fp64", reps=100)
    spmv32_ms = amg.time_spmv(lvl, "mixed", reps=100)
    spmv32_gbs = amg.spmv_bytes(lvl, "mixed") / (spmv32_ms * 1e-3) / 1e9
    log(f"SpMV fp64 in-solve {spmv_ms * 1e3:.1f} us {spmv_gbs:.0f} GB/s, back-to-back {b2b_ms * 1e3:.1f} us; "
        f"fp32 back-to-back {spmv32_ms * 1e3:.1f} us {spmv32_gbs:.0f} GB/s")
    cgbs = copy_gbs(lina)
    log(f"device copy {cgbs:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        try:
            with _CStdoutToStderr():
                cpu = cpu_baseline(s, amg, full["iters"] if full else res["fp64"]["iters"])
            log(f"cpu baseline {cpu['value']:.1f}s ({time.time() - t0:.0f}s wall)")
        except Exception as e:  # the baseline never blocks the GPU line
            log(f"cpu baseline skipped: {e}")

    value = res["mixed"]["ms"] / 1e3
    out = {
        "metric": "AMG-PCG solve time @2.1M DOF, mixed vs FP64, err<=1e-8; SpMV HBM GB/s",
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["mixed"]["ms"],
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 hierarchy / f64 outer",
        "data": "synthetic (seeded SplitMix64 notch)",
        "config": {"workload": f"degraded plane-strain elasticity P1 {args.size}x{args.size} nodes, "
                               f"AMG-PCG to relres {args.tol:g}",
                   "dof": s.n, "nnz": s.nnz, "depth": args.depth, "levels": sizes,
                   "smoother": f"multicolour SSOR on the finest level, {args.coarse_sweeps} l1-Jacobi sweeps below"
                               + (f"; levels {sorted(set(skip))} (0 = coarsest) passed through unsmoothed"
                                  if skip else ""),
                   "tol": args.tol},
        "mixed": res["mixed"],
        "fp64": res["fp64"],
        "speedup_mixed_vs_fp64": res["fp64"]["ms"] / res["mixed"]["ms"],
        "mixed_every_level_smoothed": full,
        "oracle": oracle,
        "setup_s": t_setup,
        "roofline": {"bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": spmv_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": f"k_spmv<double, double, 3, {lpr64}> (finest-level A*p of the outer CG)",
                     "bytes_per_launch": spmv_bytes, "stored_bytes_per_launch": stored_bytes, "avg_ms": spmv_ms,
                     "launches": res["mixed"]["spmv_launches"],
                     "back_to_back_ms": b2b_ms, "back_to_back_gbs": spmv_bytes / (b2b_ms * 1e-3) / 1e9,
                     "copy_gbs": cgbs},
        "iteration_roofline": {"mixed": res["mixed"]["iteration"]["frac"], "fp64": res["fp64"]["iteration"]["frac"],
                               "bytes": "outer SpMV + 2 dots + fused updates + per level (sweeps, residual, R, P) "
                                        "+ coarse inverse, from the actual level sizes"},
        "spmv_fp32": {"achieved": spmv32_gbs, "avg_ms": spmv32_ms, "kernel": "k_spmv<float, float, 3, 1>"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
