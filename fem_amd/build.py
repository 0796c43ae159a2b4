"""Build liblina_hip.so in-tree with hipcc for gfx950.

    python -m fem_amd.build        (or fem_amd.build.build())

Every source is compiled with -ffp-contract=off (no fused multiply-add behind
the reference's back) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblina_hip.so")
# the LinA drop-in: C host code with LinA's entry points over liblina_hip's C ABI
HOST = os.path.join(HERE, "host")
DROPIN_LIB = os.path.join(HERE, "liblinalg_hip.so")
DROPIN_SOURCES = ["linalg.c", "active_set.c", "pcgs.c", "multigrid.c", "bicgstab.c", "reflective_newton.c", "sparse_io.c",
                  "eigen.c", "gmres.c", "direct.c", "sparse_ops.c", "sweeps.c", "lists.c", "sparse3d.c", "rbtree.c", "file_stuff.c",
                  "spd_solver.c", "checkpoint.c"]

SOURCES = ["amg_setup.cpp", "solver.cpp", "capi.cpp", "kernels.hip", "krylov.cpp", "rn.cpp", "batch.cpp",
           "noalloc.hip", "rap.cpp", "assembly.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LINA_ARCH", "gfx950")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fopenmp", "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def _compile(src: str, build_dir: str = BUILD, extra=()) -> str:
    obj = os.path.join(build_dir, src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, out: str = LIB, build_dir: str = BUILD, extra=()) -> str:
    """Compile every source (extra: additional hipcc flags) and link `out`."""
    os.makedirs(build_dir, exist_ok=True)
    srcs = list(SOURCES)
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir, extra), srcs))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs, "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    if out == LIB:
        build_dropin(verbose)
    return out


def build_dropin(verbose: bool = True) -> str:
    """gcc: fem_amd/host/*.c -> liblinalg_hip.so linked to liblina_hip.so"""
    cc = os.environ.get("CC", shutil.which("gcc") or "cc")
    srcs = [os.path.join(HOST, s) for s in DROPIN_SOURCES]
    tmp = DROPIN_LIB + ".tmp"
    cmd = [cc, "-O2", "-fPIC", "-shared", "-std=c99", "-fopenmp", "-D_POSIX_C_SOURCE=200809L", "-Wall", "-Wno-unused-function",
           "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "include", "linalg"),
           *srcs, "-o", tmp, "-L", HERE, "-llina_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-Bsymbolic", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr)
    os.replace(tmp, DROPIN_LIB)
    if verbose:
        print(f"built {DROPIN_LIB}")
    return DROPIN_LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
