"""Host side of liblina_hip on the CPU: the AMG setup must equal the
reference's hierarchy bit for bit (golden vectors written by the compiled
reference), and the C ABI library must load and export every declared
symbol.  No device call is made here."""
import glob
import os
import re

import numpy as np
import pytest

from fem_amd import lina
from fem_amd import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the hierarchy fixtures (krylov.npz holds the Krylov drivers' vectors)
GOLDEN = sorted(p for p in glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")) if not p.endswith("krylov.npz")
          and not os.path.basename(p).startswith("direct_"))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_host_setup_matches_golden(path):
    g = np.load(path)
    depth = int(g["depth"])
    prm = lina.params("reference", int(g["dof"]), depth, coarsening=int(g["coarsening"]))
    amg = lina.AMG(g["rowptr"], g["col"], g["val"], prm, upload=False)
    assert amg.depth == depth
    for l in range(depth):
        lv = amg.level(l)
        for w in ("A", "P", "R"):
            if w not in lv:
                continue
            rp, ci, v = lv[w]
            assert np.array_equal(rp, g[f"L{l}_{w}_rp"]), (l, w)
            assert np.array_equal(ci, g[f"L{l}_{w}_ci"]), (l, w)
            assert np.array_equal(v, g[f"L{l}_{w}_v"]), (l, w)
        if l > 0:
            assert np.array_equal(lv["labels"], g[f"L{l}_labels"])
    amg.close()


def test_host_setup_matches_oracle_larger():
    from oracle.bind import Oracle
    orc = Oracle()
    for s, depth in ((P.degraded_elasticity(65), 5), (P.poisson(129), 5)):
        h = orc.amg_setup(s.rowptr, s.col, s.val, dof=s.dof, depth=depth)
        O = orc.amg_levels(h, depth)
        amg = lina.AMG(s.rowptr, s.col, s.val, lina.params("reference", s.dof, depth), upload=False)
        for l in range(depth):
            lv = amg.level(l)
            for w in ("A", "P", "R"):
                if w in lv:
                    assert np.array_equal(lv[w][1], O[l][w].col) and np.array_equal(lv[w][2], O[l][w].val)
            if l > 0:
                assert np.array_equal(lv["labels"], O[l]["labels"])
        amg.close()
        orc.amg_free(h)


def test_colouring_matches_oracle():
    from oracle.bind import Oracle
    orc = Oracle()
    s = P.degraded_elasticity(33)
    amg = lina.AMG(s.rowptr, s.col, s.val, lina.params("vcycle", 2, 4), upload=False)
    for l in range(1, 4):
        lv = amg.level(l)
        rp, ci, v = lv["A"]
        k, c = orc.colouring(rp, ci, v)
        assert k == lv["ncolours"] and np.array_equal(c, lv["colours"])


def test_bad_input_is_rejected():
    s = P.poisson(9)
    col = s.col.copy()
    col[3], col[4] = col[4], col[3]  # unsorted row
    with pytest.raises(lina.LinaError):
        lina.AMG(s.rowptr, col, s.val, lina.params("reference", 1, 3), upload=False)
    with pytest.raises(lina.LinaError):
        lina.AMG(s.rowptr, s.col, s.val, lina.params("reference", 0, 3), upload=False)


def _declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M):
        name = m.group(1)
        if m.group(0).lstrip().startswith("typedef"):  # function-pointer types
            continue
        if name not in ("if", "while", "for", "return", "sizeof", "typedef"):
            names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    import ctypes
    L = ctypes.CDLL(lina.LIB_PATH)
    missing = []
    for header in glob.glob(os.path.join(ROOT, "include", "*.h")):
        for name in _declared_functions(header):
            if not hasattr(L, name):
                missing.append((os.path.basename(header), name))
    assert not missing, missing


def test_recompute_matches_reference_and_oracle(ref):
    """AMG_recompute_coarse_matrices (Multigrid.c:1756-1780): new finest values
    with P and R kept; every level's A is bit-identical to the compiled
    reference and to the oracle."""
    from oracle.bind import Oracle
    s1 = P.damage_system(33, seed_h=11)[0]
    s2 = P.damage_system(33, seed_h=12)[0]
    assert np.array_equal(s1.rowptr, s2.rowptr) and np.array_equal(s1.col, s2.col)
    depth = 4
    M1 = ref.matrix(s1.rowptr, s1.col, s1.val)
    M2 = ref.matrix(s2.rowptr, s2.col, s2.val)
    I = ref.amg_setup(M1, dof=1, depth=depth)
    ref.L.refh_amg_recompute(I, M2)
    R = ref.amg_levels(I, depth)
    orc = Oracle()
    h = orc.amg_setup(s1.rowptr, s1.col, s1.val, dof=1, depth=depth)
    orc.L.orc_amg_recompute(h, s2.val.ctypes.data_as(orc_dp()))
    O = orc.amg_levels(h, depth)
    amg = lina.AMG(s1.rowptr, s1.col, s1.val, lina.params("reference", 1, depth), upload=False)
    amg.recompute(s2.val)
    changed = False
    for l in range(depth):
        rp, ci, v = amg.level(l)["A"]
        assert np.array_equal(ci, R[l]["A"].col) and np.array_equal(v, R[l]["A"].val), l
        assert np.array_equal(ci, O[l]["A"].col) and np.array_equal(v, O[l]["A"].val), l
        changed |= l < depth - 1 and not np.array_equal(v, lina.AMG(s1.rowptr, s1.col, s1.val, lina.params(
            "reference", 1, depth), upload=False).level(l)["A"][2])
    assert changed
    amg.close()
    orc.amg_free(h)
    ref.amg_free(I)
    ref.free(M1)
    ref.free(M2)


def orc_dp():
    import ctypes
    return ctypes.POINTER(ctypes.c_double)
