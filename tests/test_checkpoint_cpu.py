"""Crack's checkpoint file data.bin (Crack/crack.c:5095-5287) through the
drop-in's C functions and the numpy module: the byte layout spelled out
with struct, files written by either side read by the other, record
selection (first, last, the one before), mesh sizes from .node files, and
the failures (index out of range, a truncated file)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from fem_amd import checkpoint as ck
from tests.dropin_bind import DROPIN_SO

pytestmark = pytest.mark.skipif(not os.path.exists(DROPIN_SO), reason="drop-in not built")

DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int)


@pytest.fixture(scope="module")
def L():
    L = C.CDLL(DROPIN_SO)
    L.lina_checkpoint_init.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p)]
    L.lina_checkpoint_append.argtypes = [C.c_char_p, C.c_int, DP, C.c_int, C.c_double, C.c_int]
    L.lina_checkpoint_info.argtypes = [C.c_char_p, IP, IP]
    L.lina_checkpoint_mesh_sizes.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(IP), IP]
    L.lina_checkpoint_read.argtypes = [C.c_char_p, IP, C.c_int, C.c_int, C.POINTER(DP), DP, IP, IP]
    return L


def _records():
    rng = np.random.default_rng(7)
    sizes = [5, 9]
    recs = [(0.0, 1, rng.normal(size=3 * 5)), (0.25, 1, rng.normal(size=3 * 5)), (0.5, 2, rng.normal(size=3 * 9))]
    return sizes, recs


def test_layout_and_c_writer(L, tmp_path):
    sizes, recs = _records()
    p = str(tmp_path / "data.bin").encode()
    names = (C.c_char_p * 3)(b"ux", b"uy", b"damage")
    assert L.lina_checkpoint_init(p, 3, names) == 1
    for t, mi, x in recs:
        assert L.lina_checkpoint_append(p, 3, x.ctypes.data_as(DP), len(x), t, mi) == 1
    want = struct.pack("=ii", 3, len(recs)) + b"".join(struct.pack("=di", t, mi) + x.tobytes() for t, mi, x in recs)
    assert open(p, "rb").read() == want
    assert open(p + b".info").read() == "3\nux\nuy\ndamage\n"
    got = ck.read_all(p.decode(), sizes)
    assert [(t, mi) for t, mi, _ in got] == [(t, mi) for t, mi, _ in recs]
    assert all(np.array_equal(a[2], b[2]) for a, b in zip(got, recs))
    n, m = C.c_int(), C.c_int()
    assert L.lina_checkpoint_info(p, C.byref(n), C.byref(m)) == 1 and (n.value, m.value) == (3, 3)


def test_numpy_writer_c_reader(L, tmp_path):
    sizes, recs = _records()
    p = str(tmp_path / "data.bin")
    ck.init(p, 3)
    for t, mi, x in recs:
        ck.append(p, 3, x, t, mi)
    assert ck.info(p) == (3, 3)
    sz = (C.c_int * 2)(*sizes)
    for index, want in ((0, 0), (1, 1), (-1, 2), (-2, 1)):
        X, t, size, mi = DP(), C.c_double(), C.c_int(), C.c_int()
        assert L.lina_checkpoint_read(p.encode(), sz, 2, index, C.byref(X), C.byref(t), C.byref(size),
                                      C.byref(mi)) == 1
        x = np.ctypeslib.as_array(X, shape=(size.value,)).copy()
        assert (t.value, mi.value) == recs[want][:2] and np.array_equal(x, recs[want][2])
    X, t, size, mi = DP(), C.c_double(), C.c_int(), C.c_int()
    assert L.lina_checkpoint_read(p.encode(), sz, 2, 3, C.byref(X), C.byref(t), C.byref(size), C.byref(mi)) == 0
    with open(p, "r+b") as f:                    # cut the last record short
        f.truncate(os.path.getsize(p) - 8)
    assert L.lina_checkpoint_read(p.encode(), sz, 2, -1, C.byref(X), C.byref(t), C.byref(size), C.byref(mi)) == 0
    with pytest.raises(ValueError):
        ck.read_all(p, sizes)


def test_mesh_sizes_from_node_files(L, tmp_path):
    os.makedirs(tmp_path / "meshes")
    for k, (nodes, extra) in enumerate(((1089, " 2 0 1"), (4225, " 2 0 1"), (66049, ""))):
        (tmp_path / "meshes" / f"notch.{k + 1}.node").write_text(f"{nodes}{extra}\n0 0.0 0.0\n")
    arr, cnt = IP(), C.c_int()
    assert L.lina_checkpoint_mesh_sizes(str(tmp_path).encode(), b"notch", C.byref(arr), C.byref(cnt)) == 1
    assert [arr[i] for i in range(cnt.value)] == [1089, 4225, 66049] == ck.mesh_sizes(str(tmp_path), "notch")
    assert L.lina_checkpoint_mesh_sizes(str(tmp_path).encode(), b"other", C.byref(arr), C.byref(cnt)) == 0
