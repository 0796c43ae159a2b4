"""Crack's field checkpoint file data.bin in numpy (Crack/crack.c:5095-5287):
int n (fields per node), int m (records), then m records of double t,
int mesh_index (1-based) and n * mesh_size[mesh_index - 1] doubles, in the
machine's byte order with no padding.  The C side of the drop-in reads and
writes the same files (include/linalg/checkpoint.h)."""
import os

import numpy as np

_HDR = np.dtype([("n", np.int32), ("m", np.int32)])
_REC = np.dtype([("t", np.float64), ("mesh", np.int32)])


def init(path: str, n_fields: int, names=None) -> None:
    """an empty file (init_binary_save); names -> path + '.info'"""
    with open(path, "wb") as f:
        np.array([(n_fields, 0)], dtype=_HDR).tofile(f)
    if names is not None:
        with open(path + ".info", "w") as f:
            f.write(f"{n_fields}\n" + "".join(f"{s}\n" for s in names))


def append(path: str, n_fields: int, x, t: float, mesh_index: int) -> None:
    """one record at the end, the header's count + 1 (binary_save)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    with open(path, "r+b") as f:
        hdr = np.fromfile(f, dtype=_HDR, count=1)[0]
        f.seek(0)
        np.array([(n_fields, int(hdr["m"]) + 1)], dtype=_HDR).tofile(f)
        f.seek(0, os.SEEK_END)
        np.array([(t, mesh_index)], dtype=_REC).tofile(f)
        x.tofile(f)


def info(path: str):
    hdr = np.fromfile(path, dtype=_HDR, count=1)[0]
    return int(hdr["n"]), int(hdr["m"])


def read_all(path: str, mesh_sizes):
    """every record as (t, mesh_index, x)"""
    out = []
    with open(path, "rb") as f:
        hdr = np.fromfile(f, dtype=_HDR, count=1)[0]
        n, m = int(hdr["n"]), int(hdr["m"])
        for _ in range(m):
            rec = np.fromfile(f, dtype=_REC, count=1)
            if len(rec) != 1:
                raise ValueError(f"{path}: ends before its {m} records")
            mi = int(rec[0]["mesh"])
            size = n * int(mesh_sizes[mi - 1])
            x = np.fromfile(f, dtype=np.float64, count=size)
            if len(x) != size:
                raise ValueError(f"{path}: ends inside a record")
            out.append((float(rec[0]["t"]), mi, x))
    return out


def mesh_sizes(directory: str, mesh_base: str):
    """first token of <directory>/meshes/<mesh_base>.<k>.node, k = 1, 2, ..."""
    sizes, k = [], 1
    while True:
        p = os.path.join(directory, "meshes", f"{mesh_base}.{k}.node")
        if not os.path.exists(p):
            return sizes
        with open(p) as f:
            tok = f.readline().split()
        if tok:
            sizes.append(int(tok[0]))
        k += 1
