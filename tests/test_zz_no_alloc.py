"""lina_dev_dot_noalloc, the allocation-free form of lina_dev_dot: the same
kernels with their block partials and result in the library's file-scope
device scratch.  Called through the C ABI on device vectors, it returns
exactly lina_dev_dot's value on two calls in a row, the second with a
smaller input (the scratch keeps the first call's partials: nothing of them
may leak into the second), and rejects a negative length."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_noalloc_dot_equals_dot(gpu):
    L = gpu.lib()
    f = L.lina_dev_dot_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_long]
    f.restype = C.c_double
    L.lina_dev_dot.argtypes = [C.c_void_p, C.c_void_p, C.c_long]
    L.lina_dev_dot.restype = C.c_double
    rng = np.random.default_rng(9)
    n1, n2 = 3_000_001, 777
    a = gpu.DeviceVector(n1)
    b = gpu.DeviceVector(n1)
    ha, hb = rng.normal(size=n1), rng.normal(size=n1)
    a.upload(ha)
    b.upload(hb)
    for n in (n1, n2, 0):
        say("noalloc dot n =", n)
        want = L.lina_dev_dot(C.c_void_p(a.ptr), C.c_void_p(b.ptr), n)
        got = f(C.c_void_p(a.ptr), C.c_void_p(b.ptr), n)
        say("  dot", want, "noalloc", got)
        assert got == want
        assert abs(got - float(ha[:n] @ hb[:n])) <= 1e-12 * max(1.0, float(np.abs(ha[:n]) @ np.abs(hb[:n])))
    assert math.isnan(f(C.c_void_p(a.ptr), C.c_void_p(b.ptr), -1))
    assert "n must lie" in L.lina_last_error().decode()
