"""mdft_vec_multi_dot_noalloc: the multi-dot with its partial products in the
library's fixed device array, against mdft_vec_multi_dot (same kernels, memory
the library allocates), through the C-ABI, on two calls in a row, the second
with a smaller vector length; and its rejection of k beyond the array's 16
vectors.  Bar: equal to the last bit (the same reductions in the same order)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(eng, fn, n, vecs, x, k):
    arr = (C.c_void_p * k)(*[v.ptr for v in vecs[:k]])
    out = np.zeros(k)
    rc = fn(eng._h, C.c_size_t(n), k, arr, x.ptr, out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_multi_dot_matches_multi_dot():
    from mdft.gpu import DeviceBuffer, Engine
    eng = Engine((16, 16, 16), (12.0, 12.0, 12.0), 1, 2, 2.4777, 0.0333, "fp64")
    lib = eng.lib
    rng = np.random.default_rng(5)
    n1, n2 = 3_000_001, 777_777          # second call smaller (and odd: the tail element)
    k = 16
    host = [rng.standard_normal(n1) for _ in range(k)]
    hx = rng.standard_normal(n1)
    vecs = []
    for h in host:
        b = DeviceBuffer(eng, 8 * n1)
        b.upload(h)
        vecs.append(b)
    x = DeviceBuffer(eng, 8 * n1)
    x.upload(hx)
    vecs[3] = x                           # one of the vectors is x itself (|x|^2)
    for n, kk in ((n1, 16), (n2, 5)):
        rc_a, a = _call(eng, lib.mdft_vec_multi_dot_noalloc, n, vecs, x, kk)
        rc_b, b = _call(eng, lib.mdft_vec_multi_dot, n, vecs, x, kk)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b)
        ref = np.array([np.dot(hx[:n], (hx if j == 3 else host[j])[:n]) for j in range(kk)])
        print(f"n={n} k={kk}: noalloc {a[:3]} multi_dot {b[:3]} numpy {ref[:3]}", flush=True)
        assert np.array_equal(a, b)
        assert np.allclose(a, ref, rtol=1e-12, atol=1e-9)
    # k beyond the fixed array is refused with a message
    arr = (C.c_void_p * 17)(*([x.ptr] * 17))
    out = np.zeros(17)
    rc = lib.mdft_vec_multi_dot_noalloc(eng._h, C.c_size_t(n2), 17, arr, x.ptr,
                                        out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == -1
    assert b"16" in lib.mdft_last_error(eng._h)
