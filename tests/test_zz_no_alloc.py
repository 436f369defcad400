"""nqs_hip_cholesky_solve_noalloc: the allocation-free variant of
nqs_hip_cholesky_solve (device pointers in and out, every work buffer in
the library's static device scratch).  Two calls in a row, the second
smaller, each checked bit for bit against the allocating entry point on the
same matrix (the same kernels run on the same data), a rank-deficient
matrix that takes the reference's retry path, and a size beyond the scratch
refused with a message.  The test's own device buffers come from the HIP
runtime the library is linked against (the test may allocate; the entry
point does not)."""
import ctypes

import numpy as np
import pytest

from spin_based_neural_network_amd.abi import ptr

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def rt(hip):
    lib = ctypes.CDLL("libamdhip64.so.7")     # already loaded with libnqs_hip.so
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


class DevArray:
    def __init__(self, rt, host):
        self.rt, self.n = rt, host.size
        self.p = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.p), 8 * max(self.n, 1)) == 0
        h = np.ascontiguousarray(host, dtype=np.float64)
        assert rt.hipMemcpy(self.p, ptr(h), 8 * self.n, H2D) == 0

    def get(self, shape):
        out = np.zeros(shape)
        assert self.rt.hipMemcpy(ptr(out), self.p, 8 * self.n, D2H) == 0
        return out

    def free(self):
        self.rt.hipFree(self.p)


def _both(hip, rt, A, b):
    n = A.shape[0]
    Ah, xh = A.copy(), np.zeros(n)
    ok_h = hip.nqs_hip_cholesky_solve(ptr(Ah), n, ptr(b), ptr(xh))
    dT, db, dx = DevArray(rt, A), DevArray(rt, b), DevArray(rt, np.full(n, 7.0))
    ok_d = hip.nqs_hip_cholesky_solve_noalloc(dT.p, n, db.p, dx.p)
    Td, xd = dT.get((n, n)), dx.get(n)
    for d in (dT, db, dx):
        d.free()
    return ok_h, Ah, xh, ok_d, Td, xd


def test_noalloc_cholesky_equals_allocating_entry_point(hip, rt, capfd):
    rng = np.random.default_rng(17)
    for n in (1500, 700):          # two calls in a row, the second smaller
        X = rng.standard_normal((n, n + 40))
        A = np.ascontiguousarray(X @ X.T / (n + 40) + 1e-3 * np.eye(n))
        b = np.ascontiguousarray(rng.standard_normal(n))
        print(f"[noalloc] n={n}", flush=True)
        ok_h, Ah, xh, ok_d, Td, xd = _both(hip, rt, A, b)
        assert ok_h == ok_d == 1
        np.testing.assert_array_equal(xd, xh)
        np.testing.assert_array_equal(np.tril(Td), np.tril(Ah))
        np.testing.assert_allclose(A @ xd, b, atol=1e-7)
    # rank-deficient: the retry path (repair from the scratch copy, +1e-6)
    n = 600
    Y = rng.standard_normal((n, 450))
    A = np.ascontiguousarray(Y @ Y.T / n)
    b = np.ascontiguousarray(rng.standard_normal(n))
    print(f"[noalloc] rank-deficient n={n}", flush=True)
    ok_h, Ah, xh, ok_d, Td, xd = _both(hip, rt, A, b)
    assert ok_h == ok_d
    np.testing.assert_array_equal(xd, xh)
    np.testing.assert_array_equal(np.tril(Td), np.tril(Ah))
    # beyond the scratch: refused before any device work
    dummy = DevArray(rt, np.zeros(4))
    assert hip.nqs_hip_cholesky_solve_noalloc(dummy.p, 4096, dummy.p, dummy.p) == -1
    assert "needs" in capfd.readouterr().err
    dummy.free()
