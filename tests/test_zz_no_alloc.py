"""mwg_mlp_gradient_noalloc against mwg_mlp_gradient through the C ABI.

The allocation-free variant keeps its per-wave partial gradients in a
file-scope device array instead of library-allocated memory and runs the same
kernels in the same order, so its result must equal the allocating entry
point's bit for bit: checked on two calls in a row, the second with fewer
samples (the scratch keeps the first call's contents, which the second call
must overwrite), and the first call also against the float64 restatement
oracle/mlp.py within the FP32 tolerance the kernel tests use."""
import ctypes

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


class Dev:
    def __init__(self, mr):
        self.L = mr.lib()
        assert self.L.mwg_init(0) == 0
        self.L.mwg_mlp_gradient_noalloc.restype = ctypes.c_int
        self.L.mwg_mlp_gradient_noalloc.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_float, ctypes.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.mwg_malloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        assert self.L.mwg_upload(p, a.ctypes.data, a.nbytes) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n):
        out = np.zeros(n, dtype=np.float32)
        assert self.L.mwg_download(out.ctypes.data, p, out.nbytes) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.mwg_free(p)


def test_noalloc_gradient_equals_allocating_entry_point(mr):
    d = Dev(mr)
    rng = np.random.default_rng(29)
    p = rng.normal(0.0, 0.25, mlp.N_PARAMS).astype(np.float32)
    dp = d.put(p)
    first = None
    for n in (2113, 371):  # second call smaller than the first
        X = rng.normal(0.0, 1.0, (n, mlp.F)).astype(np.float32)
        Y = (rng.random(n) > 0.6).astype(np.float32)
        dx, dy = d.put(X), d.put(Y)
        g_ref = d.put(np.zeros(mlp.N_PARAMS + 1, np.float32))
        g_new = d.put(np.full(mlp.N_PARAMS + 1, np.nan, np.float32))
        assert d.L.mwg_mlp_gradient(dp, dx, dy, n, ctypes.c_float(2.5), g_ref) == 0
        assert d.L.mwg_mlp_gradient_noalloc(dp, dx, dy, n, ctypes.c_float(2.5), g_new) == 0, d.L.mwg_last_error()
        assert d.L.mwg_synchronize() == 0
        a, b = d.get(g_ref, mlp.N_PARAMS + 1), d.get(g_new, mlp.N_PARAMS + 1)
        assert np.array_equal(a, b), (n, float(np.abs(a - b).max()))
        ref = mlp.gradient(p, X, Y, 2.5)
        assert np.abs(b[:-1] - ref).max() / np.abs(ref).max() <= 1e-4
        if first is None:
            first = b
        else:
            assert not np.array_equal(first, b)  # the second call's own result, not the first's
    # argument errors are reported, not run
    assert d.L.mwg_mlp_gradient_noalloc(dp, None, None, 5, ctypes.c_float(1.0), g_new) != 0
    assert d.L.mwg_mlp_gradient_noalloc(dp, dx, dy, -1, ctypes.c_float(1.0), g_new) != 0
    d.free()
