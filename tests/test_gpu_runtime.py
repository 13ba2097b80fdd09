"""Device-layer transfers through the C ABI.

mwg_download_many brings several device ranges back with one wait: every
range must arrive intact, for the single-launch form (up to four ranges of
whole words within 1 MiB in all), for more ranges, for odd byte counts and
for ranges past 1 MiB (which fall back to one download each).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Dev:
    def __init__(self, mr):
        self.L = mr.lib()
        assert self.L.mwg_init(0) == 0
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.mwg_malloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        if a.nbytes:
            assert self.L.mwg_upload(p, a.ctypes.data, a.nbytes) == 0
        self.bufs.append(p)
        return p

    def free(self):
        for p in self.bufs:
            self.L.mwg_free(p)


def _download_many(d, srcs, arrays):
    n = len(arrays)
    outs = [np.zeros_like(a) for a in arrays]
    dst = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    src = (ctypes.c_void_p * n)(*[s.value for s in srcs])
    nbytes = (ctypes.c_size_t * n)(*[a.nbytes for a in arrays])
    rc = d.L.mwg_download_many(n, dst, src, nbytes)
    assert rc == 0, d.L.mwg_last_error()
    return outs


@pytest.mark.parametrize("shapes", [
    [(9 * 300, np.float64), (4, np.float64), (300, np.int32), (3076, np.int32)],  # one launch
    [(5, np.float64)] * 6,                                                      # more ranges than one launch takes
    [(7, np.uint8), (1, np.float64)],                                           # odd byte counts
    [(200_000, np.float64), (10, np.int32)],                                    # past 1 MiB in all
    [(0, np.float64), (3, np.int64)],                                           # an empty range
])
def test_download_many_returns_every_range(mr, shapes):
    d = Dev(mr)
    rng = np.random.default_rng(len(shapes))
    try:
        arrays = []
        for n, dt in shapes:
            if np.issubdtype(dt, np.floating):
                arrays.append(rng.normal(size=n).astype(dt))
            else:
                arrays.append(rng.integers(0, 100, n).astype(dt))
        srcs = [d.put(a) for a in arrays]
        outs = _download_many(d, srcs, arrays)
        for a, o in zip(arrays, outs):
            assert np.array_equal(a, o)
    finally:
        d.free()
