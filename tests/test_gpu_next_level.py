"""mwg_next_level through the C ABI against its documented rule.

The kernel forms a finished level's split decisions and the next level's
nodes with the host's operations in the host's order (TreeBuilder::build's
norm bookkeeping and WaveletAdaptor::splitNode, src/treebuilders/
TreeBuilder.cpp:48-80, src/treebuilders/WaveletAdaptor.h:45-56).  The rule
is restated here with Python floats (IEEE doubles, every product and sum
rounded on its own, math.sqrt correctly rounded), so every output must agree
bit for bit: the tree norm, the next screening threshold, the wavelet sum,
the split flags, and the children's ids, translations and thresholds.  Edge
cases: several chunks of the workgroup, the largest level, negative (unset)
component norms, absolute precision, zero precision, no split allowed and a
negative scaling norm.  The end-to-end check (trees equal to the host's
split) is in test_gpu_level_ahead.py.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Args(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int), ("may_split", ctypes.c_int), ("abs_prec", ctypes.c_int),
                ("prec", ctypes.c_double), ("s_norm", ctypes.c_double), ("w_norm", ctypes.c_double),
                ("scale_fac", ctypes.c_double), ("n_terms", ctypes.c_double), ("thr_floor", ctypes.c_double),
                ("base_id", ctypes.c_longlong), ("ncoef", ctypes.c_longlong),
                ("g_norms", ctypes.c_void_p), ("l", ctypes.c_void_p), ("head", ctypes.c_void_p), ("split", ctypes.c_void_p),
                ("next_off", ctypes.c_void_p), ("next_l", ctypes.c_void_p), ("next_thr", ctypes.c_void_p), ("next_count", ctypes.c_void_p)]


class Dev:
    def __init__(self, mr):
        self.L = mr.lib()
        assert self.L.mwg_init(0) == 0
        self.L.mwg_last_error.restype = ctypes.c_char_p
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.mwg_malloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        if a.nbytes:
            assert self.L.mwg_upload(p, a.ctypes.data, ctypes.c_size_t(a.nbytes)) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype=dtype)
        if out.nbytes:
            assert self.L.mwg_download(out.ctypes.data, p, ctypes.c_size_t(out.nbytes)) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.mwg_free(p)


def host_rule(norms, l, s_norm, w_norm, prec, abs_prec, scale_fac, n_terms, floor, may, base, ncoef):
    ws, step = [], 0.0
    for row in norms:
        w = 0.0
        for t in range(1, 8):
            v = float(row[t])
            if v >= 0.0:
                w += v * v
            else:
                w = -1.0
        ws.append(w)
        step += w
    wn = w_norm + step
    tsn = -1.0 if (s_norm < 0.0 or wn < 0.0) else s_norm + wn
    tn = math.sqrt(tsn) if (tsn > 0.0 and not abs_prec) else 1.0
    x = prec * tn * scale_fac
    thr = x if floor < x else floor
    gthr = prec * 1.0 * math.sqrt(tsn / n_terms) if tsn > 0.0 else tsn
    split = [1 if (may and prec > 0.0 and w >= 0.0 and math.sqrt(w) > thr) else 0 for w in ws]
    off, nl, nthr = [], [], []
    for i, sp in enumerate(split):
        if not sp:
            continue
        for c in range(8):
            j = len(off)
            off.append((base + j) * ncoef)
            nl.append([2 * int(l[i][d]) + ((c >> d) & 1) for d in range(3)])
            nthr.append(gthr)
    return dict(head=[tsn, gthr, wn, float(len(off))], split=split, off=off, l=nl, thr=nthr)


def run_case(mr, norms, l, **kw):
    d = Dev(mr)
    try:
        L = len(norms)
        a = Args()
        a.count, a.may_split, a.abs_prec = L, kw["may"], kw["abs_prec"]
        a.prec, a.s_norm, a.w_norm = kw["prec"], kw["s_norm"], kw["w_norm"]
        a.scale_fac, a.n_terms, a.thr_floor = kw["scale_fac"], 96.0, 2.0e-15
        a.base_id, a.ncoef = kw["base"], 4096
        a.g_norms = d.put(np.asarray(norms, dtype=np.float64).reshape(-1))
        a.l = d.put(np.asarray(l, dtype=np.int32).reshape(-1))
        a.head = d.put(np.full(4, np.nan))
        a.split = d.put(np.full(L, -7, dtype=np.int32))
        a.next_off = d.put(np.full(8 * L, -1, dtype=np.int64))
        a.next_l = d.put(np.full(24 * L, -1, dtype=np.int32))
        a.next_thr = d.put(np.full(8 * L, np.nan))
        a.next_count = d.put(np.full(1, -1, dtype=np.int32))
        rc = d.L.mwg_next_level(ctypes.byref(a))
        assert rc == 0, d.L.mwg_last_error()
        assert d.L.mwg_synchronize() == 0
        ref = host_rule(norms, l, kw["s_norm"], kw["w_norm"], kw["prec"], kw["abs_prec"], kw["scale_fac"], 96.0, 2.0e-15, kw["may"],
                        kw["base"], 4096)
        head = d.get(a.head, 4, np.float64)
        n = len(ref["off"])
        got = dict(head=head.tolist(), split=d.get(a.split, L, np.int32).tolist(), count=int(d.get(a.next_count, 1, np.int32)[0]),
                   off=d.get(a.next_off, n, np.int64).tolist(), l=d.get(a.next_l, 3 * n, np.int32).reshape(-1, 3).tolist(),
                   thr=d.get(a.next_thr, n, np.float64).tolist())
    finally:
        d.free()
    assert got["head"] == ref["head"], (got["head"], ref["head"])
    assert got["split"] == ref["split"]
    assert got["count"] == n
    assert got["off"] == ref["off"]
    assert got["l"] == ref["l"]
    assert got["thr"] == ref["thr"]
    return ref


def level(rng, L, split_frac, scale=3):
    """Component norms (8 + square norm) and translations of L nodes; about
    split_frac of them with wavelet norms above a threshold near 1e-5."""
    norms = np.zeros((L, 9))
    norms[:, 0] = rng.uniform(1e-3, 1.0, L)
    big = rng.random(L) < split_frac
    norms[:, 1:8] = np.where(big[:, None], rng.uniform(1e-4, 1e-2, (L, 7)), rng.uniform(0.0, 1e-7, (L, 7)))
    norms[:, 8] = (norms[:, :8] ** 2).sum(axis=1)
    l = rng.integers(-(1 << scale), 1 << scale, (L, 3))
    return norms, l


BASE = dict(may=1, abs_prec=0, prec=1e-5, s_norm=1.3, w_norm=0.2, scale_fac=2.0 ** (-(0.5 * 1.0 * (3 + 1))), base=777)


@pytest.mark.parametrize("L", [1, 63, 700, 4096])
def test_next_level_matches_the_rule(mr, L):
    rng = np.random.default_rng(L)
    norms, l = level(rng, L, 0.3)
    ref = run_case(mr, norms, l, **BASE)
    if L >= 63:
        assert 0 < len(ref["off"]) < 8 * L


def test_next_level_edge_cases(mr):
    rng = np.random.default_rng(5)
    norms, l = level(rng, 300, 0.4)
    norms[7, 3] = -1.0   # an unset component norm: that node's wavelet norm is -1 plus what follows
    norms[11, 7] = -1.0  # the last component unset: -1
    run_case(mr, norms, l, **BASE)
    run_case(mr, norms, l, **dict(BASE, abs_prec=1))
    assert not any(run_case(mr, norms, l, **dict(BASE, prec=0.0))["split"])
    assert not any(run_case(mr, norms, l, **dict(BASE, may=0))["split"])
    ref = run_case(mr, norms, l, **dict(BASE, s_norm=-1.0))
    assert ref["head"][0] == -1.0 and ref["head"][1] == -1.0


def test_next_level_rejects_more_than_4096_nodes(mr):
    d = Dev(mr)
    try:
        a = Args()
        a.count, a.n_terms, a.ncoef = 4097, 96.0, 4096
        for f in ("g_norms", "l", "head", "split", "next_off", "next_l", "next_thr", "next_count"):
            setattr(a, f, d.put(np.zeros(8)))
        assert d.L.mwg_next_level(ctypes.byref(a)) != 0
        assert b"4096" in d.L.mwg_last_error()
    finally:
        d.free()
