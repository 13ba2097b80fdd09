"""Sweep precisions of the band sum on the MI355X against the FP64 CPU oracle.

FP64 on the matrix cores and FP64 on the vector ALUs are held to the FP64
bar (identical node set and applied-tuple count, relative L2 <= 1e-12).
FP32 sweeps and three-part split BF16 sweeps (FP64 band sum, FP64
screening and transforms) are held to the same node set and a relative L2
error of 10*prec and <= 1e-6 against the oracle; the measured error is
printed for every case.
"""
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

from test_gpu_apply import assert_parity, run_pair

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture
def sweep(mr):
    def set_(name):
        mr.set_sweep_precision(name)
    yield set_
    mr.set_sweep_precision("fp64")


def test_default_precision_is_fp64(mr):
    assert mr.sweep_precision() == "fp64"
    with pytest.raises(ValueError):
        mr.set_sweep_precision("fp16")


@pytest.mark.parametrize("precision", ["fp64", "fp64_vector"])
@pytest.mark.parametrize("key,mu", [("gauss1_k5", None), ("gauss16_k7", None), ("gauss1_k9_helmholtz", 1.0)])
def test_fp64_sweeps_match_oracle(mr, orc, sweep, precision, key, mu):
    w = W.WORKLOADS[key]
    _say("case", key, mu, precision)
    sweep(precision)
    r = run_pair(mr, orc, w["sources"](), w["order"], w["box"], w["prec"], mu)
    rel = assert_parity(r, w["prec"])
    _say(key, precision, "rel L2", rel)


def reduced_precision_error(r):
    oi, oc, on = r["og"].export()
    pi, pc, pn = r["g"].export()
    same = np.array_equal(oi, pi)
    if same:
        rel = np.linalg.norm(pc - oc) / np.linalg.norm(oc)
    else:
        # compare on the common nodes when a split decision near its threshold flipped
        ko = {tuple(x): i for i, x in enumerate(oi)}
        common = [(ko[tuple(x)], j) for j, x in enumerate(pi) if tuple(x) in ko]
        a = np.array([oc[i] for i, _ in common])
        b = np.array([pc[j] for _, j in common])
        rel = np.linalg.norm(b - a) / np.linalg.norm(a)
    return same, rel, len(oi), len(pi)


@pytest.mark.parametrize("key,mu", [("gauss1_k5", None), ("gauss16_k7", None), ("gauss1_k9_poisson", None),
                                    ("gauss1_k9_helmholtz", 1.0)])
def test_fp32_sweeps_error(mr, orc, sweep, key, mu):
    w = W.WORKLOADS[key]
    _say("case", key, mu, "fp32")
    sweep("fp32")
    r = run_pair(mr, orc, w["sources"](), w["order"], w["box"], w["prec"], mu)
    same, rel, no, ng = reduced_precision_error(r)
    _say(key, "fp32", "same node set", same, "nodes", no, ng, "rel L2 vs FP64 oracle", rel)
    assert rel <= 10 * w["prec"]
    assert rel <= 1e-6
    assert same


@pytest.mark.parametrize("key,mu", [("gauss1_k5", None), ("gauss16_k7", None), ("gauss1_k9_poisson", None),
                                    ("gauss1_k9_helmholtz", 1.0)])
def test_bf16x3_sweeps_error(mr, orc, sweep, key, mu):
    """Three-part split BF16 (x = hi + mid + lo, six partial products formed
    exactly and summed in FP32): held to the same bar as the FP32 sweeps --
    identical node set, relative L2 <= 10*prec and <= 1e-6 against the FP64
    oracle."""
    w = W.WORKLOADS[key]
    _say("case", key, mu, "bf16x3")
    sweep("bf16x3")
    r = run_pair(mr, orc, w["sources"](), w["order"], w["box"], w["prec"], mu)
    same, rel, no, ng = reduced_precision_error(r)
    _say(key, "bf16x3", "same node set", same, "nodes", no, ng, "rel L2 vs FP64 oracle", rel)
    assert rel <= 10 * w["prec"]
    assert rel <= 1e-6
    assert same
