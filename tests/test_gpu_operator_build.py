"""Operator construction with the 2D operator trees formed on the MI355X
(CrossCorrelationCalculator per refinement level for all terms at once,
src/treebuilders/CrossCorrelationCalculator.cpp:36-96) against the CPU
oracle's construction (oracle/mrcpp_oracle.cpp build_operator) and against
the product's host construction.

Bar: bit-identical.  The kernel expansion, every cached operator node (matrix
and norm), the band widths at the apply precision, and the result of an apply
with either operator.
"""
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [("poisson_k5", 5, 20, 1e-5, None), ("helmholtz_k7", 7, 32, 1e-6, 1.5), ("poisson_k9", 9, 20, 1e-7, None)]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _build(mr, mra, prec, mu, target):
    mr.set_operator_target(target)
    try:
        return mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
    finally:
        mr.set_operator_target("auto")


def _same_operator(H, D, prec):
    """Every cached operator node (matrix and norm) and the band widths at prec
    identical; returns the number of nodes compared."""
    assert H.size() == D.size()
    assert np.array_equal(*(op.kernel()[0] for op in (H, D))) and np.array_equal(*(op.kernel()[1] for op in (H, D)))
    bh, bd = H.bandwidths(prec), D.bandwidths(prec)
    assert np.array_equal(bh, bd)
    checked = 0
    for term in range(H.size()):
        for depth in range(bh.shape[1]):
            w = int(bh[term, depth, 4])
            if w < 0:
                continue
            for l in range(-w - 1, w + 2):
                for comp in range(4):
                    mh, nh = H.node(term, depth, l, comp)
                    md, nd = D.node(term, depth, l, comp)
                    assert (mh is None) == (md is None)
                    if mh is None:
                        continue
                    assert np.array_equal(mh, md) and nh == nd, (term, depth, l, comp)
                    checked += 1
    return checked


@pytest.mark.parametrize("name,order,box,prec,mu", CASES, ids=[c[0] for c in CASES])
def test_device_operator_trees_are_bit_identical(mr, orc, name, order, box, prec, mu):
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    H = _build(mr, mra, prec, mu, "host")
    D = _build(mr, mra, prec, mu, "device")
    assert not H.built_on_device() and D.built_on_device()
    _say(name, "terms", D.size(), "build s host", H.build_seconds(), "device", D.build_seconds())
    # the device construction against the oracle's, node by node
    O = orc.OracleOperator(orc.OracleMRA((-box, box), order), prec, mu=mu)
    checked_oracle = _same_operator(O, D, prec)
    checked = _same_operator(H, D, prec)
    _say(name, "nodes compared: oracle", checked_oracle, "host", checked)
    assert checked_oracle > 100 and checked == checked_oracle


def test_batched_device_construction(mr, orc):
    """Poisson and Helmholtz operators of different precisions built in one
    device pass equal their own host constructions."""
    mra = mr.MultiResolutionAnalysis((-20, 20), 5)
    specs = [(None, 1e-5), (1.5, 1e-5), (0.7, 1e-6), (None, 1e-6), (2.9, 1e-4)]
    D = mr.build_operators(mra, specs, "device")
    assert len(D) == len(specs) and all(op.built_on_device() for op in D)
    ph = D[0].build_phases()
    _say("batched build s", D[0].build_seconds(), ph)
    omra = orc.OracleMRA((-20, 20), 5)
    for (mu, prec), d in zip(specs, D):
        h = _build(mr, mra, prec, mu, "host")
        assert isinstance(d, mr.PoissonOperator if mu is None else mr.HelmholtzOperator)
        assert _same_operator(h, d, prec) > 50
        assert _same_operator(orc.OracleOperator(omra, prec, mu=mu), d, prec) > 50


def test_apply_with_device_built_operator(mr):
    w = W.WORKLOADS["gauss1_k5"]
    mra = mr.MultiResolutionAnalysis((-w["box"], w["box"]), w["order"])
    gx = mr.GaussExp.from_sources(w["sources"]())
    f = mr.FunctionTree(mra)
    mr.build_grid(f, gx)
    mr.project(w["prec"], f, gx)
    out = []
    for target in ("host", "device"):
        P = _build(mr, mra, w["prec"], None, target)
        g = mr.FunctionTree(mra)
        st = mr.apply(w["prec"], g, P, f)
        out.append((g.export(), st["n_applied"]))
    (ih, ch, _), nh = out[0]
    (idd, cd, _), nd = out[1]
    assert np.array_equal(ih, idd) and np.array_equal(ch, cd) and nh == nd
