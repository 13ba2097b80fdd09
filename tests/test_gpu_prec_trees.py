"""apply with precision trees (src/treebuilders/apply.h:43,
src/treebuilders/apply.cpp:214-251) on the MI355X against the CPU oracle.

Bar as the plain apply: identical output node set and applied-tuple count,
coefficients within relative L2 1e-12 of the oracle.  An empty tree list
reproduces the plain apply bit for bit; an output tree among the precision
trees is rejected.
"""
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu

BOX, ORDER, PREC = 16, 5, 1e-4


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _pair(mr, orc, mra, omra, src, scale=1.0, prec=PREC):
    s = W.GaussianSources([w * scale for w in src.weights], src.exponents, src.centers)
    o = orc.OracleTree(omra)
    o.project_gaussians(prec, s.coefs, s.exponents, s.centers)
    p = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(s)
    mr.build_grid(p, gx)
    mr.project(prec, p, gx)
    return p, o


@pytest.fixture(scope="module")
def trees(mr, orc):
    mra = mr.MultiResolutionAnalysis((-BOX, BOX), ORDER)
    omra = orc.OracleMRA((-BOX, BOX), ORDER)
    f = _pair(mr, orc, mra, omra, W.random_sources(5, 3, center_range=3.0, beta_range=(2.0, 80.0)))
    h = _pair(mr, orc, mra, omra, W.random_sources(17, 2, center_range=5.0, beta_range=(0.5, 20.0)), scale=1e3)
    k = _pair(mr, orc, mra, omra, W.random_sources(23, 4, center_range=6.0, beta_range=(0.5, 50.0)))
    return dict(mra=mra, omra=omra, f=f, h=h, k=k)


def _check(mr, orc, t, names, mu=None):
    mra, omra = t["mra"], t["omra"]
    P = mr.PoissonOperator(mra, PREC) if mu is None else mr.HelmholtzOperator(mra, mu, PREC)
    op = orc.OracleOperator(omra, PREC, mu=mu)
    g, og = mr.FunctionTree(mra), orc.OracleTree(omra)
    st = mr.apply(PREC, g, P, t["f"][0], [t[n][0] for n in names])
    ost = orc.apply(PREC, og, op, t["f"][1], prec_trees=[t[n][1] for n in names])
    oi, oc, _ = og.export()
    pi, pc, _ = g.export()
    _say("precTrees", names, "mu", mu, "nodes", len(pi), "oracle", len(oi), "applied", st["n_applied"])
    assert np.array_equal(oi, pi), "output node sets differ"
    assert st["n_applied"] == ost["n_applied"]
    rel = np.linalg.norm(pc - oc) / np.linalg.norm(oc)
    assert rel <= 1e-12, rel
    return g, st


@pytest.mark.parametrize("names,mu", [(["h"], None), (["f"], None), (["h", "k"], None), (["k"], 1.5)],
                         ids=["other_tree", "input_itself", "two_trees", "helmholtz"])
def test_prec_trees_match_oracle(mr, orc, trees, names, mu):
    _check(mr, orc, trees, names, mu)


def test_empty_list_is_the_plain_apply(mr, trees):
    mra = trees["mra"]
    P = mr.PoissonOperator(mra, PREC)
    f = trees["f"][0]
    a, b = mr.FunctionTree(mra), mr.FunctionTree(mra)
    sa = mr.apply(PREC, a, P, f)
    sb = mr.apply(PREC, b, P, f, [])
    ai, ac, _ = a.export()
    bi, bc, _ = b.export()
    assert np.array_equal(ai, bi) and np.array_equal(ac, bc)
    assert sa["n_applied"] == sb["n_applied"]


def test_local_precision_takes_effect(mr, trees):
    mra = trees["mra"]
    P = mr.PoissonOperator(mra, PREC)
    a, b = mr.FunctionTree(mra), mr.FunctionTree(mra)
    mr.apply(PREC, a, P, trees["f"][0])
    mr.apply(PREC, b, P, trees["f"][0], [trees["h"][0]])
    assert not np.array_equal(a.export()[0], b.export()[0])


def test_output_tree_cannot_be_a_precision_tree(mr, trees):
    mra = trees["mra"]
    P = mr.PoissonOperator(mra, PREC)
    g = mr.FunctionTree(mra)
    with pytest.raises(mr.MRCPPError):
        mr.apply(PREC, g, P, trees["f"][0], [g])
