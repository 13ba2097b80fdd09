"""add and multiply (WaveletAdaptor and, with useMaxNorms, MultiplicationAdaptor
refinement) on the MI355X (src/treebuilders/add.cpp:43-80,
src/treebuilders/multiply.cpp:52-119) against the oracle's restatement, which
tests/test_arith_oracle.py pins to closed forms.

Inputs are projected on the host so both sides start from bit-identical trees.
Bar: identical node sets; coefficients within 1e-14 relative L2 of the
oracle's (both compute the nodes without fused multiply-adds in the same order;
the device's norm sums and bottom-up transform round differently).
"""
import math
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _pair(mr, orc, mra, omra, src, prec):
    mr.set_projection_target("host")
    try:
        f = mr.FunctionTree(mra)
        gx = mr.GaussExp.from_sources(src)
        mr.build_grid(f, gx)
        mr.project(prec, f, gx)
    finally:
        mr.set_projection_target("auto")
    of = orc.OracleTree(omra)
    of.project_gaussians(prec, src.coefs, src.exponents, src.centers)
    return f, of


def _compare(name, mine, ref, tol):
    mi, mc, mn = mine.export()
    oi, oc, on = ref.export()
    _say(name, "nodes", len(mi), len(oi))
    assert np.array_equal(mi, oi), "node sets differ"
    rel = np.linalg.norm(mc - oc) / max(np.linalg.norm(oc), 1e-300)
    _say(name, "rel L2", rel)
    assert rel <= tol
    assert np.allclose(mn, on, rtol=1e-12, atol=1e-14 * max(1.0, np.abs(on).max()))
    assert mine.getSquareNorm() == pytest.approx(ref.square_norm(), rel=1e-12)


CASES = [("k5", 5, 16, 1e-5), ("k7", 7, 32, 1e-6), ("k9", 9, 16, 1e-7)]


@pytest.mark.parametrize("name,order,box,prec", CASES, ids=[c[0] for c in CASES])
def test_add_and_multiply_match_the_oracle(mr, orc, name, order, box, prec):
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    omra = orc.OracleMRA((-box, box), order)
    f, of = _pair(mr, orc, mra, omra, W.random_sources(5, 4, center_range=2.0, beta_range=(1.0, 100.0)), prec)
    g, og = _pair(mr, orc, mra, omra, W.random_sources(6, 3, center_range=2.0, beta_range=(2.0, 300.0)), prec)
    h = mr.FunctionTree(mra)
    mr.add(prec, h, 1.5, f, -0.75, g)
    oh = orc.OracleTree(omra)
    orc.add(prec, oh, [(1.5, of), (-0.75, og)])
    _compare(name + " add", h, oh, 1e-14)
    p = mr.FunctionTree(mra)
    mr.multiply(prec, p, 2.0, f, g)
    op = orc.OracleTree(omra)
    orc.multiply(prec, op, [(2.0, of), (1.0, og)])
    _compare(name + " multiply", p, op, 1e-14)
    # the max-norm refinement (MultiplicationAdaptor, useMaxNorms = true)
    q = mr.FunctionTree(mra)
    mr.multiply(prec, q, 2.0, f, g, useMaxNorms=True)
    oq = orc.OracleTree(omra)
    orc.multiply(prec, oq, [(2.0, of), (1.0, og)], use_max_norms=True)
    _compare(name + " multiply useMaxNorms", q, oq, 1e-14)


def test_vector_forms_shared_inputs_and_device_resident_operands(mr, orc):
    """Three terms with one tree twice, a product of three, and an apply result
    (coefficients in HBM) as an operand."""
    w = W.WORKLOADS["gauss1_k5"]
    mra = mr.MultiResolutionAnalysis((-w["box"], w["box"]), w["order"])
    omra = orc.OracleMRA((-w["box"], w["box"]), w["order"])
    src = w["sources"]()
    f, of = _pair(mr, orc, mra, omra, src, w["prec"])
    g, og = _pair(mr, orc, mra, omra, W.random_sources(8, 2, center_range=1.0, beta_range=(5.0, 50.0)), w["prec"])
    h = mr.FunctionTree(mra)
    mr.add(w["prec"], h, [(1.0, f), (2.0, g), (-0.5, f)])
    oh = orc.OracleTree(omra)
    orc.add(w["prec"], oh, [(1.0, of), (2.0, og), (-0.5, of)])
    _compare("add3", h, oh, 1e-14)
    p = mr.FunctionTree(mra)
    mr.multiply(w["prec"], p, [(1.0, f), (1.0, g), (0.5, g)])
    op = orc.OracleTree(omra)
    orc.multiply(w["prec"], op, [(1.0, of), (1.0, og), (0.5, og)])
    _compare("mul3", p, op, 1e-14)
    # V = P f in HBM, then f V: its integral is the Coulomb self-energy sqrt(2 beta / pi)
    V = mr.FunctionTree(mra)
    mr.apply(w["prec"], V, mr.PoissonOperator(mra, w["prec"]), f)
    fv = mr.FunctionTree(mra)
    mr.multiply(w["prec"], fv, 1.0, f, V)
    beta = src.exponents[0]
    assert fv.integrate() == pytest.approx(math.sqrt(2 * beta / math.pi), rel=20 * w["prec"])
    assert fv.integrate() == pytest.approx(mr.dot(f, V), rel=1e-6)


def _np_dot(a, b, root_scale):
    """dot from exported trees: every node of a with a match in b, scaling
    part at the root scale, wavelet part everywhere."""
    (ai, ac, _), (bi, bc, _) = a.export(), b.export()
    kd = ac.shape[1] // 8
    where = {tuple(r[:4]): i for i, r in enumerate(bi)}
    tot = 0.0
    for i, r in enumerate(ai):
        j = where.get(tuple(r[:4]))
        if j is None:
            continue
        if r[0] == root_scale:
            tot += float(ac[i, :kd] @ bc[j, :kd])
        tot += float(ac[i, kd:] @ bc[j, kd:])
    return tot


def test_dot_and_integrate_of_trees_in_hbm(mr, orc):
    """dot and integrate of an apply result without copying it to the host:
    equal to the same sums formed from the exported coefficients."""
    w = W.WORKLOADS["gauss16_k7"]
    mra = mr.MultiResolutionAnalysis((-w["box"], w["box"]), w["order"])
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(w["sources"]())
    mr.build_grid(f, gx)
    mr.project(w["prec"], f, gx)
    V = mr.FunctionTree(mra)
    mr.apply(w["prec"], V, mr.PoissonOperator(mra, w["prec"]), f)
    d_dev = mr.dot(f, V)
    i_dev = V.integrate()
    root = int(f.export(False)[0][0, 0])
    d_ref = _np_dot(f, V, root)
    _say("dot", d_dev, d_ref, "integrate", i_dev)
    assert d_dev == pytest.approx(d_ref, rel=1e-13)
    idx, coefs, _ = V.export()
    roots = idx[:, 0] == root
    i_ref = float(np.sum(coefs[roots, 0])) * 2.0 ** (-1.5 * root)
    assert i_dev == pytest.approx(i_ref, rel=1e-14, abs=1e-14)
    assert mr.dot(V, V) == pytest.approx(V.getSquareNorm(), rel=1e-12)


def test_arithmetic_arguments_are_checked(mr):
    mra = mr.MultiResolutionAnalysis((-16, 16), 5)
    other = mr.MultiResolutionAnalysis((-32, 32), 5)
    f, g = mr.FunctionTree(mra), mr.FunctionTree(other)
    gx = mr.GaussExp.from_sources(W.single_gaussian(10.0))
    mr.build_grid(f, gx)
    mr.project(1e-4, f, gx)
    with pytest.raises(mr.MRCPPError):
        mr.add(1e-4, mr.FunctionTree(mra), 1.0, f, 1.0, g)  # incompatible MRA
    with pytest.raises(mr.MRCPPError):
        mr.add(1e-4, f, 1.0, f, 1.0, f)  # output among the inputs
    with pytest.raises(mr.MRCPPError):
        mr.add(1e-4, mr.FunctionTree(mra), [(1.0, f)] * 9)  # more than 8 terms
    with pytest.raises(mr.MRCPPError):
        mr.multiply(1e-4, mr.FunctionTree(mra), [])
    with pytest.raises(mr.MRCPPError):
        mr.multiply(1e-4, mr.FunctionTree(mra), [(1.0, f)] * 3, useMaxNorms=True)  # the max-norm refinement takes two trees
