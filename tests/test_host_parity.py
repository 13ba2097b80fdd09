"""The product's host layer (projection, kernels, operator trees, band widths)
against the CPU oracle: bit-identical on the same inputs.  CPU only."""
import numpy as np
import pytest

from mrcpp_amd import workloads as W

CASES = ["gauss1_k5", "gauss1_k9_helmholtz"]


def _inputs(mr, orc, name):
    w = W.WORKLOADS[name]
    src = w["sources"]()
    box = (-w["box"], w["box"])
    omra = orc.OracleMRA(box, w["order"])
    mra = mr.MultiResolutionAnalysis(box, w["order"])
    of = orc.OracleTree(omra)
    of.project_gaussians(w["prec"], src.coefs, src.exponents, src.centers)
    f = mr.FunctionTree(mra)
    g = mr.GaussExp.from_sources(src)
    mr.build_grid(f, g)
    mr.project(w["prec"], f, g)
    return w, omra, mra, of, f


@pytest.mark.parametrize("name", CASES)
def test_projection_bit_identical(mr, orc, name):
    w, omra, mra, of, f = _inputs(mr, orc, name)
    oi, oc, on = of.export()
    pi, pc, pn = f.export()
    assert np.array_equal(oi, pi)
    assert np.array_equal(oc, pc)
    assert np.array_equal(on, pn)
    assert of.square_norm() == f.getSquareNorm()
    assert f.integrate() == of.integrate()


def test_projection_sixteen_gaussians(mr, orc):
    w, omra, mra, of, f = _inputs(mr, orc, "gauss16_k7")
    oi, oc, _ = of.export()
    pi, pc, _ = f.export()
    assert np.array_equal(oi, pi) and np.array_equal(oc, pc)


@pytest.mark.parametrize("box,order,prec,mu", [(20, 5, 1e-5, None), (32, 7, 1e-6, 0.5), (20, 3, 1e-4, None)])
def test_operator_bit_identical(mr, orc, box, order, prec, mu):
    omra = orc.OracleMRA((-box, box), order)
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    OP = orc.OracleOperator(omra, prec, mu=mu)
    P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
    assert OP.size() == P.size()
    ka, kc = OP.kernel()
    pa, pc = P.kernel()
    assert np.array_equal(ka, pa) and np.array_equal(kc, pc)
    for ap in (prec, 10 * prec, 1e-3):
        assert np.array_equal(OP.bandwidths(ap), P.bandwidths(ap))
    compared = 0
    for i in range(0, P.size(), 3):
        for d in range(0, 14):
            for l in range(-4, 5):
                for c in range(4):
                    a, an = OP.node(i, d, l, c)
                    b, bn = P.node(i, d, l, c)
                    assert (a is None) == (b is None)
                    if a is None:
                        continue
                    assert np.array_equal(a, b) and an == bn
                    compared += 1
    assert compared > 1000


def test_batched_host_construction_matches_oracle(mr, orc):
    """build_operators on the host cores: each operator of a mixed batch equals
    the oracle's construction of the same operator bit for bit."""
    mra = mr.MultiResolutionAnalysis((-20, 20), 3)
    omra = orc.OracleMRA((-20, 20), 3)
    specs = [(None, 1e-4), (1.2, 1e-4), (None, 1e-3)]
    ops = mr.build_operators(mra, specs, "host")
    assert [type(op) for op in ops] == [mr.PoissonOperator, mr.HelmholtzOperator, mr.PoissonOperator]
    for (mu, prec), P in zip(specs, ops):
        assert not P.built_on_device()
        OP = orc.OracleOperator(omra, prec, mu=mu)
        assert OP.size() == P.size()
        assert np.array_equal(OP.bandwidths(prec), P.bandwidths(prec))
        compared = 0
        for i in range(P.size()):
            for d in range(0, 10):
                for l in range(-3, 4):
                    for c in range(4):
                        a, an = OP.node(i, d, l, c)
                        b, bn = P.node(i, d, l, c)
                        assert (a is None) == (b is None)
                        if a is not None:
                            assert np.array_equal(a, b) and an == bn
                            compared += 1
        assert compared > 200
    with pytest.raises(mr.MRCPPError):
        mr.build_operators(mra, [(None, 0.0)], "host")


def test_kernel_terms_match(mr, orc):
    for kind, mu in ((0, 0.0), (1, 0.7)):
        a1, c1 = orc.Oracle.kernel_terms(kind, mu, 1e-7, 1e-5, 55.4)
        a2, c2 = mr.kernel_terms(kind, mu, 1e-7, 1e-5, 55.4)
        assert np.array_equal(a1, a2) and np.array_equal(c1, c2)


def test_incompatible_or_invalid_arguments(mr):
    with pytest.raises(mr.MRCPPError):
        mr.MultiResolutionAnalysis((-20, 21), 5)  # lower bound must be 0 or -upper
    with pytest.raises(mr.MRCPPError):
        mr.MultiResolutionAnalysis((-20, 20), 50)  # beyond MaxOrder
