"""apply() on the MI355X against the CPU oracle on identical inputs.

Bar: identical output node set, identical number of applied
(f, ft, gt, term) tuples, and coefficients within relative L2 10*prec of the
oracle (the FP64 path is held to 1e-12).  Size-independent properties
(determinism, closed-form energies) are checked on the GPU result itself.
"""
import math
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    # name, workload key or inline spec, mu
    ("gauss1_k5_poisson", "gauss1_k5", None),
    ("gauss1_k5_helmholtz", "gauss1_k5", 1.0),
    ("gauss16_k7_poisson", "gauss16_k7", None),
    ("gauss1_k9_poisson", "gauss1_k9_poisson", None),
    ("gauss1_k9_helmholtz", "gauss1_k9_helmholtz", 1.0),
]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def run_pair(mr, orc, src, order, box, prec, mu, max_iter=-1, abs_prec=False, grid_sources=None):
    b = (-box, box)
    omra = orc.OracleMRA(b, order)
    of = orc.OracleTree(omra)
    of.project_gaussians(prec, src.coefs, src.exponents, src.centers)
    op = orc.OracleOperator(omra, prec, mu=mu)
    og = orc.OracleTree(omra)
    mra = mr.MultiResolutionAnalysis(b, order)
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(src)
    mr.build_grid(f, gx)
    mr.project(prec, f, gx)
    P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
    g = mr.FunctionTree(mra)
    if grid_sources is not None:
        og.build_grid_gaussians(grid_sources.coefs, grid_sources.exponents, grid_sources.centers)
        mr.build_grid(g, mr.GaussExp.from_sources(grid_sources))
    ost = orc.apply(prec, og, op, of, max_iter=max_iter, abs_prec=abs_prec)
    st = mr.apply(prec, g, P, f, maxIter=max_iter, absPrec=abs_prec)
    return dict(omra=omra, of=of, og=og, ost=ost, mra=mra, f=f, g=g, P=P, st=st)


def assert_parity(r, prec):
    oi, oc, on = r["og"].export()
    pi, pc, pn = r["g"].export()
    assert np.array_equal(oi, pi), "output node sets differ"
    assert r["st"]["n_applied"] == r["ost"]["n_applied"]
    rel = np.linalg.norm(pc - oc) / np.linalg.norm(oc)
    assert rel <= 10 * prec
    assert rel <= 1e-12, rel
    assert np.allclose(pn, on, rtol=1e-10, atol=1e-13)
    assert r["g"].getSquareNorm() == pytest.approx(r["og"].square_norm(), rel=1e-12)
    return rel


@pytest.mark.parametrize("name,key,mu", CASES, ids=[c[0] for c in CASES])
def test_apply_matches_oracle(mr, orc, name, key, mu):
    w = W.WORKLOADS[key]
    _say("case", name)
    r = run_pair(mr, orc, w["sources"](), w["order"], w["box"], w["prec"], mu)
    rel = assert_parity(r, w["prec"])
    _say(name, "nodes", r["g"].getNNodes(), "applied", r["st"]["n_applied"], "rel L2", rel)


def test_apply_small_order3_many_sources(mr, orc):
    """Order 3, 6 sources with centres spread over the box (generated band nodes, ragged levels)."""
    src = W.random_sources(99, 6, center_range=10.0, beta_range=(0.5, 200.0))
    r = run_pair(mr, orc, src, 3, 16, 1e-4, None)
    assert_parity(r, 1e-4)
    assert r["st"]["n_generated"] > 0


def test_apply_on_existing_grid(mr, orc):
    """An output grid present before apply is computed in the first pass (all
    scales at once), as ConvolutionCalculator::getInitialWorkVector does."""
    src = W.single_gaussian(100.0, (0.3, -0.2, 0.1))
    grid = W.single_gaussian(30.0, (-1.0, 0.5, 0.0))
    r = run_pair(mr, orc, src, 5, 20, 1e-5, None, grid_sources=grid)
    assert_parity(r, 1e-5)


@pytest.mark.parametrize("max_iter,abs_prec", [(0, False), (2, False), (-1, True)])
def test_apply_refinement_controls(mr, orc, max_iter, abs_prec):
    r = run_pair(mr, orc, W.single_gaussian(100.0), 5, 20, 1e-5, None, max_iter=max_iter, abs_prec=abs_prec)
    assert_parity(r, 1e-5)
    if max_iter == 0:
        assert r["g"].getNNodes() == 8


def test_apply_deterministic_and_closed_form(mr):
    w = W.WORKLOADS["gauss1_k5"]
    mra = mr.MultiResolutionAnalysis((-20, 20), 5)
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(w["sources"]())
    mr.build_grid(f, gx)
    mr.project(w["prec"], f, gx)
    P = mr.PoissonOperator(mra, w["prec"])
    outs = []
    for _ in range(2):
        g = mr.FunctionTree(mra)
        mr.apply(w["prec"], g, P, f)
        outs.append(g.export())
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    # Coulomb self-energy of the unit Gaussian (beta = 100): sqrt(2 beta / pi)
    assert mr.dot(g, f) == pytest.approx(math.sqrt(200.0 / math.pi), rel=10 * w["prec"])
    # the potential itself, erf(sqrt(beta) r) / r, to the adaptive precision
    # relative to the output norm (the reference's own check of this case)
    from scipy.special import erf
    scale = 10 * w["prec"] * math.sqrt(g.getSquareNorm())
    for r in (0.02, 0.05, 0.1, 0.3, 1.0, 3.0, 7.0, 15.0):
        p = np.array([r, 0.013, -0.021])
        rr = float(np.linalg.norm(p))
        assert abs(g.evalf(p) - erf(10.0 * rr) / rr) < scale, r


def test_apply_rejects_incompatible_mra(mr):
    a = mr.MultiResolutionAnalysis((-20, 20), 5)
    b = mr.MultiResolutionAnalysis((-32, 32), 5)
    f = mr.FunctionTree(a)
    gx = mr.GaussExp([mr.GaussFunc(10.0, 1.0)])
    mr.build_grid(f, gx)
    mr.project(1e-3, f, gx)
    P = mr.PoissonOperator(a, 1e-3)
    with pytest.raises(mr.MRCPPError):
        mr.apply(1e-3, mr.FunctionTree(b), P, f)
