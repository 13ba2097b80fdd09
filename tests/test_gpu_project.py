"""project() on the MI355X (ProjectionCalculator per level on the device,
src/treebuilders/project.cpp:85-102) against the CPU oracle's projection.

Bar: identical node set, coefficients within relative L2 1e-13 (the device
forms each Gaussian as a product of three 1D exponentials and sums with fused
multiply-adds, so values differ from the oracle's in the last bits), norms
within 1e-12; the result does not depend on where it was computed beyond that.
"""
import math
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    # name, sources, order, box, prec
    ("gauss1_k5", lambda: W.single_gaussian(100.0), 5, 20, 1e-5),
    ("gauss16_k7", lambda: W.random_sources(20240607, 16, 4.0), 7, 32, 1e-6),
    ("gauss1_k9", lambda: W.single_gaussian(100.0), 9, 20, 1e-7),
    ("mixed40_k3", lambda: W.random_sources(41, 40, center_range=8.0, beta_range=(0.3, 3000.0)), 3, 16, 1e-4),
]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _pair(mr, orc, src, order, box, prec, max_iter=-1, abs_prec=False, omra=None):
    omra = omra or orc.OracleMRA((-box, box), order)
    of = orc.OracleTree(omra)
    of.project_gaussians(prec, src.coefs, src.exponents, src.centers)
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(src)
    mr.build_grid(f, gx)
    mr.project(prec, f, gx, max_iter, abs_prec)
    return of, f


@pytest.mark.parametrize("name,srcf,order,box,prec", CASES, ids=[c[0] for c in CASES])
def test_projection_matches_oracle(mr, orc, name, srcf, order, box, prec):
    mr.set_projection_target("auto")
    of, f = _pair(mr, orc, srcf(), order, box, prec)
    assert mr.last_projection_on_device()
    oi, oc, on = of.export()
    pi, pc, pn = f.export()
    _say(name, "nodes", len(pi), "oracle", len(oi))
    assert np.array_equal(oi, pi), "projected node sets differ"
    rel = np.linalg.norm(pc - oc) / np.linalg.norm(oc)
    _say(name, "rel L2", rel)
    assert rel <= 1e-13, rel
    # component norms to 1e-12 of the node's norm (small wavelet norms are
    # differences of nearly equal values and carry the coefficients' absolute error)
    assert np.all(np.abs(pn[:, :8] - on[:, :8]) <= 1e-12 * np.sqrt(on[:, 8:9]) + 1e-15)
    assert np.allclose(pn[:, 8], on[:, 8], rtol=1e-12, atol=1e-15)
    assert f.getSquareNorm() == pytest.approx(of.square_norm(), rel=1e-12)


def test_projection_device_equals_host(mr):
    src = W.random_sources(5, 6, center_range=3.0, beta_range=(1.0, 300.0))
    mra = mr.MultiResolutionAnalysis((-16, 16), 5)
    gx = mr.GaussExp.from_sources(src)
    out = {}
    for target in ("host", "device"):
        mr.set_projection_target(target)
        f = mr.FunctionTree(mra)
        mr.build_grid(f, gx)
        mr.project(1e-5, f, gx)
        assert mr.last_projection_on_device() == (target == "device")
        out[target] = f.export()
    mr.set_projection_target("auto")
    (hi, hc, _), (di, dc, _) = out["host"], out["device"]
    assert np.array_equal(hi, di)
    assert np.linalg.norm(dc - hc) / np.linalg.norm(hc) <= 1e-13


def test_projection_charge_and_refinement_limits(mr):
    """Total charge of normalised Gaussians (integrate, FunctionTree.cpp:438-454),
    maxIter = 0 (no refinement) and absPrec."""
    src = W.random_sources(9, 5, center_range=3.0, beta_range=(2.0, 200.0))
    mra = mr.MultiResolutionAnalysis((-16, 16), 7)
    gx = mr.GaussExp.from_sources(src)
    f = mr.FunctionTree(mra)
    mr.build_grid(f, gx)
    mr.project(1e-7, f, gx)
    assert mr.last_projection_on_device()
    assert f.integrate() == pytest.approx(sum(src.weights), rel=1e-6, abs=1e-7)
    g = mr.FunctionTree(mra)
    mr.build_grid(g, gx)
    n0 = g.getNNodes()
    mr.project(1e-7, g, gx, 0)
    assert g.getNNodes() == n0
    h = mr.FunctionTree(mra)
    mr.build_grid(h, gx)
    mr.project(1e-7, h, gx, -1, True)
    assert h.getNNodes() >= n0


def test_apply_on_device_projected_input(mr, orc):
    """The apply path on a device-projected input against the oracle end to end."""
    w = W.WORKLOADS["gauss1_k5"]
    src, prec = w["sources"](), w["prec"]
    omra = orc.OracleMRA((-w["box"], w["box"]), w["order"])
    of, f = _pair(mr, orc, src, w["order"], w["box"], prec, omra=omra)
    og = orc.OracleTree(omra)
    orc.apply(prec, og, orc.OracleOperator(omra, prec), of)
    mra = mr.MultiResolutionAnalysis((-w["box"], w["box"]), w["order"])
    g = mr.FunctionTree(mra)
    mr.apply(prec, g, mr.PoissonOperator(mra, prec), f)
    oi, oc, _ = og.export()
    pi, pc, _ = g.export()
    assert np.array_equal(oi, pi)
    assert np.linalg.norm(pc - oc) / np.linalg.norm(oc) <= 1e-12
    # Coulomb self-energy of the normalised Gaussian: sqrt(2 beta / pi)
    beta = src.exponents[0]
    assert mr.dot(g, f) == pytest.approx(math.sqrt(2 * beta / math.pi), rel=10 * prec)
