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


def test_apply_deep_tree_beyond_depth_18(mr, orc):
    """A Gaussian of exponent 1e12 (width 7e-7) in the box [-2, 2]^3: the input
    tree reaches 23 levels below the root and the output 20, past the 64-bit
    key range of depth 18; MRCPP allows 30 (api/constants.h:33).  Same bar as
    the other applies against the oracle, and the Coulomb self-energy
    sqrt(2 beta / pi) to 1e-3."""
    beta = 1.0e12
    src = W.GaussianSources([1.0], [beta], [(0.1, -0.2, 0.05)])
    r = run_pair(mr, orc, src, 5, 2, 1e-4, None)
    rel = assert_parity(r, 1e-4)
    fi, _, _ = r["f"].export(False)
    gi, _, _ = r["g"].export(False)
    fdepth = int(fi[:, 0].max() - fi[:, 0].min())
    gdepth = int(gi[:, 0].max() - gi[:, 0].min())
    e = mr.dot(r["g"], r["f"])
    e_ana = math.sqrt(2.0 * beta / math.pi)
    _say("deep tree: input depth", fdepth, "output depth", gdepth, "rel L2", rel, "energy rel", abs(e - e_ana) / e_ana)
    assert fdepth > 18 and gdepth > 18
    assert abs(e - e_ana) <= 1e-3 * e_ana


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


def test_reused_input_equals_fresh_input(mr):
    """Generated band nodes are removed from the input tree's device table
    after each apply (their slots emptied): applies of different operators
    and precisions on one input equal the same applies on a freshly
    projected copy, bit for bit."""
    src = W.WORKLOADS["gauss16_k7"]["sources"]()
    mra = mr.MultiResolutionAnalysis((-32, 32), 7)
    gx = mr.GaussExp.from_sources(src)

    def projected():
        f = mr.FunctionTree(mra)
        mr.build_grid(f, gx)
        mr.project(1e-6, f, gx)
        return f

    shared = projected()
    for prec, mu in ((1e-6, None), (1e-5, 2.0), (1e-7, None), (1e-6, None)):
        P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
        outs = []
        for f in (shared, projected()):
            g = mr.FunctionTree(mra)
            st = mr.apply(prec, g, P, f)
            outs.append((g.export(), st["n_applied"], st["n_generated"]))
        (a, na, ga), (b, nb, gb) = outs
        _say("prec", prec, "mu", mu, "applied", na, nb, "generated", ga, gb)
        assert na == nb and ga == gb
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


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


def test_concurrent_threads_serialised(mr):
    """Two host threads calling apply() at once (ctypes releases the GIL) get
    the trees the same calls give one after the other: the C ABI serialises
    every call that uses the device."""
    import threading
    mra = mr.MultiResolutionAnalysis((-20, 20), 5)
    jobs = []
    for seed, mu in ((5, None), (6, 1.5)):
        src = W.random_sources(seed, 4, center_range=3.0, beta_range=(2.0, 200.0))
        f = mr.FunctionTree(mra)
        gx = mr.GaussExp.from_sources(src)
        mr.build_grid(f, gx)
        mr.project(1e-5, f, gx)
        P = mr.PoissonOperator(mra, 1e-5) if mu is None else mr.HelmholtzOperator(mra, mu, 1e-5)
        jobs.append((f, P))
    seq = []
    for f, P in jobs:
        g = mr.FunctionTree(mra)
        mr.apply(1e-5, g, P, f)
        seq.append(g.export())
    results = [[None] * 3 for _ in jobs]
    errors = []

    def worker(j):
        try:
            f, P = jobs[j]
            for r in range(3):
                g = mr.FunctionTree(mra)
                mr.apply(1e-5, g, P, f)
                results[j][r] = g.export()
        except Exception as e:  # reported below
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(j,)) for j in range(len(jobs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(120)
    assert not errors, errors
    for j, (si, sc, _) in enumerate(seq):
        for r in range(3):
            ri, rc, _ = results[j][r]
            assert np.array_equal(ri, si) and np.array_equal(rc, sc), (j, r)


STEP0_CASES = [
    ("gauss1_k5_poisson", W.single_gaussian(100.0), 5, 20, 1e-5, None),
    ("gauss1_k5_helmholtz", W.single_gaussian(100.0), 5, 20, 1e-5, 1.0),
    ("gauss4_k7_poisson", W.random_sources(11, 4, 4.0), 7, 32, 1e-6, None),
    ("gauss16_k7_poisson", W.WORKLOADS["gauss16_k7"]["sources"](), 7, 32, 1e-6, None),
]


@pytest.mark.parametrize("name,src,order,box,prec,mu", STEP0_CASES, ids=[c[0] for c in STEP0_CASES])
def test_first_step_matches_brute_force(mr, name, src, order, box, prec, mu):
    """maxIter = 0 on the GPU against the brute-force restatement of the
    first breadth-first step (tests/step0_brute.py): the applied-tuple count
    exactly, the root coefficients to 1e-13."""
    from tests.step0_brute import step0
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(src)
    mr.build_grid(f, gx)
    mr.project(prec, f, gx)
    P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
    g = mr.FunctionTree(mra)
    st = mr.apply(prec, g, P, f, maxIter=0)
    fi, fc, fn = f.export()
    n0 = int(fi[:, 0].min())

    def roots(idx, c, n):
        sel = sorted(np.where(idx[:, 0] == n0)[0], key=lambda r: (idx[r, 3], idx[r, 2], idx[r, 1]))
        return idx[sel, 1:4], c[sel], n[sel]

    fl, fcoef, fnorm = roots(fi, fc, fn)
    gi, gc, gn = g.export()
    assert (gi[:, 0] == n0).all() and len(gi) == len(fl)
    gl, gcoef, _ = roots(gi, gc, gn)
    assert np.array_equal(gl, fl)
    count, ref = step0(fl, fcoef, fnorm, order + 1, P.bandwidths(prec), lambda i, l, c: P.node(i, 0, l, c)[0])
    rel = np.linalg.norm(gcoef - ref) / np.linalg.norm(ref)
    _say(name, "applied", st["n_applied"], "brute force", count, "rel L2", rel)
    assert st["n_applied"] == count
    assert rel <= 1e-13


def test_applied_count_grows_with_precision(mr):
    src = W.single_gaussian(100.0)
    mra = mr.MultiResolutionAnalysis((-20, 20), 5)
    gx = mr.GaussExp.from_sources(src)
    counts, nodes = [], []
    for prec in (1e-3, 1e-4, 1e-5, 1e-6):
        f = mr.FunctionTree(mra)
        mr.build_grid(f, gx)
        mr.project(prec, f, gx)
        g = mr.FunctionTree(mra)
        st = mr.apply(prec, g, mr.PoissonOperator(mra, prec), f)
        counts.append(st["n_applied"])
        nodes.append(st["n_output_nodes"])
    _say("applied tuples", counts, "output nodes", nodes)
    assert all(a < b for a, b in zip(counts, counts[1:])), counts
    assert all(a <= b for a, b in zip(nodes, nodes[1:])), nodes


def test_reference_poisson_energy_kat(mr):
    """MRCPP's own Poisson test (tests/operators/poisson_operator.cpp:129-154)
    at its parameters, in a Legendre basis of order 5: the world of
    tests/factory_functions.h (root scale 1, corner (-1, 0, 1), 1 x 2 x 3 root
    boxes), a normalised Gaussian of exponent 1e4 at (-0.2, 0.5, 1.0)
    projected at 1e-4, the Poisson operator built at 1e-4 and applied at
    1e-3; dot(P f, f) equals the analytic Coulomb self-energy
    sqrt(2 beta / pi) (GaussFunc::calcCoulombEnergy,
    src/functions/GaussFunc.cpp:210-237) to relative 1e-3."""
    beta = 1.0e4
    src = W.GaussianSources([1.0], [beta], [(-0.2, 0.5, 1.0)])
    mra = mr.MultiResolutionAnalysis((1, (-1, 0, 1), (1, 2, 3)), 5)
    f = mr.FunctionTree(mra)
    mr.project(1e-4, f, mr.GaussExp.from_sources(src))
    P = mr.PoissonOperator(mra, 1e-4)
    g = mr.FunctionTree(mra)
    st = mr.apply(1e-3, g, P, f)
    e_num = mr.dot(g, f)
    e_ana = math.sqrt(2.0 * beta / math.pi)
    _say("KAT: E_num", e_num, "E_ana", e_ana, "rel", abs(e_num - e_ana) / e_ana, "nodes", st["n_output_nodes"], "terms", P.size())
    assert abs(f.integrate() - 1.0) < 1e-3
    assert abs(e_num - e_ana) <= 1e-3 * abs(e_ana)
