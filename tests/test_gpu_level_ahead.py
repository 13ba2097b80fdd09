"""Levels formed ahead on the GPU give the same trees as the host's split.

With the level-ahead pipeline (opt-in, for a single plain tree), right
after a step's band sums the device forms the step's split decisions, the
next level's nodes and that level's pair screen, all before the host splits;
the host takes them only where its own split agrees bit for bit.  The trees
must equal the host-only path's bit for bit, with the same applied tuples,
candidate pairs and generated band nodes, and the inner steps must have been
formed ahead.  Parity of the default path with the oracle is held by the
apply tests.
"""
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture
def ahead_on(mr):
    yield
    mr.set_level_ahead(False)
    mr.set_sweep_precision("fp64")


def _apply(mr, src, order, box, prec, mu, ahead, max_iter=-1, abs_prec=False, precision="fp64"):
    mr.set_level_ahead(ahead)
    mr.set_sweep_precision(precision)
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(src)
    mr.build_grid(f, gx)
    mr.project(prec, f, gx)
    P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
    g = mr.FunctionTree(mra)
    st = mr.apply(prec, g, P, f, maxIter=max_iter, absPrec=abs_prec)
    return g.export(), st


def _workload(key):
    w = W.WORKLOADS[key]
    return w["sources"](), w["order"], w["box"], w["prec"]


CASES = [
    # id, (sources, order, box, prec), mu, max_iter, abs_prec, sweep precision
    ("gauss16_k7_poisson", "gauss16_k7", None, -1, False, "fp64"),
    ("gauss16_k7_helmholtz", "gauss16_k7", 1.5, -1, False, "fp64"),
    ("gauss16_k7_fp32", "gauss16_k7", None, -1, False, "fp32"),
    ("gauss16_k7_maxiter3", "gauss16_k7", None, 3, False, "fp64"),
    ("gauss16_k7_absprec", "gauss16_k7", None, -1, True, "fp64"),
    ("gauss1_k5_poisson", "gauss1_k5", None, -1, False, "fp64"),
    ("gauss1_k9_poisson", "gauss1_k9_poisson", None, -1, False, "fp64"),
    ("order3_spread", None, None, -1, False, "fp64"),
]


@pytest.mark.parametrize("name,key,mu,max_iter,abs_prec,precision", CASES, ids=[c[0] for c in CASES])
def test_ahead_equals_host_split(mr, ahead_on, name, key, mu, max_iter, abs_prec, precision):
    if key is None: # generated band nodes, ragged levels
        src, order, box, prec = W.random_sources(99, 6, center_range=10.0, beta_range=(0.5, 200.0)), 3, 16, 1e-4
    else:
        src, order, box, prec = _workload(key)
    _say("case", name)
    (ai, ac, an), ast = _apply(mr, src, order, box, prec, mu, True, max_iter, abs_prec, precision)
    (hi, hc, hn), hst = _apply(mr, src, order, box, prec, mu, False, max_iter, abs_prec, precision)
    _say(name, "levels", ast["n_levels"], "formed ahead", ast["n_ahead"], "nodes", len(ai), "applied", ast["n_applied"],
         "pairs", ast["n_pairs"], hst["n_pairs"])
    assert hst["n_ahead"] == 0
    assert ast["n_ahead"] >= 1
    assert ast["n_levels"] == hst["n_levels"]
    assert np.array_equal(ai, hi)
    assert np.array_equal(ac, hc), float(np.abs(ac - hc).max())
    assert np.array_equal(an, hn)
    for k in ("n_applied", "n_pairs", "n_generated", "n_output_nodes"):
        assert ast[k] == hst[k], k


def test_ahead_switch(mr, ahead_on):
    assert not mr.level_ahead()
    mr.set_level_ahead(True)
    assert mr.level_ahead()


def test_ahead_over_a_sequence_of_trees(mr, ahead_on):
    """Uneven order-7 trees applied one after another with levels formed
    ahead (the pair arrays grow between the screen of a level formed ahead and
    its band sums), then again on the host's split: the same trees."""
    specs = W.batch_sources()[32:56]
    mra = mr.MultiResolutionAnalysis((-32, 32), 7)
    inputs = []
    for sp in specs:
        f = mr.FunctionTree(mra)
        gx = mr.GaussExp.from_sources(sp["sources"])
        mr.build_grid(f, gx)
        mr.project(sp["prec"], f, gx)
        P = mr.PoissonOperator(mra, sp["prec"]) if sp["mu"] is None else mr.HelmholtzOperator(mra, sp["mu"], sp["prec"])
        inputs.append((sp["prec"], P, f))
    out = {}
    for ahead in (True, False):
        mr.set_level_ahead(ahead)
        out[ahead] = []
        for prec, P, f in inputs:
            g = mr.FunctionTree(mra)
            st = mr.apply(prec, g, P, f)
            out[ahead].append((g.export(), st))
    n_ahead = 0
    for k, (((ai, ac, _), ast), ((hi, hc, _), hst)) in enumerate(zip(out[True], out[False])):
        assert np.array_equal(ai, hi), k
        assert np.array_equal(ac, hc), k
        assert ast["n_applied"] == hst["n_applied"], k
        n_ahead += ast["n_ahead"]
    _say("trees", len(inputs), "steps formed ahead", n_ahead)
    assert n_ahead > 0
