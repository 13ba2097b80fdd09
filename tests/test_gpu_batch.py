"""apply_batch on the MI355X: many trees of uneven size, operator and
precision in one run of shared launches.

Bar: every tree of the batch has the node set and applied-tuple count of its
own apply() and of the CPU oracle, and coefficients equal to its own apply()
up to rounding (relative L2 <= 1e-13; unit boundaries may differ when a level
batch holds very many groups, which reorders the band sum).
"""
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def make_trees(mr, n, order=5, box=20, seed=31):
    rng = np.random.default_rng(seed)
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    out = []
    for i in range(n):
        src = W.random_sources(1000 + i, int(rng.integers(1, 7)), center_range=4.0, beta_range=(1.0, 300.0))
        prec = [1e-4, 1e-5][i % 2]
        f = mr.FunctionTree(mra)
        gx = mr.GaussExp.from_sources(src)
        mr.build_grid(f, gx)
        mr.project(prec, f, gx)
        mu = None if i % 3 == 0 else 0.5 + 0.5 * i
        P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
        out.append(dict(src=src, prec=prec, mu=mu, f=f, P=P))
    return mra, out


def test_batch_matches_single_applies(mr):
    mra, trees = make_trees(mr, 7)
    solo, solo_st = [], []
    for t in trees:
        g = mr.FunctionTree(mra)
        solo_st.append(mr.apply(t["prec"], g, t["P"], t["f"]))
        solo.append(g.export())
    outs = [mr.FunctionTree(mra) for _ in trees]
    st = mr.apply_batch([(t["prec"], o, t["P"], t["f"]) for t, o in zip(trees, outs)])
    for i, (t, o) in enumerate(zip(trees, outs)):
        bi, bc, bn = o.export()
        si, sc, sn = solo[i]
        _say("tree", i, "mu", t["mu"], "prec", t["prec"], "nodes", len(bi), "applied", st[i]["n_applied"])
        assert np.array_equal(bi, si)
        assert st[i]["n_applied"] == solo_st[i]["n_applied"]
        assert st[i]["n_output_nodes"] == solo_st[i]["n_output_nodes"]
        rel = np.linalg.norm(bc - sc) / np.linalg.norm(sc)
        assert rel <= 1e-13, rel


def test_batch_matches_oracle(mr, orc):
    mra, trees = make_trees(mr, 4, seed=7)
    outs = [mr.FunctionTree(mra) for _ in trees]
    st = mr.apply_batch([(t["prec"], o, t["P"], t["f"]) for t, o in zip(trees, outs)])
    for i, (t, o) in enumerate(zip(trees, outs)):
        omra = orc.OracleMRA((-20, 20), 5)
        of = orc.OracleTree(omra)
        of.project_gaussians(t["prec"], t["src"].coefs, t["src"].exponents, t["src"].centers)
        og = orc.OracleTree(omra)
        ost = orc.apply(t["prec"], og, orc.OracleOperator(omra, t["prec"], mu=t["mu"]), of)
        oi, oc, _ = og.export()
        pi, pc, _ = o.export()
        assert np.array_equal(oi, pi)
        assert st[i]["n_applied"] == ost["n_applied"]
        assert np.linalg.norm(pc - oc) / np.linalg.norm(oc) <= 1e-12


def test_batch_shared_input_and_errors(mr):
    mra, trees = make_trees(mr, 2, seed=3)
    t = trees[0]
    a, b = mr.FunctionTree(mra), mr.FunctionTree(mra)
    # the same input and operator in two tasks
    st = mr.apply_batch([(t["prec"], a, t["P"], t["f"]), (t["prec"], b, t["P"], t["f"])])
    ai, ac, _ = a.export()
    bi, bc, _ = b.export()
    assert np.array_equal(ai, bi) and np.array_equal(ac, bc)
    assert st[0]["n_applied"] == st[1]["n_applied"]
    # an output tree twice is rejected
    c = mr.FunctionTree(mra)
    with pytest.raises(mr.MRCPPError):
        mr.apply_batch([(t["prec"], c, t["P"], t["f"]), (t["prec"], c, t["P"], t["f"])])
    assert mr.apply_batch([]) == []
