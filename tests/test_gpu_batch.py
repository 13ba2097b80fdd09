"""apply_batch on the MI355X: many trees of uneven size, operator and
precision in one run of shared launches.

Bar: every tree of the batch has the node set and applied-tuple count of its
own apply() and of the CPU oracle, and coefficients equal to its own apply()
up to rounding (relative L2 <= 1e-13; unit boundaries may differ when a level
batch holds very many groups, which reorders the band sum), bit for bit for a
small batch.
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
        # a small batch uses the single applies' work units, and the band
        # nodes of all its trees generated together are the trees' own: the
        # same coefficients bit for bit
        assert np.array_equal(bc, sc), float(np.abs(bc - sc).max())
        assert st[i]["n_generated"] == solo_st[i]["n_generated"]


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
    # and both equal the input's single apply (planned in task order on one mirror)
    s = mr.FunctionTree(mra)
    mr.apply(t["prec"], s, t["P"], t["f"])
    si, sc, _ = s.export()
    assert np.array_equal(ai, si) and np.array_equal(ac, sc)
    # an output tree twice is rejected
    c = mr.FunctionTree(mra)
    with pytest.raises(mr.MRCPPError):
        mr.apply_batch([(t["prec"], c, t["P"], t["f"]), (t["prec"], c, t["P"], t["f"])])
    assert mr.apply_batch([]) == []


def _batch64(mr):
    """The 64 uneven order-7 trees of the batch workload: 1..32 Gaussians,
    centres U[-6,6]^3, precisions 1e-5..1e-7, even index Poisson, odd index
    Helmholtz mu ~ U[0.5,3]; one operator per distinct (mu, prec)."""
    specs = W.batch_sources()
    mra = mr.MultiResolutionAnalysis((-32, 32), 7)
    ops, trees = {}, []
    for sp in specs:
        f = mr.FunctionTree(mra)
        gx = mr.GaussExp.from_sources(sp["sources"])
        mr.build_grid(f, gx)
        mr.project(sp["prec"], f, gx)
        key = (sp["mu"], sp["prec"])
        if key not in ops:
            ops[key] = mr.PoissonOperator(mra, sp["prec"]) if sp["mu"] is None else mr.HelmholtzOperator(mra, sp["mu"], sp["prec"])
        trees.append(dict(src=sp["sources"], prec=sp["prec"], mu=sp["mu"], f=f, P=ops[key]))
    return mra, trees


def test_batch64_k7(mr, orc):
    """The full 64-tree order-7 batch through one apply_batch: every tree has
    the node set and applied-tuple count of its own FP64 apply(); four trees
    (Poisson and Helmholtz, all three precisions) equal the CPU oracle; the
    FP32-sweep batch keeps every node set within 1e-6."""
    mra, trees = _batch64(mr)
    mr.set_sweep_precision("fp64")
    solo, solo_st = [], []
    for t in trees:
        g = mr.FunctionTree(mra)
        solo_st.append(mr.apply(t["prec"], g, t["P"], t["f"]))
        solo.append(g.export())
    outs = [mr.FunctionTree(mra) for _ in trees]
    st = mr.apply_batch([(t["prec"], o, t["P"], t["f"]) for t, o in zip(trees, outs)])
    batched = [o.export() for o in outs]
    worst = 0.0
    for i, (t, (bi, bc, _), (si, sc, _)) in enumerate(zip(trees, batched, solo)):
        assert np.array_equal(bi, si), "tree %d: node set differs from its own apply" % i
        assert st[i]["n_applied"] == solo_st[i]["n_applied"], i
        rel = np.linalg.norm(bc - sc) / np.linalg.norm(sc)
        worst = max(worst, rel)
        assert rel <= 1e-13, (i, rel)
    _say("64 trees: node sets and applied counts equal their own applies; worst rel L2 %.2e" % worst,
         "nodes", sum(len(b[0]) for b in batched), "applied", sum(s["n_applied"] for s in st))
    # oracle: the smallest tree of each (operator, precision) class among four
    picks = []
    for kind, prec in ((None, 1e-5), ("H", 1e-6), (None, 1e-7), ("H", 1e-5)):
        cand = [i for i, t in enumerate(trees) if t["prec"] == prec and ((t["mu"] is None) == (kind is None))]
        picks.append(min(cand, key=lambda i: len(trees[i]["src"])))
    for i in picks:
        t = trees[i]
        omra = orc.OracleMRA((-32, 32), 7)
        of = orc.OracleTree(omra)
        of.project_gaussians(t["prec"], t["src"].coefs, t["src"].exponents, t["src"].centers)
        og = orc.OracleTree(omra)
        ost = orc.apply(t["prec"], og, orc.OracleOperator(omra, t["prec"], mu=t["mu"]), of)
        oi, oc, _ = og.export()
        bi, bc, _ = batched[i]
        rel = np.linalg.norm(bc - oc) / np.linalg.norm(oc)
        _say("tree", i, "mu", t["mu"], "prec", t["prec"], "gaussians", len(t["src"]), "nodes", len(oi), "applied", ost["n_applied"], "rel", rel)
        assert np.array_equal(oi, bi), "tree %d: node set differs from the oracle" % i
        assert st[i]["n_applied"] == ost["n_applied"], i
        assert rel <= 1e-12, (i, rel)
    # FP32 sweeps: same node sets, every tree within 1e-6 of FP64
    mr.set_sweep_precision("fp32")
    try:
        outs32 = [mr.FunctionTree(mra) for _ in trees]
        mr.apply_batch([(t["prec"], o, t["P"], t["f"]) for t, o in zip(trees, outs32)])
    finally:
        mr.set_sweep_precision("fp64")
    worst = 0.0
    for i, (o, (si, sc, _)) in enumerate(zip(outs32, solo)):
        ci, cc, _ = o.export()
        assert np.array_equal(ci, si), "tree %d: FP32 node set differs" % i
        worst = max(worst, np.linalg.norm(cc - sc) / np.linalg.norm(sc))
    _say("FP32 batch: worst rel L2 against FP64 %.2e" % worst)
    assert worst <= 1e-6
