"""The oracle's apply with precision trees (src/treebuilders/apply.cpp:214-251)
pinned by what the definition fixes.

precFunc(idx) = 1 / max over the trees of sqrt(getMaxSquareNorm(idx)), where
getMaxSquareNorm is the largest 2^(3n) |c|^2 of the node and its real
descendants, and a node below a tree's leaves is generated from its leaf.

Pins:
  * every real node's factor against the subtree maxima recomputed from the
    exported tree (brute force);
  * generated nodes through Parseval: the 8 children (and 64 grandchildren)
    generated below a leaf carry the leaf's whole square norm;
  * degree -1 homogeneity: doubling a tree halves every factor, exactly;
  * the maximum over trees: [h, h/2] gives the factors and the apply of [h];
  * an empty tree list is the plain apply, bit for bit.
"""
import numpy as np
import pytest

from mrcpp_amd import workloads as W

BOX, ORDER, PREC = 16, 5, 1e-4


def _tree(orc, omra, src, scale=1.0):
    t = orc.OracleTree(omra)
    t.project_gaussians(PREC, np.asarray(src.coefs) * scale, src.exponents, src.centers)
    return t


@pytest.fixture(scope="module")
def setup(orc):
    omra = orc.OracleMRA((-BOX, BOX), ORDER)
    f_src = W.random_sources(5, 3, center_range=3.0, beta_range=(2.0, 80.0))
    h_src = W.random_sources(17, 2, center_range=5.0, beta_range=(0.5, 20.0))
    return omra, f_src, h_src


def _subtree_max(idx, norms):
    """Largest 2^(3n) |c|^2 over each node and its descendants, by (n, l)."""
    val = {}
    for row, nr in zip(idx, norms):
        n, l = int(row[0]), tuple(int(x) for x in row[1:4])
        val[(n, l)] = 2.0 ** (3 * n) * nr[8]
    best = dict(val)
    for (n, l) in sorted(val, key=lambda k: -k[0]):
        p = (n - 1, tuple(x >> 1 for x in l))
        if p in best:
            best[p] = max(best[p], best[(n, l)])
    return best


def test_real_node_factors_are_subtree_maxima(orc, setup):
    omra, _, h_src = setup
    h = _tree(orc, omra, h_src)
    idx, _, norms = h.export()
    best = _subtree_max(idx, norms)
    assert len(best) == len(idx) > 8
    for (n, l), m in best.items():
        assert orc.prec_factor([h], n, l) == pytest.approx(1.0 / np.sqrt(m), rel=1e-15)


def test_generated_node_factors_keep_the_leaf_norm(orc, setup):
    omra, _, h_src = setup
    h = _tree(orc, omra, h_src)
    idx, _, norms = h.export()
    leaves = [(int(r[0]), tuple(int(x) for x in r[1:4]), nr[8]) for r, nr in zip(idx, norms) if r[4]]
    rng = np.random.default_rng(3)
    for k in rng.choice(len(leaves), size=6, replace=False):
        n, l, sq = leaves[k]
        for depth in (1, 2):
            tot = 0.0
            for c in range(8 ** depth):
                cl = [(l[d] << depth) + ((c >> (depth * d)) & ((1 << depth) - 1)) for d in range(3)]
                fac = orc.prec_factor([h], n + depth, cl)
                tot += fac ** -2 / 2.0 ** (3 * (n + depth))
            assert tot == pytest.approx(sq, rel=1e-11), (n, l, depth)


def test_factors_are_homogeneous_of_degree_minus_one(orc, setup):
    omra, _, h_src = setup
    h = _tree(orc, omra, h_src)
    h2 = _tree(orc, omra, h_src, scale=2.0)
    i1, c1, _ = h.export()
    i2, c2, _ = h2.export()
    assert np.array_equal(i1, i2) and np.array_equal(c2, 2.0 * c1)  # premise: the projection is exactly linear
    for row in i1[::7]:
        n, l = int(row[0]), row[1:4].tolist()
        assert orc.prec_factor([h2], n, l) == 0.5 * orc.prec_factor([h], n, l)
        below = [2 * x + 1 for x in l]
        assert orc.prec_factor([h2], n + 1, below) == 0.5 * orc.prec_factor([h], n + 1, below)


def test_factor_is_the_maximum_over_trees(orc, setup):
    omra, f_src, h_src = setup
    h = _tree(orc, omra, h_src)
    hh = _tree(orc, omra, h_src, scale=0.5)
    f = _tree(orc, omra, f_src)
    idx, _, _ = f.export()
    for row in idx[::5]:
        n, l = int(row[0]), row[1:4].tolist()
        a, b = orc.prec_factor([h], n, l), orc.prec_factor([f], n, l)
        assert orc.prec_factor([h, hh], n, l) == a
        assert orc.prec_factor([h, f], n, l) == min(a, b)
    op = orc.OracleOperator(omra, PREC)
    g1, g2 = orc.OracleTree(omra), orc.OracleTree(omra)
    s1 = orc.apply(PREC, g1, op, f, prec_trees=[h])
    s2 = orc.apply(PREC, g2, op, f, prec_trees=[h, hh])
    (i1, c1, _), (i2, c2, _) = g1.export(), g2.export()
    assert np.array_equal(i1, i2) and np.array_equal(c1, c2) and s1["n_applied"] == s2["n_applied"]


def test_empty_tree_list_is_the_plain_apply(orc, setup):
    omra, f_src, _ = setup
    f = _tree(orc, omra, f_src)
    op = orc.OracleOperator(omra, PREC)
    g1, g2 = orc.OracleTree(omra), orc.OracleTree(omra)
    s1 = orc.apply(PREC, g1, op, f)
    s2 = orc.apply(PREC, g2, op, f, prec_trees=[])
    (i1, c1, _), (i2, c2, _) = g1.export(), g2.export()
    assert np.array_equal(i1, i2) and np.array_equal(c1, c2) and s1["n_applied"] == s2["n_applied"]


def test_local_precision_changes_the_grid(orc, setup):
    """A precision tree concentrated away from the source coarsens the output
    where it is small and refines it where it is large."""
    omra, f_src, h_src = setup
    f = _tree(orc, omra, f_src)
    h = _tree(orc, omra, h_src, scale=1e3)
    op = orc.OracleOperator(omra, PREC)
    g0, g1 = orc.OracleTree(omra), orc.OracleTree(omra)
    orc.apply(PREC, g0, op, f)
    orc.apply(PREC, g1, op, f, prec_trees=[h])
    i0, _, _ = g0.export()
    i1, _, _ = g1.export()
    assert not np.array_equal(i0, i1)
