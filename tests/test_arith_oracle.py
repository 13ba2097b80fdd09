"""The oracle's add and multiply (src/treebuilders/add.cpp:43-80,
src/treebuilders/multiply.cpp:52-100) pinned against closed forms.  CPU only.

* add on a fixed grid (prec < 0, no refinement) is exactly linear: evalf of the
  result equals the combination of the inputs' evalf to rounding at any point,
  and the integral is the combination of the integrals;
* adaptive add of f and -f is zero;
* the product of two Gaussians is a Gaussian: its integral has a closed form,
  matched to the projection precision, and equals the inputs' dot product;
* multiply by the constant function (a tree projected from a wide flat
  Gaussian) leaves a function unchanged to that constant's accuracy.
"""
import math

import numpy as np
import pytest


def _gauss_tree(orc, m, beta, pos, weight=1.0, prec=1e-7):
    t = orc.OracleTree(m)
    t.project_gaussians(prec, [weight * (beta / math.pi) ** 1.5], [beta], [pos])
    return t


def _lib(orc):
    return orc.Oracle.lib()


def _on_grid(orc, m, betas, centers, weights):
    """A Gaussian sum projected onto the grid build_grid makes for all of
    `betas` (project with prec < 0 computes the grid's nodes and stops)."""
    t = orc.OracleTree(m)
    t.build_grid_gaussians([1.0] * len(betas), betas, centers)
    t.project_gaussians(-1.0, [w * (b / math.pi) ** 1.5 for w, b in zip(weights, betas)], betas, centers)
    return t


def test_add_on_a_shared_grid_is_linear(orc):
    """Inputs and output on one grid: every node of the sum is the combination
    of the inputs' nodes (to rounding), so are the values and the integral."""
    m = orc.OracleMRA((-16, 16), 5)
    betas, centers = [3.0, 20.0], [(0.3, -0.2, 0.1), (-0.4, 0.5, 0.0)]
    f = _on_grid(orc, m, betas, centers, [1.0, 0.0])
    g = _on_grid(orc, m, betas, centers, [0.0, 2.0])
    out = orc.OracleTree(m)
    out.build_grid_gaussians([1.0, 1.0], betas, centers)
    orc.add(-1.0, out, [(1.5, f), (-0.25, g)])
    fi, fc, _ = f.export()
    gi, gc, _ = g.export()
    oi, oc, _ = out.export()
    assert np.array_equal(fi, oi) and np.array_equal(gi, oi)
    want = 1.5 * fc - 0.25 * gc
    assert np.abs(oc - want).max() <= 1e-14 * np.abs(want).max()
    L = _lib(orc)
    rng = np.random.default_rng(3)
    for r in rng.uniform(-1.5, 1.5, size=(20, 3)):
        w = 1.5 * L.orc_tree_evalf(f.h, *r) - 0.25 * L.orc_tree_evalf(g.h, *r)
        assert abs(L.orc_tree_evalf(out.h, *r) - w) <= 1e-13 * max(1.0, abs(w))
    assert L.orc_tree_integrate(out.h) == pytest.approx(1.5 * L.orc_tree_integrate(f.h) - 0.25 * L.orc_tree_integrate(g.h), abs=1e-13)


def test_add_without_refinement_keeps_the_grid(orc):
    """prec < 0: the output keeps its grid; scaling coefficients, hence the
    integral, still combine exactly."""
    m = orc.OracleMRA((-16, 16), 5)
    f = _gauss_tree(orc, m, 3.0, (0.3, -0.2, 0.1), prec=1e-5)
    g = _gauss_tree(orc, m, 20.0, (-0.4, 0.5, 0.0), 2.0, prec=1e-5)
    out = orc.OracleTree(m)
    out.build_grid_gaussians([1.0, 1.0], [3.0, 20.0], [(0.3, -0.2, 0.1), (-0.4, 0.5, 0.0)])
    n0 = out.num_nodes()
    orc.add(-1.0, out, [(1.5, f), (-0.25, g)])
    assert out.num_nodes() == n0
    L = _lib(orc)
    assert L.orc_tree_integrate(out.h) == pytest.approx(1.5 * L.orc_tree_integrate(f.h) - 0.25 * L.orc_tree_integrate(g.h), abs=1e-12)


def test_add_of_a_function_and_its_negative_is_zero(orc):
    m = orc.OracleMRA((-16, 16), 5)
    f = _gauss_tree(orc, m, 10.0, (0.1, 0.2, -0.3), prec=1e-5)
    out = orc.OracleTree(m)
    orc.add(1e-5, out, [(1.0, f), (-1.0, f)])
    assert out.square_norm() == 0.0
    assert out.num_nodes() == 8  # nothing to refine


def test_product_of_gaussians_closed_form(orc):
    m = orc.OracleMRA((-16, 16), 7)
    a, b = 3.0, 5.0
    A, B = np.array([0.3, -0.2, 0.1]), np.array([-0.4, 0.5, 0.0])
    f = _gauss_tree(orc, m, a, A)
    g = _gauss_tree(orc, m, b, B, 2.0)
    out = orc.OracleTree(m)
    orc.multiply(1e-7, out, [(1.0, f), (1.0, g)])
    L = _lib(orc)
    exact = 2.0 * (a * b / math.pi ** 2) ** 1.5 * (math.pi / (a + b)) ** 1.5 * math.exp(-a * b / (a + b) * np.sum((A - B) ** 2))
    assert L.orc_tree_integrate(out.h) == pytest.approx(exact, rel=1e-9)
    assert L.orc_tree_dot(f.h, g.h) == pytest.approx(exact, rel=1e-9)
    # pointwise: the product's value at a few points
    for r in ((0.0, 0.1, 0.0), (-0.2, 0.3, 0.05), (0.5, -0.5, 0.2)):
        r = np.array(r)
        want = 2.0 * (a * b / math.pi ** 2) ** 1.5 * math.exp(-a * np.sum((r - A) ** 2) - b * np.sum((r - B) ** 2))
        assert L.orc_tree_evalf(out.h, *r) == pytest.approx(want, rel=1e-5, abs=1e-7)


def test_multiply_by_a_constant(orc):
    m = orc.OracleMRA((-16, 16), 5)
    f = _gauss_tree(orc, m, 10.0, (0.1, 0.2, -0.3), prec=1e-6)
    one = orc.OracleTree(m)
    # exp(-1e-6 r^2) is 1 to 2.6e-4 relative over the box; on |r| < 1, to 1e-6
    one.project_gaussians(1e-6, [1.0], [1e-6], [(0.0, 0.0, 0.0)])
    out = orc.OracleTree(m)
    orc.multiply(1e-6, out, [(3.0, f), (1.0, one)])
    L = _lib(orc)
    assert L.orc_tree_integrate(out.h) == pytest.approx(3.0 * L.orc_tree_integrate(f.h), rel=1e-5)
    for r in ((0.1, 0.2, -0.3), (0.2, 0.0, -0.2)):
        assert L.orc_tree_evalf(out.h, *r) == pytest.approx(3.0 * L.orc_tree_evalf(f.h, *r), rel=1e-5)


def _scaled_maxima(idx, norms):
    """Per exported node (scale, lx, ly, lz): the largest 2^(3n) |c|^2 and
    2^(3n) |w|^2 over the node and its descendants, and whether it is a leaf."""
    key = {tuple(int(x) for x in r[:4]): i for i, r in enumerate(idx)}
    sq = norms[:, 8] * 2.0 ** (3 * idx[:, 0])
    wsq = np.sum(norms[:, 1:8] ** 2, axis=1) * 2.0 ** (3 * idx[:, 0])
    ms, mw, leaf = sq.copy(), wsq.copy(), np.ones(len(idx), bool)
    for i in np.argsort(-idx[:, 0]):  # deepest first: children before parents
        n, l = int(idx[i, 0]), idx[i, 1:4]
        if n > idx[:, 0].min():
            p = key[(n - 1, *(int(x) >> 1 for x in l))]
            ms[p] = max(ms[p], ms[i])
            mw[p] = max(mw[p], mw[i])
            leaf[p] = False
    return key, ms, mw, leaf


def test_multiply_with_max_norms(orc):
    """multiply(..., useMaxNorms=true) (MultiplicationAdaptor,
    src/treebuilders/MultiplicationAdaptor.h:46-66): the product of two
    Gaussians integrates to the closed-form overlap and matches the Gaussian
    product pointwise; the output never goes deeper than the deeper input; a
    node both inputs hold as real nodes is split exactly when
    maxW0 maxS1 + maxW1 maxS0 + maxW0 maxW1 > prec (recomputed here from the
    exported inputs) and neither is a leaf of both; a coarse prec refines
    nothing."""
    m = orc.OracleMRA((-16, 16), 5)
    a, b = 3.0, 40.0
    A, B = np.array([0.3, -0.2, 0.1]), np.array([-0.4, 0.5, 0.0])
    f = _gauss_tree(orc, m, a, A, prec=1e-6)
    g = _gauss_tree(orc, m, b, B, 2.0, prec=1e-6)
    prec = 1e-6
    out = orc.OracleTree(m)
    orc.multiply(prec, out, [(1.0, f), (1.0, g)], use_max_norms=True)
    L = _lib(orc)
    exact = 2.0 * (a * b / math.pi ** 2) ** 1.5 * (math.pi / (a + b)) ** 1.5 * math.exp(-a * b / (a + b) * np.sum((A - B) ** 2))
    assert L.orc_tree_integrate(out.h) == pytest.approx(exact, rel=1e-5)
    for r in ((0.0, 0.1, 0.0), (-0.2, 0.3, 0.05), (-0.4, 0.45, 0.02)):
        r = np.array(r)
        want = 2.0 * (a * b / math.pi ** 2) ** 1.5 * math.exp(-a * np.sum((r - A) ** 2) - b * np.sum((r - B) ** 2))
        assert L.orc_tree_evalf(out.h, *r) == pytest.approx(want, rel=1e-3, abs=1e-5)
    fi, _, fn = f.export()
    gi, _, gn = g.export()
    oi, _, _ = out.export()
    assert oi[:, 0].max() <= max(fi[:, 0].max(), gi[:, 0].max())
    fk, fms, fmw, fleaf = _scaled_maxima(fi, fn)
    gk, gms, gmw, gleaf = _scaled_maxima(gi, gn)
    ok, _, _, oleaf = _scaled_maxima(oi, np.zeros((len(oi), 9)))
    checked = 0
    for key, i in ok.items():
        if key not in fk or key not in gk:
            continue
        p, q = fk[key], gk[key]
        w0, w1, s0, s1 = math.sqrt(fmw[p]), math.sqrt(gmw[q]), math.sqrt(fms[p]), math.sqrt(gms[q])
        split = w0 * s1 + w1 * s0 + w0 * w1 > prec and not (fleaf[p] and gleaf[q])
        if key[0] + 2 > m.max_scale:
            continue
        assert split == (not oleaf[i]), key
        checked += 1
    assert checked > 20
    coarse = orc.OracleTree(m)
    orc.multiply(1e3, coarse, [(1.0, f), (1.0, g)], use_max_norms=True)
    assert len(coarse.export()[0]) == 8  # the 2 x 2 x 2 roots only
