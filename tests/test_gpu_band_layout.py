"""The two work distributions of the order-7 matrix-core band kernel give
bit-identical trees.

'rows' (each wave owns 16 rows of every sweep, one workgroup barrier per
group) and 'tiles' (sweep tiles shared out over the waves) run the same
products in the same order for every output element, so their trees must be
equal bit for bit, in FP64 and FP32 sweeps, for Poisson and Helmholtz, for a
single tree and for a batch.  Parity of the default layout against the
oracle is held by the apply, batch and precision tests.
"""
import sys

import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture
def layouts(mr):
    yield
    mr.set_band_layout("auto")
    mr.set_sweep_precision("fp64")


def test_layout_switch(mr, layouts):
    assert mr.band_layout() == "auto"
    mr.set_band_layout("tiles")
    assert mr.band_layout() == "tiles"
    with pytest.raises(ValueError):
        mr.set_band_layout("columns")


def _apply(mr, src, order, box, prec, mu, layout, precision):
    mr.set_band_layout(layout)
    mr.set_sweep_precision(precision)
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(src)
    mr.build_grid(f, gx)
    mr.project(prec, f, gx)
    P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
    g = mr.FunctionTree(mra)
    st = mr.apply(prec, g, P, f)
    return g.export(), st


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mu", [None, 1.5])
def test_rows_equal_tiles(mr, layouts, precision, mu):
    src = W.WORKLOADS["gauss16_k7"]["sources"]()
    _say("case gauss16_k7", precision, "mu", mu)
    (ti, tc, tn), tst = _apply(mr, src, 7, 32, 1e-6, mu, "tiles", precision)
    (ri, rc, rn), rst = _apply(mr, src, 7, 32, 1e-6, mu, "rows", precision)
    _say("nodes", len(ti), len(ri), "applied", tst["n_applied"], rst["n_applied"])
    assert np.array_equal(ti, ri)
    assert tst["n_applied"] == rst["n_applied"]
    assert np.array_equal(tc, rc), float(np.abs(tc - rc).max())
    assert np.array_equal(tn, rn)


def test_rows_equal_tiles_batch(mr, layouts):
    """A batch of uneven Poisson and Helmholtz trees of three precisions
    (one level of many segments per breadth-first step)."""
    specs = W.batch_sources()[:8]
    mra = mr.MultiResolutionAnalysis((-32, 32), 7)
    out = {}
    for layout in ("tiles", "rows"):
        mr.set_band_layout(layout)
        tasks, trees = [], []
        for b in specs:
            prec, mu = b["prec"], b["mu"]
            f = mr.FunctionTree(mra)
            gx = mr.GaussExp.from_sources(b["sources"])
            mr.build_grid(f, gx)
            mr.project(prec, f, gx)
            P = mr.PoissonOperator(mra, prec) if mu is None else mr.HelmholtzOperator(mra, mu, prec)
            g = mr.FunctionTree(mra)
            tasks.append((prec, g, P, f))
            trees.append(g)
        mr.apply_batch(tasks)
        out[layout] = [t.export() for t in trees]
    for k, ((ti, tc, tn), (ri, rc, rn)) in enumerate(zip(out["tiles"], out["rows"])):
        assert np.array_equal(ti, ri), k
        assert np.array_equal(tc, rc), (k, float(np.abs(tc - rc).max()))
        assert np.array_equal(tn, rn), k


def test_unit_groups_change_only_the_last_bits(mr, layouts):
    """The minimum groups per work unit reorders only the additions of a
    node's partial sums: same node set and applied count, coefficients equal
    to 1e-14."""
    src = W.WORKLOADS["gauss16_k7"]["sources"]()
    try:
        (ri, rc, _), rst = _apply(mr, src, 7, 32, 1e-6, None, "auto", "fp64")
        for ug, wg in ((8, 0), (64, 2), (1024, 0)):
            mr.set_band_launch(ug, wg)
            (ti, tc, _), tst = _apply(mr, src, 7, 32, 1e-6, None, "auto", "fp64")
            rel = np.linalg.norm(tc - rc) / np.linalg.norm(rc)
            _say("unit groups", ug, "wg/cu", wg, "rel", rel)
            assert np.array_equal(ti, ri) and tst["n_applied"] == rst["n_applied"]
            assert rel <= 1e-14
    finally:
        mr.set_band_launch(32, 0)


def test_autotune_band(mr, layouts):
    src = W.WORKLOADS["gauss16_k7"]["sources"]()
    mra = mr.MultiResolutionAnalysis((-32, 32), 7)
    f = mr.FunctionTree(mra)
    gx = mr.GaussExp.from_sources(src)
    mr.build_grid(f, gx)
    mr.project(1e-6, f, gx)
    P = mr.PoissonOperator(mra, 1e-6)
    calls = []

    def run():
        calls.append(1)
        mr.apply(1e-6, mr.FunctionTree(mra), P, f)

    try:
        res = mr.autotune_band(run, repeats=2)
        cands = res["candidates"]
        _say("autotune best", res["best"], "of", len(cands))
        assert len(cands) == 12 and len(calls) == 12 * 3
        assert res["best"]["seconds"] == min(c["seconds"] for c in cands)
        assert mr.band_layout() == res["best"]["layout"]
        assert mr.band_launch() == (res["best"]["unit_groups"], res["best"]["workgroups_per_cu"])
        # the tuned shape computes the same tree (to the last bits)
        g = mr.FunctionTree(mra)
        mr.set_sweep_precision("fp64")
        mr.apply(1e-6, g, P, f)
        mr.set_band_launch(32, 0)
        mr.set_band_layout("auto")
        h = mr.FunctionTree(mra)
        mr.apply(1e-6, h, P, f)
        (gi, gc, _), (hi, hc, _) = g.export(), h.export()
        assert np.array_equal(gi, hi)
        assert np.linalg.norm(gc - hc) / np.linalg.norm(hc) <= 1e-14
        # an exception in the workload propagates and the settings are restored
        mr.set_band_launch(48, 3)

        def bad():
            raise RuntimeError("workload failed")

        with pytest.raises(RuntimeError):
            mr.autotune_band(bad, repeats=1)
        assert mr.band_launch() == (48, 3) and mr.band_layout() == "auto"
    finally:
        mr.set_band_launch(32, 0)
