"""The oracle's screening pinned by what the reference's rules fix on their own.

1. First step (maxIter = 0): the applied-tuple count and the root
   coefficients of the oracle's apply equal the brute-force restatement in
   tests/step0_brute.py (band widths, the f-component skip, the
   per-component width test and direction-0 maxDeltaL, tensorApplyOperComp's
   index order), for Poisson and Helmholtz at orders 5 and 7.
2. The applied-tuple count grows as the precision tightens (a smaller
   threshold applies a superset of tuples on a refined grid).

Node-set identity with MRCPP itself beyond these stays unpinned: the
reference cannot be built here (Eigen and Catch2 are downloads).
"""
import numpy as np
import pytest

from mrcpp_amd import workloads as W
from tests.step0_brute import step0

CASES = [
    # name, sources, order, box, prec, mu
    ("gauss1_k5_poisson", W.single_gaussian(100.0), 5, 20, 1e-5, None),
    ("gauss1_k5_helmholtz", W.single_gaussian(100.0), 5, 20, 1e-5, 1.0),
    ("gauss4_k7_poisson", W.random_sources(11, 4, 4.0), 7, 32, 1e-6, None),
]


def roots_of(idx, coefs, norms, root_scale):
    sel = np.where(idx[:, 0] == root_scale)[0]
    order = sorted(sel, key=lambda r: (idx[r, 3], idx[r, 2], idx[r, 1]))
    return idx[order, 1:4], coefs[order], norms[order]


@pytest.mark.parametrize("name,src,order,box,prec,mu", CASES, ids=[c[0] for c in CASES])
def test_oracle_first_step_matches_brute_force(orc, name, src, order, box, prec, mu):
    omra = orc.OracleMRA((-box, box), order)
    of = orc.OracleTree(omra)
    of.project_gaussians(prec, src.coefs, src.exponents, src.centers)
    op = orc.OracleOperator(omra, prec, mu=mu)
    og = orc.OracleTree(omra)
    st = orc.apply(prec, og, op, of, max_iter=0)
    fi, fc, fn = of.export()
    root_scale = int(fi[:, 0].min())
    fl, fcoef, fnorm = roots_of(fi, fc, fn, root_scale)
    gi, gc, gn = og.export()
    assert (gi[:, 0] == root_scale).all() and len(gi) == len(fl), "maxIter = 0 keeps the roots only"
    gl, gcoef, _ = roots_of(gi, gc, gn, root_scale)
    assert np.array_equal(gl, fl)
    widths = op.bandwidths(prec)
    count, ref = step0(fl, fcoef, fnorm, order + 1, widths, lambda i, l, c: op.node(i, 0, l, c)[0])
    assert st["n_applied"] == count
    rel = np.linalg.norm(gcoef - ref) / np.linalg.norm(ref)
    assert rel <= 1e-13, rel


def test_oracle_applied_count_grows_with_precision(orc):
    src = W.single_gaussian(100.0)
    omra = orc.OracleMRA((-20, 20), 5)
    counts, nodes = [], []
    for prec in (1e-3, 1e-4, 1e-5):
        of = orc.OracleTree(omra)
        of.project_gaussians(prec, src.coefs, src.exponents, src.centers)
        og = orc.OracleTree(omra)
        st = orc.apply(prec, og, orc.OracleOperator(omra, prec), of)
        counts.append(st["n_applied"])
        nodes.append(st["n_g"])
    assert counts[0] < counts[1] < counts[2], counts
    assert nodes[0] <= nodes[1] <= nodes[2], nodes
