"""Tree files of an apply result whose coefficients live in HBM
(FunctionTree::saveTree / saveTreeTXT, src/trees/FunctionTree.cpp:310-435):
the writers bring the coefficients to the host first, and the files read
back to the same tree (binary: bit for bit; text: 14 significant digits)."""
import numpy as np
import pytest

from mrcpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_files_of_a_device_resident_apply_result(mr, tmp_path):
    # the text format needs a world cube [-2^m, 2^m] with unit scaling factors
    w = dict(W.WORKLOADS["gauss1_k5"], box=16)
    mra = mr.MultiResolutionAnalysis((-w["box"], w["box"]), w["order"])
    gx = mr.GaussExp.from_sources(w["sources"]())
    f = mr.FunctionTree(mra)
    mr.build_grid(f, gx)
    mr.project(w["prec"], f, gx)
    g = mr.FunctionTree(mra)
    mr.apply(w["prec"], g, mr.PoissonOperator(mra, w["prec"]), f)
    base, txt = str(tmp_path / "g"), str(tmp_path / "g.txt")
    g.saveTree(base)
    g.saveTreeTXT(txt)
    ref = g.export()
    b = mr.FunctionTree(mra)
    b.loadTree(base)
    got = b.export()
    for name, x, y in zip(("idx", "coefs", "norms"), ref, got):
        same = np.array_equal(x, y, equal_nan=True)
        if not same:
            print(name, x.shape, y.shape, "nan", np.isnan(x).sum() if x.dtype.kind == "f" else 0,
                  "max diff", np.nanmax(np.abs(x.astype(float) - y.astype(float))) if x.shape == y.shape else None, flush=True)
        assert same, name
    t = mr.FunctionTree(mra)
    t.loadTreeTXT(txt)
    ti, tc, _ = t.export()
    assert np.array_equal(ti, ref[0])
    assert np.linalg.norm(tc - ref[1]) / np.linalg.norm(ref[1]) <= 1e-11
    # a world of another size is rejected, as by the reference reader
    with pytest.raises(mr.MRCPPError):
        mr.FunctionTree(mr.MultiResolutionAnalysis((-32, 32), w["order"])).loadTreeTXT(txt)
    # a loaded tree is a valid apply input
    h = mr.FunctionTree(mra)
    mr.apply(w["prec"], h, mr.PoissonOperator(mra, w["prec"]), b)
    assert h.getNNodes() > 0


def _text_tree(path):
    """Header lines, box lines (scale and translation) and value rows of a
    MADNESS-convention text tree: after the six header lines, one box line
    and one line of values per end node."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    head, body = lines[:6], lines[6:]
    boxes = [ln.split() for ln in body[0::2]]
    vals = np.array([np.array(ln.split(), dtype=float) for ln in body[1::2]])
    return head, boxes, vals


def test_text_file_of_apply_matches_the_oracle(mr, orc, tmp_path):
    """saveTreeTXT of the GPU apply result (coefficients in HBM) against the
    oracle writer's file of the oracle's apply of the same input: the same
    header, boxes and order (every integer token), values equal to 1e-12 of
    the largest (the two results agree to about 1e-15)."""
    w = dict(W.WORKLOADS["gauss1_k5"], box=16)
    src = w["sources"]()
    b = (-w["box"], w["box"])
    omra = orc.OracleMRA(b, w["order"])
    of = orc.OracleTree(omra)
    of.project_gaussians(w["prec"], src.coefs, src.exponents, src.centers)
    og = orc.OracleTree(omra)
    orc.apply(w["prec"], og, orc.OracleOperator(omra, w["prec"]), of)
    mra = mr.MultiResolutionAnalysis(b, w["order"])
    gx = mr.GaussExp.from_sources(src)
    f = mr.FunctionTree(mra)
    mr.build_grid(f, gx)
    mr.project(w["prec"], f, gx)
    g = mr.FunctionTree(mra)
    mr.apply(w["prec"], g, mr.PoissonOperator(mra, w["prec"]), f)
    p, o = str(tmp_path / "gpu.txt"), str(tmp_path / "oracle.txt")
    g.saveTreeTXT(p)
    og.save_txt(o)
    ph, pb, pv = _text_tree(p)
    oh, ob, ov = _text_tree(o)
    assert ph == oh
    assert pb == ob
    assert pv.shape == ov.shape
    assert np.abs(pv - ov).max() <= 1e-12 * np.abs(ov).max()
