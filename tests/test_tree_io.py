"""Tree files (FunctionTree::saveTree / loadTree, saveTreeTXT / loadTreeTXT,
src/trees/FunctionTree.cpp:127-435).  CPU only: the files are written and read
on the host.

The text format is pinned three ways: every value in a file written from a
projected tree is the projected function at the Gauss-Legendre point that the
record's MADNESS box and index name (closed form: a projection samples the
function exactly there); the product's file equals, byte for byte, the one the
oracle's restatement of the reference writer makes from the same tree; and
product and oracle read the same files, odd ones included, to identical trees.
"""
import math
import os

import numpy as np
import pytest

from mrcpp_amd import workloads as W


def _projected(mr, orc, order=5, box=16, prec=1e-5, src=None, oracle_tree=False):
    src = src or W.random_sources(9, 5, center_range=3.0, beta_range=(2.0, 200.0))
    mra = mr.MultiResolutionAnalysis((-box, box), order)
    gx = mr.GaussExp.from_sources(src)
    f = mr.FunctionTree(mra)
    mr.build_grid(f, gx)
    mr.project(prec, f, gx)
    omra = orc.OracleMRA((-box, box), order)
    of = None
    if oracle_tree:
        of = orc.OracleTree(omra)
        of.project_gaussians(prec, src.coefs, src.exponents, src.centers)
    return src, mra, f, omra, of


def _records(path):
    """(header lines, [(n_in, l_in[3], values[ncoefs])])"""
    with open(path) as fh:
        lines = fh.read().split("\n")
    head = lines[:6]
    count = int(head[5])
    recs = []
    for i in range(count):
        idx = [int(x) for x in lines[6 + 2 * i].split()]
        vals = np.array([float(x) for x in lines[7 + 2 * i].split()])
        recs.append((idx[0], idx[1:4], vals))
    return head, recs


def _gauss_sum(src, pts):
    out = np.zeros(len(pts))
    for c, b, r0 in zip(src.coefs, src.exponents, src.centers):
        out += c * np.exp(-b * np.sum((pts - np.asarray(r0)) ** 2, axis=1))
    return out


def test_text_values_are_samples_of_the_function(mr, orc, tmp_path):
    src, mra, f, _, _ = _projected(mr, orc)
    path = str(tmp_path / "f.txt")
    f.saveTreeTXT(path)
    head, recs = _records(path)
    assert head[0] == "3" and head[1:4] == ["-16 16"] * 3 and head[4] == "6"
    assert len(recs) == 8 * f.export(False)[0][:, 4].sum()
    kp1, L, rscale = 6, 16, -4
    x01 = (np.polynomial.legendre.leggauss(kp1)[0] + 1.0) / 2.0
    # file entry i: x index k - i // kp1^2, y index k - (i // kp1) % kp1, z index k - i % kp1
    i = np.arange(kp1 ** 3)
    ix, iy, iz = kp1 - 1 - i // kp1 ** 2, kp1 - 1 - (i // kp1) % kp1, kp1 - 1 - i % kp1
    fmax = 0.0
    worst = 0.0
    for n_in, l_in, vals in recs:
        n = n_in + rscale - 1  # MRCPP scale of the record's box
        h = 2.0 ** -n
        lo = np.array(l_in) * h - L  # MADNESS translations count from the lower corner -L
        pts = np.stack([lo[0] + h * x01[ix], lo[1] + h * x01[iy], lo[2] + h * x01[iz]], axis=1)
        ref = _gauss_sum(src, pts)
        fmax = max(fmax, np.abs(ref).max())
        worst = max(worst, np.abs(vals - ref).max())
    assert worst <= 1e-11 * fmax, (worst, fmax)


def test_text_file_matches_the_oracle_writer(mr, orc, tmp_path):
    _, _, f, _, of = _projected(mr, orc, oracle_tree=True)
    p, o = str(tmp_path / "p.txt"), str(tmp_path / "o.txt")
    f.saveTreeTXT(p)
    of.save_txt(o)
    with open(p, "rb") as a, open(o, "rb") as b:
        assert a.read() == b.read()


def test_text_round_trip_and_oracle_reader(mr, orc, tmp_path):
    _, mra, f, omra, _ = _projected(mr, orc, order=7, box=32, prec=1e-6,
                                    src=W.random_sources(20240607, 16, 4.0))
    path = str(tmp_path / "f.txt")
    f.saveTreeTXT(path)
    g = mr.FunctionTree(mra)
    g.loadTreeTXT(path)
    fi, fc, _ = f.export()
    gi, gc, gn = g.export()
    assert np.array_equal(fi, gi)
    # 14 significant digits per value, through the value -> coefficient maps
    assert np.linalg.norm(gc - fc) / np.linalg.norm(fc) <= 1e-11
    assert g.getSquareNorm() == pytest.approx(f.getSquareNorm(), rel=1e-11)
    og = orc.OracleTree(omra)
    og.load_txt(path)
    oi, oc, on = og.export()
    assert np.array_equal(oi, gi) and np.array_equal(oc, gc) and np.array_equal(on, gn)
    assert og.square_norm() == g.getSquareNorm()


def test_text_reader_on_ragged_files(mr, orc, tmp_path):
    """Files the writer does not produce: a box holding values and values below
    it (left incomplete, as the reference leaves it), and no records at all."""
    _, mra, f, omra, _ = _projected(mr, orc, order=3, box=16, prec=1e-3)
    path = str(tmp_path / "f.txt")
    f.saveTreeTXT(path)
    with open(path) as fh:
        lines = fh.read().split("\n")
    # values for the 8 children of the first record's box
    n_in, *l_in = [int(x) for x in lines[6].split()]
    vals = lines[7]
    extra = []
    for c in range(8):
        extra.append("%d %d %d %d " % (n_in + 1, 2 * l_in[0] + (c & 1), 2 * l_in[1] + ((c >> 1) & 1), 2 * l_in[2] + (c >> 2)))
        extra.append(vals)
    lines[5] = str(int(lines[5]) + 8)
    ragged = str(tmp_path / "ragged.txt")
    with open(ragged, "w") as fh:
        fh.write("\n".join(lines[:-1] + extra + [""]))
    empty = str(tmp_path / "empty.txt")
    with open(empty, "w") as fh:
        fh.write("\n".join(lines[:5] + ["0", ""]))
    for fname in (ragged, empty):
        g = mr.FunctionTree(mra)
        g.loadTreeTXT(fname)
        og = orc.OracleTree(omra)
        og.load_txt(fname)
        gi, gc, gn = g.export()
        oi, oc, on = og.export()
        assert np.array_equal(oi, gi) and np.array_equal(oc, gc)
        assert np.array_equal(np.nan_to_num(on), np.nan_to_num(gn))
    assert g.getSquareNorm() == 0.0 and g.getNNodes() == 8


def test_text_reader_rejects_other_worlds(mr, orc, tmp_path):
    _, _, f, _, _ = _projected(mr, orc)
    path = str(tmp_path / "f.txt")
    f.saveTreeTXT(path)
    # another order; a world cube [-32, 32] (the reader compares the cube with
    # 2^-rootScale only, FunctionTree.cpp:141-153)
    for box, order in ((16, 7), (32, 5)):
        g = mr.FunctionTree(mr.MultiResolutionAnalysis((-box, box), order))
        with pytest.raises(mr.MRCPPError):
            g.loadTreeTXT(path)
    with pytest.raises(mr.MRCPPError):
        f.loadTreeTXT(str(tmp_path / "missing.txt"))


def test_binary_round_trip(mr, orc, tmp_path):
    _, mra, f, _, _ = _projected(mr, orc, order=7, box=32, prec=1e-6, src=W.random_sources(20240607, 16, 4.0))
    base = str(tmp_path / "f")
    f.saveTree(base)
    assert os.path.exists(base + ".tree")
    g = mr.FunctionTree(mra)
    g.loadTree(base)
    a, b = f.export(), g.export()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert g.getSquareNorm() == f.getSquareNorm()
    assert g.integrate() == f.integrate()
    for other in (mr.MultiResolutionAnalysis((-32, 32), 5), mr.MultiResolutionAnalysis((-16, 16), 7)):
        with pytest.raises(mr.MRCPPError):
            mr.FunctionTree(other).loadTree(base)
    with open(base + ".tree", "r+b") as fh:
        fh.truncate(os.path.getsize(base + ".tree") - 100)
    with pytest.raises(mr.MRCPPError):
        mr.FunctionTree(mra).loadTree(base)
    with pytest.raises(mr.MRCPPError):
        mr.FunctionTree(mra).loadTree(str(tmp_path / "missing"))


def test_rescale_on_the_host(mr, orc):
    """FunctionTree::rescale: end nodes times c, bottom-up transform, norms:
    every coefficient scales by c, norms by |c|, the integral by c."""
    _, mra, f, _, _ = _projected(mr, orc, order=5, box=16, prec=1e-5)
    i0, c0, n0 = f.export()
    q0, s0 = f.integrate(), f.getSquareNorm()
    f.rescale(-2.5)
    i1, c1, n1 = f.export()
    assert np.array_equal(i0, i1)
    assert np.abs(c1 + 2.5 * c0).max() <= 1e-14 * np.abs(c0).max()
    # component norms to 1e-12 of the node's norm (recompressed small wavelet
    # norms are differences of nearly equal values)
    scale = 2.5 * np.abs(n0[:, :8]).max(axis=1, keepdims=True)
    assert np.all(np.abs(n1[:, :8] - 2.5 * n0[:, :8]) <= 1e-12 * scale + 1e-15 * scale.max())
    assert f.integrate() == pytest.approx(-2.5 * q0, rel=1e-13)
    assert f.getSquareNorm() == pytest.approx(6.25 * s0, rel=1e-13)
