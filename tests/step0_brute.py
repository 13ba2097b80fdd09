"""Brute-force restatement of the first breadth-first step of apply().

At the first step of apply() on a fresh output tree (MWTree squareNorm = -1,
src/trees/MWTree.cpp:58) the screening threshold is -1
(src/treebuilders/ConvolutionCalculator.cpp:241-246), so the norm test
`upperBound > gThreshold` (:323) passes for every tuple that survives the
other rules.  The applied tuples and the root coefficients then follow from
the band widths, the operator blocks and the input roots alone:

* every (g root, f root, ft, gt, term) with |f_ft| >= MachineZero
  (ConvolutionCalculator.cpp:258-260; (ft, gt) = (0, 0) allowed at depth 0);
* term i skipped when max_d |dl_d| exceeds its direction-0 maximum width
  (applyOperComp, :277-288);
* tuple skipped when some |dl_d| exceeds the width of component
  2 gbit_d + fbit_d (applyOperator / OperatorTree::isOutsideBand, :297-317,
  src/trees/OperatorTree.cpp:145-147);
* G[gt][j0, j1, j2] += sum F[ft][i0, i1, i2] Ox(i0, j0) Oy(i1, j1) Oz(i2, j2)
  with O(a, j) the column-major operator block of component 2 gbit + fbit at
  translation dl_d (tensorApplyOperComp, :333-382).

With maxIter = 0 the tree builder stops after this step and the transforms
have nothing to do (roots only), so the count and the coefficients are what
apply() returns.  Written from the reference's rules, independently of the
oracle's and the product's code; the operator blocks and widths come from
the tree being checked.
"""
import numpy as np

MACHINE_ZERO = 1.0e-14


def step0(roots_l, roots_coef, roots_norm, kp1, widths, block):
    """roots_l: [R, 3] translations of the world's roots (the input's and the
    output's); roots_coef: [R, 8 kp1^3] input root coefficients; roots_norm:
    [R, >= 8] their component norms; widths: [M, depth, 5] band widths
    (T, C, B, A, max) at the apply precision; block(term, l, comp) -> [kp1,
    kp1] operator block O(a, j) at depth 0.  Returns (applied count,
    [R, 8 kp1^3] output root coefficients in the same root order)."""
    M = widths.shape[0]
    n3 = kp1 ** 3
    R = len(roots_l)
    out = np.zeros((R, 8, kp1, kp1, kp1))  # [g][gt][j2][j1][j0]
    cache = {}

    def O(i, l, c):
        k = (i, l, c)
        if k not in cache:
            cache[k] = np.asarray(block(i, l, c), dtype=np.float64)
        return cache[k]

    count = 0
    w0 = widths[:, 0, :]  # depth 0
    for g in range(R):
        for f in range(R):
            dl = [int(roots_l[f][d] - roots_l[g][d]) for d in range(3)]
            maxdl = max(abs(x) for x in dl)
            for ft in range(8):
                if roots_norm[f][ft] < MACHINE_ZERO:
                    continue
                # F[i2][i1][i0] from the flat block (i0 fastest)
                F = roots_coef[f][ft * n3:(ft + 1) * n3].reshape(kp1, kp1, kp1)
                for gt in range(8):
                    comp = [2 * ((gt >> d) & 1) + ((ft >> d) & 1) for d in range(3)]
                    terms = [i for i in range(M) if maxdl <= w0[i, 4] and all(abs(dl[d]) <= w0[i, comp[d]] for d in range(3))]
                    if not terms:
                        continue
                    count += len(terms)
                    Ox = np.stack([O(i, dl[0], comp[0]) for i in terms])
                    Oy = np.stack([O(i, dl[1], comp[1]) for i in terms])
                    Oz = np.stack([O(i, dl[2], comp[2]) for i in terms])
                    # G[j2][j1][j0] = sum_t sum F[i2][i1][i0] Ox[t](i0,j0) Oy[t](i1,j1) Oz[t](i2,j2)
                    T1 = np.einsum("cba,tax->tcbx", F, Ox)
                    T2 = np.einsum("tcbx,tby->tcyx", T1, Oy)
                    out[g, gt] += np.einsum("tcyx,tcz->zyx", T2, Oz)
    return count, out.reshape(R, 8 * n3)
