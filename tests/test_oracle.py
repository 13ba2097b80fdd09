"""The CPU oracle against the reference's known-answer tests and closed forms.

Pins: kernel term counts and accuracy (tests/operators/poisson_operator.cpp:47-127,
helmholtz_operator.cpp:48-129 of MRCPP), the Coulomb energy of a Gaussian
(poisson_operator.cpp:129-154), analytic Poisson and Yukawa potentials of a
Gaussian source, brute-force operator matrix elements, band-width monotonicity
and thread-count independence of apply.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg
from scipy.special import erf, erfc

from mrcpp_amd import workloads as W


def test_poisson_kernel_kat(orc):
    a, c = orc.Oracle.kernel_terms(0, 0.0, 1.0e-4, 1.0e-3, 1.0)
    assert len(a) == 26
    x = 1.0e-3
    while x < 1.0:
        v = np.sum(c * np.exp(-a * x * x))
        assert v == pytest.approx(1.0 / x, rel=2.0e-4)
        x *= 1.5


def test_helmholtz_kernel_kat(orc):
    mu = 0.01
    a, c = orc.Oracle.kernel_terms(1, mu, 1.0e-4, 1.0e-3, 1.0)
    assert len(a) == 33
    x = 1.0e-3
    while x < 1.0:
        v = np.sum(c * np.exp(-a * x * x))
        assert v == pytest.approx(math.exp(-mu * x) / x, rel=2.0e-4)
        x *= 1.5


@pytest.mark.parametrize("box,order,prec,mu,M", [
    (20, 5, 1e-5, None, 79), (32, 7, 1e-6, None, 96), (32, 7, 1e-6, 0.5, 78), (32, 7, 1e-6, 1.0, 75),
    (32, 7, 1e-6, 3.0, 71), (20, 9, 1e-7, None, 113), (20, 9, 1e-7, 1.0, 92)])
def test_separation_ranks(orc, box, order, prec, mu, M):
    """Separation ranks of the benchmark operators (closed-form kernel set-up:
    k_prec = prec/10, r_min = sqrt(k_prec 2^-maxScale), r_max = box diagonal)."""
    mra = orc.OracleMRA((-box, box), order)
    k_prec = prec / 10
    r_min = math.sqrt(k_prec * 2.0 ** (-mra.max_scale))
    r_max = math.sqrt(3) * 2 * box
    a, _ = orc.Oracle.kernel_terms(0 if mu is None else 1, mu or 0.0, k_prec, r_min, r_max)
    assert len(a) == M


def _phi(j, x):
    c = np.zeros(j + 1)
    c[j] = 1.0
    return np.sqrt(2 * j + 1) * npleg.legval(2 * x - 1, c)


def _psi(k, i, x):
    """Multiwavelet i of order k on [0,1] from the G filters: sqrt(2) (G0 phi(2x) + G1 phi(2x-1))."""
    import os
    from mrcpp_amd import FILTER_DIR
    K = k + 1
    G0 = np.fromfile(os.path.join(FILTER_DIR, "L_G0_%d" % k), dtype="<f8").reshape(K, K)
    G1 = (-1.0) ** np.add.outer(np.arange(K), np.arange(K)) * (-1.0) ** K * G0
    left = (x < 0.5)
    v = np.zeros_like(x)
    for j in range(K):
        v += np.where(left, G0[i, j] * _phi(j, 2 * x), G1[i, j] * _phi(j, 2 * x - 1))
    return np.sqrt(2.0) * v


# (term, depth, dl): terms resolved by a 120-point rule at that depth
@pytest.mark.parametrize("term,depth,dl", [(5, 0, 0), (20, 1, 1), (20, 2, -1), (20, 4, 2), (30, 4, -1), (30, 6, 1)])
def test_operator_blocks_brute_force(orc, term, depth, dl):
    """Component c = 2 g_bit + f_bit of term m at depth n and translation
    dl = l_f - l_g equals h int int u_i(s) v_j(t) K_m(h (dl + s - t)) ds dt,
    h = sf 2^-n, u (v) the scaling function or wavelet of the f (g) side and
    K_m(x) = cbrt(c_m) exp(-a_m x^2), the 1D factor of the separated term."""
    k = 5
    mra = orc.OracleMRA((-20, 20), k)
    op = orc.OracleOperator(mra, 1e-5)
    a, c = op.kernel()
    h = mra.scaling_factor * 2.0 ** (-(mra.root_scale + depth))
    xg, wg = npleg.leggauss(60)
    # split [0,1] at 1/2 so the piecewise wavelets are integrated exactly
    s = np.concatenate([(xg + 1) / 4, (xg + 1) / 4 + 0.5])
    ws = np.concatenate([wg / 4, wg / 4])
    S, Tt = np.meshgrid(s, s, indexing="ij")
    Wt = np.outer(ws, ws)
    K = np.cbrt(c[term]) * np.exp(-a[term] * (h * (dl + S - Tt)) ** 2)
    for comp in range(4):
        fbit, gbit = comp & 1, comp >> 1
        O, _ = op.node(term, depth, dl, comp)
        assert O is not None
        u = [(_psi(k, i, S) if fbit else _phi(i, S)) for i in range(k + 1)]
        v = [(_psi(k, j, Tt) if gbit else _phi(j, Tt)) for j in range(k + 1)]
        ref = np.array([[h * np.sum(Wt * u[i] * v[j] * K) for j in range(k + 1)] for i in range(k + 1)])
        scale = max(np.abs(ref).max(), 1e-300)
        # the kernel is projected to prec/10 before the cross-correlation
        assert np.abs(O - ref).max() < 1e-6 * max(scale, np.abs(op.node(term, depth, dl, 0)[0]).max()), comp
    T, _ = op.node(term, depth, dl, 0)
    if dl != 0:
        Tm, _ = op.node(term, depth, -dl, 0)
        assert np.abs(Tm - T.T).max() < 1e-12 * np.abs(T).max()  # T(-dl) = T(dl)^T for even kernels


def test_band_widths_monotone_in_precision(orc):
    """Looser precision never widens the band (poisson_operator.cpp:106-121)."""
    mra = orc.OracleMRA((-20, 20), 5)
    op = orc.OracleOperator(mra, 1e-4)
    b1 = op.bandwidths(1.0)[:, :, 4]
    b2 = op.bandwidths(1e-3)[:, :, 4]
    b3 = op.bandwidths(1e-6)[:, :, 4]
    assert np.all(b1 <= b2) and np.all(b2 <= b3)


@pytest.fixture(scope="module")
def gauss1(orc):
    """Unit-charge Gaussian beta=100 at the origin, k=5, [-20,20]^3, prec 1e-5."""
    w = W.WORKLOADS["gauss1_k5"]
    src = w["sources"]()
    mra = orc.OracleMRA((-20, 20), 5)
    f = orc.OracleTree(mra)
    f.project_gaussians(w["prec"], src.coefs, src.exponents, src.centers)
    return mra, f, w


def test_projection_charge(gauss1):
    _, f, _ = gauss1
    assert f.integrate() == pytest.approx(1.0, abs=1e-12)


def test_poisson_apply_closed_forms(orc, gauss1):
    mra, f, w = gauss1
    beta, prec = 100.0, w["prec"]
    P = orc.OracleOperator(mra, prec)
    g = orc.OracleTree(mra)
    st = orc.apply(prec, g, P, f)
    assert st["n_applied"] > 0
    # Coulomb self-energy of a normalised Gaussian: sqrt(2 beta / pi)
    E = g.dot(f)
    assert E == pytest.approx(math.sqrt(2 * beta / math.pi), rel=10 * prec)
    for r in (0.05, 0.1, 0.3, 1.0, 3.0, 7.0):
        p = np.array([r, 0.02, -0.03])
        rr = np.linalg.norm(p)
        # point values are accurate to the adaptive precision relative to the output norm
        assert abs(g.evalf(p) - erf(math.sqrt(beta) * rr) / rr) < 10 * prec * math.sqrt(g.square_norm())


def test_helmholtz_apply_yukawa(orc, gauss1):
    """Yukawa potential of a normalised Gaussian:
    V(r) = e^{mu^2/4b} / (2r) [e^{-mu r} erfc(mu/(2 sqrt b) - sqrt b r) - e^{mu r} erfc(mu/(2 sqrt b) + sqrt b r)]."""
    mra, f, w = gauss1
    beta, mu, prec = 100.0, 1.0, w["prec"]
    H = orc.OracleOperator(mra, prec, mu=mu)
    g = orc.OracleTree(mra)
    orc.apply(prec, g, H, f)
    sb = math.sqrt(beta)
    for r in (0.05, 0.2, 1.0, 3.0):
        p = np.array([r, -0.01, 0.02])
        rr = np.linalg.norm(p)
        v = math.exp(mu * mu / (4 * beta)) / (2 * rr) * (math.exp(-mu * rr) * erfc(mu / (2 * sb) - sb * rr)
                                                         - math.exp(mu * rr) * erfc(mu / (2 * sb) + sb * rr))
        assert abs(g.evalf(p) - v) < 10 * prec * math.sqrt(g.square_norm())
    # energy against the Gaussian itself: int int rho rho e^{-mu r}/r, by 1D radial quadrature of V rho
    rs = np.linspace(1e-6, 1.0, 20001)
    V = np.exp(mu * mu / (4 * beta)) / (2 * rs) * (np.exp(-mu * rs) * erfc(mu / (2 * sb) - sb * rs)
                                                    - np.exp(mu * rs) * erfc(mu / (2 * sb) + sb * rs))
    rho = (beta / math.pi) ** 1.5 * np.exp(-beta * rs * rs)
    E_ana = np.trapezoid(4 * math.pi * rs * rs * V * rho, rs)
    assert g.dot(f) == pytest.approx(E_ana, rel=1e-5)


def test_apply_independent_of_thread_count(orc, gauss1):
    mra, f, w = gauss1
    P = orc.OracleOperator(mra, w["prec"])
    res = []
    for nt in (1, 4):
        orc.Oracle.set_threads(nt)
        g = orc.OracleTree(mra)
        orc.apply(w["prec"], g, P, f)
        res.append(g.export())
    orc.Oracle.set_threads(8)
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


def test_max_iter_zero_keeps_grid(orc, gauss1):
    """maxIter = 0: one level, no refinement (apply.cpp docs: maxIter = 0 means NO refinement)."""
    mra, f, w = gauss1
    P = orc.OracleOperator(mra, w["prec"])
    g = orc.OracleTree(mra)
    st = orc.apply(w["prec"], g, P, f, max_iter=0)
    assert st["level_nodes"] == [8] and g.num_nodes() == 8
