"""The multiwavelet filter and cross-correlation tables against the mathematics
that defines them (Legendre scaling functions phi_j(x) = sqrt(2j+1) P_j(2x-1)
on [0,1], Alpert multiwavelets)."""
import os

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from mrcpp_amd import FILTER_DIR

ORDERS = [1, 3, 5, 7, 9]


def phi(j, x):
    c = np.zeros(j + 1)
    c[j] = 1.0
    v = np.sqrt(2 * j + 1) * npleg.legval(2 * x - 1, c)
    return np.where((x >= 0) & (x <= 1), v, 0.0)


def load(name, shape):
    return np.fromfile(os.path.join(FILTER_DIR, name), dtype="<f8").reshape(shape)


@pytest.mark.parametrize("k", ORDERS)
def test_h0_two_scale_relation(k):
    """H0(i,j) = int_0^{1/2} phi_i(x) sqrt(2) phi_j(2x) dx (exact quadrature)."""
    K = k + 1
    H0 = load("L_H0_%d" % k, (K, K))
    xg, wg = npleg.leggauss(2 * K + 4)
    x, w = (xg + 1) / 4, wg / 4
    ref = np.array([[np.sum(w * phi(i, x) * np.sqrt(2) * phi(j, 2 * x)) for j in range(K)] for i in range(K)])
    assert np.abs(H0 - ref).max() < 1e-13


@pytest.mark.parametrize("k", ORDERS)
def test_filter_bank_orthogonal(k):
    """[[H0 H1],[G0 G1]] is orthogonal with H1, G1 from the Legendre symmetry
    (src/core/MWFilter.cpp:237-243)."""
    K = k + 1
    H0 = load("L_H0_%d" % k, (K, K))
    G0 = load("L_G0_%d" % k, (K, K))
    s = (-1.0) ** np.add.outer(np.arange(K), np.arange(K))
    H1 = s * H0
    G1 = s * (-1.0) ** K * G0
    F = np.block([[H0, H1], [G0, G1]])
    assert np.abs(F @ F.T - np.eye(2 * K)).max() < 1e-13


@pytest.mark.parametrize("k", ORDERS)
def test_wavelets_vanishing_moments(k):
    """Row i of G0 defines a wavelet with K+i vanishing moments about 1/2
    (Alpert's construction), so wavelets are orthogonal to all polynomials of
    degree <= k and increasingly smooth."""
    K = k + 1
    G0 = load("L_G0_%d" % k, (K, K))
    s = (-1.0) ** np.add.outer(np.arange(K), np.arange(K))
    G1 = s * (-1.0) ** K * G0
    xg, wg = npleg.leggauss(60)
    xl, wl = (xg + 1) / 4, wg / 4
    xr = xl + 0.5
    for i in range(K):
        fl = sum(G0[i, j] * np.sqrt(2) * phi(j, 2 * xl) for j in range(K))
        fr = sum(G1[i, j] * np.sqrt(2) * phi(j, 2 * xr - 1) for j in range(K))
        for m in range(K + i):
            mom = np.sum(wl * fl * (xl - 0.5) ** m) + np.sum(wl * fr * (xr - 0.5) ** m)
            assert abs(mom) < 1e-12, (i, m)
        norm = np.sum(wl * fl * fl) + np.sum(wl * fr * fr)
        assert abs(norm - 1.0) < 1e-12


@pytest.mark.parametrize("k", [1, 3, 5, 7])
def test_cross_correlation_right(k):
    """C+_{ijp} = int_0^1 dz int_0^1 dx phi_i(x) phi_j(x - z) phi_p(z), with
    phi_p of order 2k+1 (src/core/CrossCorrelationCache.cpp:79-143)."""
    K = k + 1
    R = load("L_c_right_%d" % k, (K * K, 2 * K))
    xg, wg = npleg.leggauss(3 * K + 4)
    zs, wz = (xg + 1) / 2, wg / 2
    C = np.zeros((K * K, 2 * K))
    for z, wzz in zip(zs, wz):
        xs, wx = z + (1 - z) * (xg + 1) / 2, (1 - z) * wg / 2  # x in [z, 1] where phi_j(x - z) is supported
        pz = np.array([phi(p, np.array([z]))[0] for p in range(2 * K)])
        for i in range(K):
            for j in range(K):
                C[i * K + j] += wzz * np.sum(wx * phi(i, xs) * phi(j, xs - z)) * pz
    assert np.abs(C - R).max() < 1e-13


@pytest.mark.parametrize("k", [1, 3, 5, 7])
def test_cross_correlation_left(k):
    """C-_{ijp} = int_0^1 dz int_0^1 dx phi_i(x) phi_j(x - z + 1) phi_p(z): the
    left table pairs with the kernel node one to the left (shift z - 1)."""
    K = k + 1
    L = load("L_c_left_%d" % k, (K * K, 2 * K))
    xg, wg = npleg.leggauss(3 * K + 4)
    zs, wz = (xg + 1) / 2, wg / 2
    C = np.zeros((K * K, 2 * K))
    for z, wzz in zip(zs, wz):
        xs, wx = z * (xg + 1) / 2, z * wg / 2  # x in [0, z]: x - z + 1 in [0, 1]
        pz = np.array([phi(p, np.array([z]))[0] for p in range(2 * K)])
        for i in range(K):
            for j in range(K):
                C[i * K + j] += wzz * np.sum(wx * phi(i, xs) * phi(j, xs - z + 1)) * pz
    assert np.abs(C - L).max() < 1e-13
