"""CPU restatement of the refinement predictor's network, for tests only.

Test infrastructure: this module checks the device kernels mwg_mlp_forward,
mwg_mlp_gradient and mwg_mlp_adam (mrcpp_amd/csrc/mlp.hip); nothing in the
product imports it.  It has no reference counterpart (MRCPP has no
predictor): the forward pass, the weighted cross-entropy and its gradient are
restated from their definitions in float64, and the gradient is pinned
against central finite differences in tests/test_predictor_oracle.py.

Parameter layout (5057 floats): W1[12][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.
"""
import numpy as np

F, H = 12, 64
N_PARAMS = F * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    """Logits z and the hidden pre-activations."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    return z, (X, a1, h1, a2, h2)


def softplus(z):
    return np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(-np.abs(z))))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss(p, X, Y, pos_weight):
    z, _ = forward(p, X)
    Y = np.asarray(Y, dtype=np.float64)
    return float(np.mean(pos_weight * Y * softplus(-z) + (1.0 - Y) * softplus(z)))


def gradient(p, X, Y, pos_weight):
    """Gradient of loss() by backpropagation, as one flat vector."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z, (X, a1, h1, a2, h2) = forward(p, X)
    Y = np.asarray(Y, dtype=np.float64)
    n = len(Y)
    s = sigmoid(z)
    dz = (pos_weight * Y * (s - 1.0) + (1.0 - Y) * s) / n
    gw3 = h2.T @ dz
    gb3 = dz.sum()
    d2 = np.outer(dz, w3) * (a2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
