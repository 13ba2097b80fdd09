"""The CPU restatement of the refinement predictor's network (oracle/mlp.py)
pinned by central finite differences, so it can check the device gradients."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = rng.normal(0.0, 0.3, mlp.N_PARAMS)
    X = rng.normal(0.0, 1.0, (9, mlp.F))
    Y = (rng.random(9) > 0.5).astype(float)
    g = mlp.gradient(p, X, Y, 2.5)
    idx = rng.choice(mlp.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [mlp.N_PARAMS - 1, 0, mlp.F * mlp.H]])  # b3, W1[0][0], b1[0]
    eps = 1e-6
    for i in idx:
        e = np.zeros(mlp.N_PARAMS)
        e[i] = eps
        fd = (mlp.loss(p + e, X, Y, 2.5) - mlp.loss(p - e, X, Y, 2.5)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_closed_form_cases():
    p = np.zeros(mlp.N_PARAMS)
    X = np.ones((4, mlp.F))
    # all weights zero: z = 0, p = 1/2, loss = mean(w y log 2 + (1 - y) log 2)
    Y = np.array([1.0, 0.0, 1.0, 1.0])
    assert np.isclose(mlp.loss(p, X, Y, 3.0), np.log(2.0) * (3 * 3.0 + 1) / 4)
    # only the output bias: z = b3
    p[-1] = 2.0
    z = 2.0
    assert np.isclose(mlp.loss(p, X, Y, 1.0), np.mean(np.where(Y > 0, np.log1p(np.exp(-z)), np.log1p(np.exp(z)))))
    g = mlp.gradient(p, X, Y, 1.0)
    assert np.isclose(g[-1], np.mean(mlp.sigmoid(z) - Y))
    assert np.allclose(g[:-1], 0.0)


def test_adam_first_step_is_signed_lr():
    p = np.ones(5)
    g = np.array([1.0, -2.0, 0.5, 0.0, 3.0])
    q, m, v = mlp.adam(p, g, np.zeros(5), np.zeros(5), 1, 0.01)
    assert np.allclose(q[g != 0], 1.0 - 0.01 * np.sign(g[g != 0]), atol=1e-7)
    assert q[3] == 1.0
