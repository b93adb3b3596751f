"""Pins the numpy restatement of the router MLP (oracle/mlp.py) that checks the GPU
forward/backward kernels: central finite differences in float64, and the closed-form
gradient of the linear special case."""
import numpy as np

from oracle import mlp


def random_params(rng):
    p = rng.normal(0, 1, size=49)
    p[16:32] = rng.uniform(-1, 0.2, size=16)
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    for _ in range(5):
        p = random_params(rng)
        x = rng.uniform(0, 1, size=300)
        t = np.sort(rng.uniform(0, 1, size=300))
        loss, g = mlp.loss_and_grad(p, x, t)
        num = np.zeros(49)
        eps = 1e-6
        for q in range(49):
            d = np.zeros(49)
            d[q] = eps
            num[q] = (mlp.loss_and_grad(p + d, x, t)[0] - mlp.loss_and_grad(p - d, x, t)[0]) / (2 * eps)
        # a pre-activation within eps of a ReLU kink would make the difference one-sided
        z = np.outer(x, p[:16]) + p[16:32]
        if np.min(np.abs(z)) > 1e-4:
            np.testing.assert_allclose(g, num, rtol=1e-5, atol=1e-8)
        assert loss >= 0


def test_linear_special_case_closed_form():
    # One active unit with w1 = 1, b1 = 0 and w2 = a, b2 = b on x >= 0 is y = a x + b:
    # dL/da = mean((a x + b - t) x), dL/db = mean(a x + b - t).
    rng = np.random.default_rng(1)
    x = rng.uniform(0.01, 1, size=200)
    t = rng.uniform(0, 1, size=200)
    p = np.zeros(49)
    p[0] = 1.0          # w1[0]
    p[16 + 1:32] = -1.0  # other units stay off on x in [0, 1]
    a, b = 0.7, -0.2
    p[32] = a
    p[48] = b
    _, g = mlp.loss_and_grad(p, x, t)
    e = a * x + b - t
    assert np.isclose(g[32], np.mean(e * x))
    assert np.isclose(g[48], np.mean(e))
    assert np.isclose(g[0], np.mean(e * a * x))   # through w1[0]
    assert np.isclose(g[16], np.mean(e * a))      # through b1[0]
    assert np.allclose(g[33:48], 0.0)             # inactive units get nothing
