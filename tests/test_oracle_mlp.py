"""Pins the FP64 warm-start network oracle: analytic gradients against central finite
differences, loss decrease under Adam on a smooth target, the Adam update against its
closed form on one coordinate."""
import numpy as np


def test_gradient_matches_finite_differences():
    import mlp
    rng = np.random.default_rng(0)
    theta = mlp.init(1) + 0.05 * rng.standard_normal(mlp.n_params())
    x = rng.standard_normal((32, 8))
    t = rng.standard_normal(32)
    _, g = mlp.loss_and_grad(theta, x, t)
    idx = rng.choice(mlp.n_params(), 60, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros_like(theta)
        e[k] = h
        fd = (mlp.loss_and_grad(theta + e, x, t)[0] - mlp.loss_and_grad(theta - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(fd)), k


def test_adam_closed_form_and_training_progress():
    import mlp
    theta = np.array([1.0])
    g = np.array([0.5])
    th, m, v = mlp.adam(theta, g, np.zeros(1), np.zeros(1), 1, lr=0.1)
    # First bias-corrected step moves by lr·sign(g) (|m̂|/sqrt(v̂) = 1).
    assert abs(th[0] - (1.0 - 0.1 * 0.5 / (0.5 + 1e-8))) < 1e-12
    rng = np.random.default_rng(2)
    x = rng.standard_normal((512, 8))
    t = np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]
    theta = mlp.init(3)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    l0 = mlp.loss_and_grad(theta, x, t)[0]
    for step in range(1, 301):
        loss, g = mlp.loss_and_grad(theta, x, t)
        theta, m, v = mlp.adam(theta, g, m, v, step, lr=3e-3)
    assert loss < 0.2 * l0
