"""Pins the FP64 warm-start network oracle: analytic gradients against central finite
differences, loss decrease under Adam on a smooth target, the Adam update against its
closed form on one coordinate."""
import numpy as np


def active_pattern(mlp, theta, x):
    _, (_, h1, h2, h3) = mlp.forward(theta, x)
    return np.concatenate([(h1 > 0).ravel(), (h2 > 0).ravel(), (h3 > 0).ravel()])


def test_gradient_matches_finite_differences():
    # The loss is smooth between ReLU kinks: central differences whose two evaluations keep
    # every unit on the same side of its kink must equal the analytic gradient.
    import mlp
    rng = np.random.default_rng(0)
    theta = mlp.init(1) + 0.05 * rng.standard_normal(mlp.n_params())
    x = rng.standard_normal((32, 8))
    t = rng.standard_normal(32)
    _, g = mlp.loss_and_grad(theta, x, t)
    base = active_pattern(mlp, theta, x)
    idx = rng.choice(mlp.n_params(), 80, replace=False)
    h = 1e-6
    checked = 0
    for k in idx:
        e = np.zeros_like(theta)
        e[k] = h
        if not (np.array_equal(active_pattern(mlp, theta + e, x), base) and
                np.array_equal(active_pattern(mlp, theta - e, x), base)):
            continue
        fd = (mlp.loss_and_grad(theta + e, x, t)[0] - mlp.loss_and_grad(theta - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(fd)), k
        checked += 1
    assert checked >= 60
    # Units switched off get no gradient through them: a W1 row of a unit that is inactive
    # for every sample has a zero gradient.
    W1, b1 = mlp.unflatten(theta)[:2]
    dead = np.all(x @ W1.T + b1 <= 0, 0)
    gW1 = g[:512].reshape(64, 8)
    assert np.all(gW1[dead] == 0)


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
