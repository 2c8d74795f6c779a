"""Pins the FP64 warm-start network oracle: analytic gradients against central finite
differences, loss decrease under Adam on a smooth target, the Adam update against its
closed form on one coordinate; the bf16 rounding on known values and the bf16-pipeline
statement against the FP64 gradient with its rounding switched off."""
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


def test_bf16_rounding_known_values():
    import mlp
    # bf16 spacing at 1 is 2^-7: 1 + 2^-8 is a tie (to even: 1), just above it rounds up.
    v = np.array([1.0, 1 + 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, 1 + 3 * 2.0 ** -8, -2.5,
                  3.0e38, 1e-30])
    r = mlp.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2.0 ** -7
    assert r[3] == 1 + 2.0 ** -6  # tie between 1 + 2^-7 and 1 + 2^-6: to the even one
    assert r[4] == -2.5
    assert abs(r[5] / 3.0e38 - 1) <= 2.0 ** -8 and abs(r[6] / 1e-30 - 1) <= 2.0 ** -8


def test_bf16_pipeline_without_rounding_is_the_network_gradient():
    # The bf16-pipeline statement with its rounding replaced by the identity keeps only the
    # FP32 conversions of the GPU step: within FP32 rounding of the FP64 gradient, block by
    # block (a dropped term or a transposed factor would be off by order 1).
    import mlp
    rng = np.random.default_rng(4)
    theta = mlp.init(2) + 0.02 * rng.standard_normal(mlp.n_params())
    x = rng.standard_normal((300, 8))
    t = rng.standard_normal(300)
    l0, g0 = mlp.loss_and_grad(theta, x, t)
    l1, g1 = mlp.loss_and_grad_bf16(theta, x.astype(np.float32), t, rnd=lambda a: a)
    assert abs(l1 - l0) <= 1e-6 * l0
    o = 0
    for shape in mlp.shapes():
        k = int(np.prod(shape))
        a, b = g1[o:o + k], g0[o:o + k]
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b), shape
        o += k
    # With the rounding, the gradient moves by bf16-sized amounts, not more.
    l2, g2 = mlp.loss_and_grad_bf16(theta, x, t)
    cos = g2 @ g0 / (np.linalg.norm(g2) * np.linalg.norm(g0))
    assert 0.99 < cos < 1 - 1e-9
