"""The CPU oracle of the learned-guess network, pinned by central finite
differences and by the symmetries its inputs are built to respect."""
import numpy as np

from oracle import mlp


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(scale=0.3, size=mlp.NPARAMS)
    x = rng.normal(size=(16, 8))
    t = rng.normal(size=(16, 2))
    loss, g = mlp.loss_and_grad(p, x, t)
    idx = rng.choice(mlp.NPARAMS, size=60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (mlp.loss_and_grad(p + e, x, t)[0] - mlp.loss_and_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), i


def test_gelu_is_the_tanh_form():
    x = np.linspace(-6, 6, 101)
    ref = 0.5 * x * (1 + np.tanh(np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)))
    assert np.max(np.abs(mlp.gelu(x) - ref)) < 1e-14


def test_features_are_scale_and_rotation_aware():
    rng = np.random.default_rng(2)
    c = np.abs(rng.normal(size=(5, 8))) + 1.0
    c[:, 4] += 3.0
    f = mlp.features(c)
    # rho, p, B^2 -> lambda (rho, p, B^2) scales D, S, tau by lambda and B by sqrt(lambda)
    lam = 7.5
    c2 = c.copy()
    c2[:, :5] *= lam
    c2[:, 5:] *= np.sqrt(lam)
    f2 = mlp.features(c2)
    assert np.allclose(f2[:, 1:], f[:, 1:], rtol=1e-13, atol=1e-13)
    assert np.allclose(f2[:, 0] - f[:, 0], 0.2 * np.log(lam))
