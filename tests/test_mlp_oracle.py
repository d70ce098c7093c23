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


def test_gelu_is_within_its_bound_of_the_exact_gelu():
    """The activation is GELU, x Phi(x) with Phi the normal distribution
    function (math.erf), to 1.3e-3 over |x| <= 40: inside the clamp, at it and
    beyond it on both sides (where it is x and 0)."""
    from math import erf, sqrt

    x = np.concatenate([np.linspace(-40.0, 40.0, 400_001), [-mlp.GELU_C, mlp.GELU_C]])
    exact = x * 0.5 * (1.0 + np.array([erf(v / sqrt(2.0)) for v in x]))
    err = np.abs(mlp.gelu(x) - exact)
    assert err.max() <= 1.3e-3
    # it is not a looser function that happens to fit: the error is that of
    # the fitted polynomial, largest near the clamp, tiny near zero
    assert err[np.abs(x) < 0.25].max() <= 5e-4
    # the clamp saturates: Phi~(c) is 1 (to the FP32 rounding of the
    # coefficients), so the activation is x far right and 0 far left
    assert abs(mlp.gelu(np.array([1e4]))[0] / 1e4 - 1.0) <= 1e-6
    assert abs(mlp.gelu(np.array([-1e4]))[0]) <= 1e-6 * 1e4


def test_gelu_grad_matches_central_differences():
    x = np.linspace(-6.0, 6.0, 2401)
    x = x[np.abs(np.abs(x) - mlp.GELU_C) > 1e-4]  # the clamp's kink
    h = 1e-6
    num = (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(num - mlp.gelu_grad(x))) <= 1e-8


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
