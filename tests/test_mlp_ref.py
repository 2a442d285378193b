"""Pins the CPU reference MLP (oracle/mlp_ref.py): analytic gradient vs
central finite differences of its float64 loss (CPU only)."""
import numpy as np

from oracle import mlp_ref


def _problem(seed, m=24):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.2, mlp_ref.NPARAMS)
    X = rng.normal(0, 1, (m, mlp_ref.NFEAT))
    y = (rng.uniform(size=m) < 0.4).astype(np.float64)
    c = rng.normal(3, 1, m)
    return p, X, y, c


def test_analytic_gradient_matches_finite_differences():
    p, X, y, c = _problem(0)
    lam = 0.3
    L, g = mlp_ref.loss_grad(p, X, y, c, lam, rounding=False)
    assert abs(L - mlp_ref.loss(p, X, y, c, lam, rounding=False)) < 1e-12
    rng = np.random.default_rng(1)
    idx = rng.choice(mlp_ref.NPARAMS, 300, replace=False)
    # include every bias and a few entries of each weight matrix
    idx = np.unique(np.concatenate([idx, np.arange(80 * 128, 80 * 128 + 128), np.arange(mlp_ref.NPARAMS - 2,
                                                                                         mlp_ref.NPARAMS)]))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp_ref.loss(p + e, X, y, c, lam, False) - mlp_ref.loss(p - e, X, y, c, lam, False)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e-5], np.float32)
    r = mlp_ref.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[3] == 1.0078125 and r[4] == -2.5
    assert r[2] == 1.0078125  # above the halfway point rounds up
    assert abs(r[5] - 3.0e-5) < 3.0e-5 * 2 ** -8
