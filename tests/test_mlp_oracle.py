"""Pins of the numpy MLP oracle: its exact (float64, no bf16) backward pass
equals central finite differences of its forward pass, and the bf16
rounding helper is round-to-nearest-even on known bit patterns."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39, np.inf], np.float32)
    out = mlp.bf16(vals)
    # 1 + 2^-8 is exactly halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert out[0] == 1.0 and out[1] == 1.0
    # 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert out[2] == np.float32(1.0 + 2 ** -7)
    assert out[3] == -2.5 and out[5] == np.inf
    assert np.isnan(mlp.bf16(np.array([np.nan], np.float32)))[0]


def test_exact_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mlp.NPARAM) * 0.3
    X = rng.standard_normal((40, 16))
    T = rng.standard_normal(40)
    loss, g = mlp.loss_grad(p, X, T, 1.0 / 40, bf=False)
    idx = rng.choice(mlp.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [mlp.OFFS["b4"][0], mlp.OFFS["w4"][0]]])
    eps = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mlp.loss_grad(pp, X, T, 1.0 / 40, bf=False)[0] - mlp.loss_grad(pm, X, T, 1.0 / 40, bf=False)[0]) / (
            2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
