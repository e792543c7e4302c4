"""Pins the CPU oracle of the ray-cost MLP (oracle/mlp.py): the analytic
float64 gradient against central finite differences, BF16 rounding against
exact cases, the BF16-rounding gradient against the float64 one, and the
first Adam step against its closed form (lr * sign(g) when eps is negligible)."""
import numpy as np

from oracle import mlp


def _net(seed, n=37):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.N_PARAMS)
    p[mlp.OFF_W1:mlp.OFF_B1] = rng.uniform(-0.6, 0.6, mlp.N_IN * mlp.N_H)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.uniform(-0.1, 0.1, mlp.N_H)
    p[mlp.OFF_W2:mlp.OFF_B2] = rng.uniform(-0.3, 0.3, mlp.N_H * mlp.N_H)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.uniform(-0.1, 0.1, mlp.N_H)
    p[mlp.OFF_W3:mlp.OFF_B3] = rng.uniform(-0.3, 0.3, mlp.N_H)
    p[mlp.OFF_B3] = 0.2
    x = rng.uniform(-1, 1, (n, mlp.N_IN))
    t = rng.uniform(0, 8, n)
    return p, x, t


def test_grad64_matches_finite_differences():
    p, x, t = _net(1)
    g = mlp.grad64(p, x, t)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(mlp.N_PARAMS, 80, replace=False), [mlp.OFF_B3, mlp.OFF_B1, mlp.OFF_W1]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (mlp.loss64(pp, x, t) - mlp.loss64(pm, x, t)) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])


def test_bf16_rounding_exact_cases():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -9: 1.0,              # below half an ulp
        1.0 + 2.0 ** -8: 1.0,              # tie, even stays
        1.0 + 3 * 2.0 ** -8: 1.0 + 2 * 2.0 ** -7,  # tie, odd rounds up
        -2.5: -2.5,
        0.0: 0.0,
    }
    for a, b in cases.items():
        assert mlp.bf16(np.float32(a)) == np.float32(b), (a, mlp.bf16(np.float32(a)), b)
    v = np.float32(1.0) + ulp
    assert mlp.bf16(v) == v and one < v


def test_bf16_gradient_tracks_float64():
    p, x, t = _net(3, n=256)
    x = mlp.bf16(x.astype(np.float32))
    g64 = mlp.grad64(p.astype(np.float32).astype(np.float64), x.astype(np.float64), t)
    _, gb = mlp.grad_bf16(p, x, t)
    cos = float(np.dot(g64, gb) / (np.linalg.norm(g64) * np.linalg.norm(gb)))
    assert cos > 0.999
    assert np.linalg.norm(gb - g64) / np.linalg.norm(g64) < 0.03


def test_adam_first_step_is_lr_sign():
    rng = np.random.default_rng(4)
    p = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    z = np.zeros(100, np.float32)
    p1, m, v = mlp.adam(p, g, z, z, 1, lr=1e-3)
    assert np.allclose(p1, p - np.float32(1e-3) * np.sign(g), atol=1e-7)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
