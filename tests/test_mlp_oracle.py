"""Pins the corrector MLP's CPU oracle (CPU only).

Backward is pinned by central finite differences of the float64 network;
Adam by its closed form on a one-dimensional quadratic; the BF16 rounding
by known bit patterns; the scaled correction target by the analytic
quadrupole term it must approach far from a cell.
"""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_patterns():
    assert mlp.bf16(np.float32(1.0)) == 1.0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even keeps 1
    assert mlp.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3*2^-8 is a tie between 1+2^-7 and 1+2^-6: nearest even rounds up
    assert mlp.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert mlp.bf16(np.float32(-3.0)) == -3.0
    x = np.float32(0.1)
    assert abs(mlp.bf16(x) - 0.1) < 2 ** -9 * 0.1 * 2


def test_parameter_count_and_layout():
    p = mlp.init_params(1)
    assert p.shape == (9091,) and p.dtype == np.float32
    layers = mlp.unpack(p)
    assert [W.shape for W, _ in layers] == [(64, 8), (64, 64), (64, 64), (3, 64)]
    assert np.array_equal(mlp.pack(layers), p.astype(np.float64))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init_params(3).astype(np.float64)
    X = rng.standard_normal((7, 8))
    Y = rng.standard_normal((7, 3)) * 0.1
    loss, g = mlp.backward(p, X, Y)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = mlp.backward(p + e, X, Y)
        lm, _ = mlp.backward(p - e, X, Y)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_gelu_grad_matches_finite_differences():
    z = np.linspace(-5, 5, 101)
    fd = (mlp.gelu(z + 1e-6) - mlp.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(fd, mlp.gelu_grad(z), atol=1e-8)


def test_adam_closed_form():
    # first step of Adam moves every coordinate by lr * sign(grad) (bias-corrected)
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    p1, m, v = mlp.adam(p, np.zeros(3), np.zeros(3), g, 1, lr=1e-2, eps=0.0)
    assert np.allclose(p1, p - 1e-2 * np.sign(g))
    # minimising (x - 3)^2 converges
    x, m, v = np.array([0.0]), np.zeros(1), np.zeros(1)
    for t in range(1, 3000):
        x, m, v = mlp.adam(x, m, v, 2 * (x - 3.0), t, lr=1e-2)
    assert abs(x[0] - 3.0) < 1e-3


def test_target_approaches_quadrupole_far_away():
    rng = np.random.default_rng(5)
    bodies = np.c_[rng.random((200, 3)) * [1.0, 0.5, 0.2], np.full(200, 1 / 200)]
    m = bodies[:, 3].sum()
    com = (bodies[:, :3] * bodies[:, 3:]).sum(0) / m
    d = bodies[:, :3] - com
    S = (bodies[:, 3, None, None] * d[:, :, None] * d[:, None, :]).sum(0)
    size = 1.0
    for dist, tol in ((8.0, 0.1), (30.0, 0.03)):
        x = com + dist * np.array([0.6, 0.64, 0.48])
        t = mlp.target(x, bodies, com, m, size, 0.0)
        qc = mlp.quadrupole_correction(x, com, m, size, S)
        # the remainder is the octupole and higher: O(s/r) relative to the quadrupole
        assert np.linalg.norm(t - qc) <= tol * np.linalg.norm(qc), dist
    r = mlp.residual_target(x, bodies, com, m, size, S, 0.0)
    assert np.linalg.norm(r) <= 0.03 * np.linalg.norm(t)
    f = mlp.features(x, com, m, size, S)
    assert f.shape == (8,) and abs(np.linalg.norm(f[:3]) - 1) < 1e-12
    assert abs(f[3] - size / np.linalg.norm(com - x)) < 1e-12
