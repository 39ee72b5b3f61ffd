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
    S2, O = mlp.moments(bodies, com)
    assert np.allclose(S2, S, rtol=1e-12, atol=1e-15)
    f = mlp.features(x, com, m, size, O)
    assert f.shape == (8,) and abs(np.linalg.norm(f[:3]) - 1) < 1e-12
    assert abs(f[3] - size / np.linalg.norm(com - x)) < 1e-12


def test_third_moments_shift_and_octupole_term():
    """The third moments are symmetric and shift by the parallel-axis rule;
    the octupole field they define (built from O:uu and friends by the
    textbook expansion) closes most of what the quadrupole leaves."""
    rng = np.random.default_rng(11)
    b = np.c_[rng.random((400, 3)) * [1.0, 0.4, 0.7] + [[0.3, 0.0, 0.0]] * (rng.random((400, 1)) < 0.3),
              rng.random(400) + 0.2]
    m = b[:, 3].sum()
    c = (b[:, :3] * b[:, 3:]).sum(0) / m
    S, O = mlp.moments(b, c)
    assert np.allclose(O, O.transpose(1, 0, 2)) and np.allclose(O, O.transpose(0, 2, 1))
    # tensor from its 10 components
    v10 = [O[0, 0, 0], O[1, 1, 1], O[2, 2, 2], O[0, 0, 1], O[0, 0, 2], O[0, 1, 1], O[1, 1, 2],
           O[0, 2, 2], O[1, 2, 2], O[0, 1, 2]]
    assert np.allclose(mlp.third_tensor(v10), O)
    # parallel-axis: moments of two halves combined equal the whole's
    h = b[:200], b[200:]
    tot = np.zeros((3, 3, 3))
    for part in h:
        mc = part[:, 3].sum()
        cc = (part[:, :3] * part[:, 3:]).sum(0) / mc
        Sc, Oc = mlp.moments(part, cc)
        dd = cc - c
        tot += (Oc + np.einsum("ij,k->ijk", Sc, dd) + np.einsum("ik,j->ijk", Sc, dd)
                + np.einsum("jk,i->ijk", Sc, dd) + mc * np.einsum("i,j,k->ijk", dd, dd, dd))
    assert np.allclose(tot, O, rtol=1e-10, atol=1e-12)
    # octupole field, from the third-order term of 1/|R - e| (R = x - com,
    # u = -R/r): a = [7.5 O:uu - 1.5 T - 17.5 (O:uuu) u + 7.5 (T.u) u] / r^5,
    # T_i = O_ijj.  The residual beyond the quadrupole is this term plus
    # O((s/r)^4) relative corrections.
    left = {}
    for dist in (6.0, 12.0):
        x = c + dist * np.array([0.36, 0.48, 0.8])
        res = mlp.residual_target(x, b, c, m, 1.0, S, 0.0)
        d = c - x
        r = np.linalg.norm(d)
        u = d / r
        ouu = np.einsum("ijk,j,k->i", O, u, u)
        ouuu = ouu @ u
        tr_u = np.einsum("ijj->i", O) @ u   # O_ijj u_i
        trv = np.einsum("ijj->i", O)        # O_ijj
        oct = (7.5 * ouu - 1.5 * trv - 17.5 * ouuu * u + 7.5 * tr_u * u) / r ** 5
        scale = m / r ** 5     # m s^3 / r^5 with s = 1
        left[dist] = np.linalg.norm(res - oct / scale) / np.linalg.norm(res)
        # the features carry the whole octupole term
        f = mlp.features(x, c, m, 1.0, O)
        assert np.allclose(mlp.octupole(f), oct / scale, rtol=1e-10, atol=1e-13)
    # 95 % of the residual is the octupole, and what is left is O(s/r) of it
    assert left[6.0] < 0.08 and left[12.0] < 0.6 * left[6.0], left


def test_features_equal_sums_over_bodies():
    """The moment features, restated as sums over the cell's bodies instead of
    tensor contractions: (O:uu)_i = sum m (d.u)^2 d_i, T_i = sum m |d|^2 d_i,
    u.O:uu = sum m (d.u)^3, with d = body - com."""
    rng = np.random.default_rng(4)
    b = np.c_[rng.random((150, 3)) * [0.8, 0.5, 0.3], rng.random(150) + 0.1]
    m = b[:, 3].sum()
    c = (b[:, :3] * b[:, 3:]).sum(0) / m
    S, O = mlp.moments(b, c)
    size = 0.9
    x = c + np.array([3.0, -2.0, 1.5])
    f = mlp.features(x, c, m, size, O)
    u = (c - x) / np.linalg.norm(c - x)
    d = b[:, :3] - c
    du = d @ u
    ouu = (b[:, 3, None] * (du ** 2)[:, None] * d).sum(0)
    tv = (b[:, 3, None] * (d * d).sum(1)[:, None] * d).sum(0)
    uuu = (b[:, 3] * du ** 3).sum()
    assert np.allclose(f[:3], u, rtol=1e-14, atol=1e-15)
    assert abs(f[3] - size / np.linalg.norm(c - x)) < 1e-14
    assert np.allclose(f[4:7] * m * size ** 3, 7.5 * ouu - 1.5 * tv, rtol=1e-12, atol=1e-14)
    assert abs(f[7] * m * size ** 3 - uuu) <= 1e-12 * abs(uuu)
