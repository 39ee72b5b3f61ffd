"""CPU oracle of the far-field corrector MLP (test infrastructure only).

Network: 8 -> 64 -> 64 -> 64 -> 3, GELU (tanh form) after each hidden layer.
Parameters are one flat float32 vector in this order (row-major [out][in]):
W1[64][8], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[3][64], b4[3]
(8896 weights + 195 biases = 9091).

The GPU kernels multiply BF16 operands on the matrix cores and accumulate
in FP32.  `forward`/`backward` here take `bf16=True` to round the same
operands at the same points (weights, layer inputs, back-propagated deltas)
and otherwise work in float64, so the GPU's gradients can be compared with a
tight tolerance; with `bf16=False` they are the exact float64 network, which
finite differences pin (tests/test_mlp_oracle.py).
"""
import numpy as np

SIZES = [(64, 8), (64, 64), (64, 64), (3, 64)]
N_PARAMS = sum(o * i + o for o, i in SIZES)
assert N_PARAMS == 9091

GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def unpack(params):
    p = np.asarray(params, np.float64)
    out, k = [], 0
    for o, i in SIZES:
        W = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def bf16(x):
    """Round to bfloat16 (nearest even) and back, as the GPU's conversion does."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def init_params(seed=0):
    """He-style normal initialisation (biases zero), float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    layers = []
    for o, i in SIZES:
        W = rng.standard_normal((o, i)) * np.sqrt(2.0 / i)
        layers.append((W, np.zeros(o)))
    return pack(layers).astype(np.float32)


def forward(params, X, bf16_ops=False):
    """Returns (Y, cache). X: (B, 8)."""
    q = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    h = q(X)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = h @ q(W).T + b
        cache.append((h, z))
        if li < len(layers) - 1:
            h = q(gelu(z))
    return z, cache


def backward(params, X, Y, bf16_ops=False):
    """Gradient of the loss mean_b sum_k (y_hat - y)^2 over the batch.
    Returns (loss, flat gradient in the parameter order)."""
    q = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params)
    out, cache = forward(params, X, bf16_ops)
    B = X.shape[0]
    diff = out - np.asarray(Y, np.float64)
    loss = float((diff * diff).sum() / B)
    dz = 2.0 * diff / B
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        h, z = cache[li]
        dzq = q(dz)
        gW = dzq.T @ h
        gb = dz.sum(0)
        grads.append((gW, gb))
        if li > 0:
            dh = dzq @ q(W)
            z_prev = cache[li - 1][1]
            dz = dh * gelu_grad(z_prev)
    grads.reverse()
    return loss, pack(grads)


def adam(params, m, v, grad, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) with bias correction; t counts from 1.
    Returns new (params, m, v) in float64."""
    p = np.asarray(params, np.float64)
    m = beta1 * np.asarray(m, np.float64) + (1 - beta1) * grad
    v = beta2 * np.asarray(v, np.float64) + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ---- features and targets of one (point, cell) interaction ------------------

def third_tensor(o):
    """The symmetric 3x3x3 tensor of the 10 third moments in the library's
    order: xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz."""
    xxx, yyy, zzz, xxy, xxz, xyy, yyz, xzz, yzz, xyz = (float(v) for v in o[:10])
    T = np.empty((3, 3, 3))
    val = {(0, 0, 0): xxx, (1, 1, 1): yyy, (2, 2, 2): zzz, (0, 0, 1): xxy, (0, 0, 2): xxz,
           (0, 1, 1): xyy, (1, 1, 2): yyz, (0, 2, 2): xzz, (1, 2, 2): yzz, (0, 1, 2): xyz}
    for i in range(3):
        for j in range(3):
            for k in range(3):
                T[i, j, k] = val[tuple(sorted((i, j, k)))]
    return T


def moments(bodies, com):
    """Second and third moments (3x3, 3x3x3) of bodies (x, y, z, m) about com."""
    b = np.asarray(bodies, np.float64)
    d = b[:, :3] - np.asarray(com, np.float64)
    m = b[:, 3]
    S = np.einsum("k,ki,kj->ij", m, d, d)
    O = np.einsum("k,ki,kj,kl->ijl", m, d, d, d)
    return S, O


def features(x, com, mass, size, O):
    """x: (3,) target point; com, mass, size: cell COM, mass, edge length;
    O (3,3,3): third moments about the COM.  Returns the 8 inputs:
    u = (com - x)/r (3), s/r, v = (7.5 w - 1.5 T) / (m s^3) (3) and
    a = u.w / (m s^3), with w_i = O_ijk u_j u_k and T_i = O_ijj."""
    d = np.asarray(com, np.float64) - np.asarray(x, np.float64)
    r = np.linalg.norm(d)
    u = d / r
    O = np.asarray(O, np.float64)
    k3 = mass * size ** 3
    w = np.einsum("ijk,j,k->i", O, u, u)
    T = np.einsum("ijj->i", O)
    return np.r_[u, size / r, (7.5 * w - 1.5 * T) / k3, (u @ w) / k3]


def octupole(f):
    """The cell's octupole term in target() units from its features:
    v + u (20 a - 5 u.v), the textbook term [7.5 w - 1.5 T - 17.5 (u.w) u +
    7.5 (T.u) u] / r^5 divided by m s^3 / r^5 (T.u = 5 a - u.v / 1.5)."""
    f = np.asarray(f, np.float64)
    u, v, a = f[..., 0:3], f[..., 4:7], f[..., 7:8]
    return v + u * (20.0 * a - 5.0 * (u * v).sum(-1, keepdims=True))


def target(x, bodies, com, mass, size, eps2, G=1.0):
    """Scaled correction: (exact - softened monopole) / (G m s^3 / r^5)."""
    x = np.asarray(x, np.float64)
    b = np.asarray(bodies, np.float64)
    d = b[:, :3] - x
    r2 = (d * d).sum(1) + eps2
    exact = G * (b[:, 3:4] * d / (r2 * np.sqrt(r2))[:, None]).sum(0)
    dc = np.asarray(com, np.float64) - x
    rc2 = dc @ dc
    mono = G * mass * dc / (rc2 + eps2) ** 1.5
    scale = G * mass * size ** 3 / rc2 ** 2.5
    return (exact - mono) / scale


def quadrupole_correction(x, com, mass, size, S, G=1.0, eps2=0.0):
    """The analytic quadrupole term (softened r as the walk evaluates it) in
    the same scaled units as target()."""
    d = np.asarray(com, np.float64) - np.asarray(x, np.float64)
    r2 = d @ d
    rs2 = r2 + eps2
    S = np.asarray(S, np.float64)
    sd = S @ d
    q = d @ sd
    a = G * ((7.5 * q / rs2 - 1.5 * np.trace(S)) * d - 3.0 * sd) / rs2 ** 2.5
    return a / (G * mass * size ** 3 / r2 ** 2.5)


def residual_target(x, bodies, com, mass, size, S, eps2, G=1.0):
    """What the network learns: target() minus the analytic quadrupole term."""
    return (target(x, bodies, com, mass, size, eps2, G)
            - quadrupole_correction(x, com, mass, size, S, G, eps2))
