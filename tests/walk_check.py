"""CPU restatement of the production walk's interaction sets (test helper).

The production walk accepts node j for a body box [lo, hi] (the 64-body
group's bounding box, or the body itself for the per-body walk) when j is a
leaf or size_j^2 < theta^2 d^2, with d the distance from the box to j's
centre of mass and size_j = extent / 2^level_j -- or, under the mass-radius
opening rule, size_j replaced by the node's mass radius (mass_radii).  The
test is restated here in
float32 with the kernels' operation order (no fused multiply-add), so the
accepted set is the kernels' own; the forces of that set are then summed in
float64.  A GPU walk that matches this sum to FP32 rounding has the right
traversal, the right node records and the right evaluation formulas.

Node records are the walk arrays: com (x, y, z, m), link (first child, or
-(first body + 1) for a leaf; count << 5 | level), second moments
(Sxx, Syy, Szz, Sxy, Sxz, Syz, tr S, -) about the COM.
"""
import numpy as np

f32 = np.float32


def world_extent(lo, hi):
    w, h, d = (f32(hi[k]) - f32(lo[k]) for k in range(3))
    return f32(max(w, max(h, d)))


def accepted(com, link, extent, lo, hi, theta, rad=None):
    """Accepted node indices for the box [lo, hi]: (leaves, cells), sorted.
    rad: the nodes' mass radii for the mass-radius opening rule (b = rad in
    place of size in b^2 < theta^2 d^2), None for the size rule."""
    lo = np.asarray(lo, f32)
    hi = np.asarray(hi, f32)
    th2 = f32(theta) * f32(theta)
    frontier = np.array([0], np.int64)
    leaves, cells = [], []
    while frontier.size:
        c = com[frontier, :3].astype(f32)
        dd = np.maximum(np.maximum(lo - c, c - hi), f32(0))
        d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
        lk = link[frontier]
        level = lk[:, 1] & 31
        size = np.ldexp(f32(extent), -level).astype(f32)
        if rad is not None:
            size = np.asarray(rad, f32)[frontier]
        leaf = lk[:, 0] < 0
        acc = leaf | (size * size < th2 * d2)
        leaves.append(frontier[acc & leaf])
        cells.append(frontier[acc & ~leaf])
        op = ~acc
        first = lk[op, 0].astype(np.int64)
        cnt = (lk[op, 1] >> 5).astype(np.int64)
        if first.size:
            frontier = np.concatenate([np.arange(a, a + k) for a, k in zip(first, cnt)])
        else:
            frontier = np.zeros(0, np.int64)
    return np.sort(np.concatenate(leaves)), np.sort(np.concatenate(cells))


def mass_radii(com, link, pos):
    """Each node's mass radius as the build forms it, in float32 with its
    operation order: a leaf of one body 0, of several the largest
    sqrt((dx dx + dy dy) + dz dz) from its centre of mass to its bodies; an
    internal node the largest sqrt(...) + radius over its children, d from
    its centre of mass to the child's.  Nodes are level-major, so children
    follow their parents and a reverse sweep finishes them first."""
    com = np.asarray(com, f32)
    p = np.asarray(pos[:, :3], f32)
    M = com.shape[0]
    rad = np.zeros(M, f32)
    first = link[:, 0].astype(np.int64)
    cnt = (link[:, 1] >> 5).astype(np.int64)
    level = link[:, 1] & 31
    for L in np.unique(level)[::-1]:
        idx = np.flatnonzero(level == L)
        lf = idx[(first[idx] < 0) & (cnt[idx] > 1)]
        for j in lf:
            b0 = -first[j] - 1
            d = p[b0:b0 + cnt[j]] - com[j, :3]
            r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
            rad[j] = r.max()
        it = idx[first[idx] >= 0]
        if it.size:
            k = cnt[it]
            par = np.repeat(it, k)
            ch = np.concatenate([np.arange(a, a + b) for a, b in zip(first[it], k)])
            d = com[ch, :3] - com[par, :3]
            t = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) + rad[ch]
            starts = np.concatenate([[0], np.cumsum(k)[:-1]])
            rad[it] = np.maximum.reduceat(t.astype(f32), starts)
    return rad


def monopole(x, com, idx, eps2):
    d = com[idx, :3].astype(np.float64) - x
    r2 = (d * d).sum(1) + eps2
    return (com[idx, 3:4].astype(np.float64) * d / (r2 * np.sqrt(r2))[:, None]).sum(0)


def quadrupole(x, com, quad, idx, eps2):
    """Monopole plus quadrupole of cells idx at point x (softened r)."""
    d = com[idx, :3].astype(np.float64) - x
    m = com[idx, 3].astype(np.float64)
    q = quad[idx].astype(np.float64)
    S = np.empty((idx.size, 3, 3))
    S[:, 0, 0], S[:, 1, 1], S[:, 2, 2] = q[:, 0], q[:, 1], q[:, 2]
    S[:, 0, 1] = S[:, 1, 0] = q[:, 3]
    S[:, 0, 2] = S[:, 2, 0] = q[:, 4]
    S[:, 1, 2] = S[:, 2, 1] = q[:, 5]
    tr = q[:, 6]
    r2 = (d * d).sum(1) + eps2
    sd = np.einsum("kij,kj->ki", S, d)
    qq = (d * sd).sum(1)
    inv5 = r2 ** -2.5
    cd = m / (r2 * np.sqrt(r2)) + inv5 * (7.5 * qq / r2 - 1.5 * tr)
    return (cd[:, None] * d - 3.0 * inv5[:, None] * sd).sum(0)


def _pairings(n, k):
    """All ways to take k disjoint unordered pairs from positions 0..n-1:
    a list of (pairs, rest)."""
    import itertools
    out = []
    for pairs in itertools.combinations(itertools.combinations(range(n), 2), k):
        used = [p for pr in pairs for p in pr]
        if len(set(used)) == 2 * k:
            out.append((pairs, tuple(c for c in range(n) if c not in used)))
    return out


def _dfact(n):
    return 1 if n <= 0 else n * _dfact(n - 2)


def kernel_derivative(y, eps2, n):
    """n-th derivative tensor of K(y) = (|y|^2 + eps2)^(-1/2) for points y
    (k, 3): shape (k, 3, ..., 3).  Closed form: sum over j of
    (-1)^(n-j) (2n-2j-1)!! rho^-(2n-2j+1) times the sum, over the ways of
    pairing 2j of the n indices, of the Kronecker deltas of the pairs times
    the y components of the other indices."""
    y = np.asarray(y, np.float64)
    r2 = (y * y).sum(1) + eps2
    T = np.zeros((y.shape[0],) + (3,) * n)
    import itertools
    for j in range(n // 2 + 1):
        coef = (-1) ** (n - j) * _dfact(2 * n - 2 * j - 1)
        power = r2 ** (-(2 * n - 2 * j + 1) / 2.0)
        pats = _pairings(n, j)
        for ind in itertools.product(range(3), repeat=n):
            val = np.zeros(y.shape[0])
            for pairs, rest in pats:
                if all(ind[a] == ind[b] for a, b in pairs):
                    v = np.ones(y.shape[0])
                    for r in rest:
                        v = v * y[:, ind[r]]
                    val += v
            T[(slice(None),) + ind] += coef * power * val
    return T


def local_coefficients(g, com, quad, idx, eps2):
    """Local expansion about the point g of the field of cells idx (monopole
    and second moments S about their centres of mass): the symmetric tensors
    C1 (3), C2 (3x3), C3 (3x3x3), C4 (3x3x3x3) with
    a(g + delta) = C1 + C2 delta + C3 delta delta / 2 + C4 delta^3 / 6, where
    C_n = sum over cells of m D_n(y) + S_ab D_(n+2)(y)_(...ab) / 2 for n <= 3
    and C4 = sum of m D_4(y): y = g - centre of mass, D_n the n-th derivative
    tensor of the softened kernel (the field of a cell at x is
    m grad K(x - c) + S_ab grad d_a d_b K(x - c) / 2).  The truncation error
    is O(delta^4) for the monopoles and O(delta^3) for the second moments."""
    y = np.asarray(g, np.float64)[None, :] - com[idx, :3].astype(np.float64)
    m = com[idx, 3].astype(np.float64)
    q = quad[idx].astype(np.float64)
    S = np.empty((idx.size, 3, 3))
    S[:, 0, 0], S[:, 1, 1], S[:, 2, 2] = q[:, 0], q[:, 1], q[:, 2]
    S[:, 0, 1] = S[:, 1, 0] = q[:, 3]
    S[:, 0, 2] = S[:, 2, 0] = q[:, 4]
    S[:, 1, 2] = S[:, 2, 1] = q[:, 5]
    D = {n: kernel_derivative(y, eps2, n) for n in range(1, 6)}
    C1 = np.einsum("k,ki->i", m, D[1]) + 0.5 * np.einsum("kab,kiab->i", S, D[3])
    C2 = np.einsum("k,kij->ij", m, D[2]) + 0.5 * np.einsum("kab,kijab->ij", S, D[4])
    C3 = np.einsum("k,kijl->ijl", m, D[3]) + 0.5 * np.einsum("kab,kijlab->ijl", S, D[5])
    C4 = np.einsum("k,kijlm->ijlm", m, D[4])
    return C1, C2, C3, C4


def local_eval(C, delta):
    """The expansion C of local_coefficients at offsets delta (b, 3)."""
    C1, C2, C3, C4 = C
    d = np.asarray(delta, np.float64)
    return (C1[None, :] + np.einsum("ij,bj->bi", C2, d)
            + 0.5 * np.einsum("ijl,bj,bl->bi", C3, d, d)
            + np.einsum("ijlm,bj,bl,bm->bi", C4, d, d, d) / 6.0)


def far_split(cells, com, lo, hi, far):
    """The walk's far cells of a group box [lo, hi]: h^2 < far^2 R^2 with h
    the box's half diagonal and R the distance from the box centre to the
    cell's centre of mass, in float32 with the kernel's operation order.
    Returns (far mask over the cells, box centre g)."""
    lo = np.asarray(lo, f32)
    hi = np.asarray(hi, f32)
    g = (lo + hi) * f32(0.5)
    e = hi - lo
    h2 = f32(0.25) * ((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2])
    c = com[cells, :3].astype(f32)
    d = c - g
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    phi2 = f32(far) * f32(far)
    return h2 < phi2 * r2, g


def potential_fp64(x, box, com, link, quad, extent, theta, eps2, mode, rad=None):
    """Potential (without -G) at point x for the interaction set of `box`:
    sum m / r over monopoles; quadrupole cells (mode 1) add
    (3 d.S d - r^2 tr S) / (2 r^5); r softened, as the walk evaluates it."""
    lo, hi = box
    leaves, cells = accepted(com, link, extent, lo, hi, theta, rad)
    x = np.asarray(x, np.float64)
    mono = leaves if mode else np.concatenate([leaves, cells])
    d = com[mono, :3].astype(np.float64) - x
    r2 = (d * d).sum(1) + eps2
    phi = (com[mono, 3].astype(np.float64) / np.sqrt(r2)).sum()
    if mode and cells.size:
        d = com[cells, :3].astype(np.float64) - x
        m = com[cells, 3].astype(np.float64)
        q = quad[cells].astype(np.float64)
        sd = np.stack([q[:, 0] * d[:, 0] + q[:, 3] * d[:, 1] + q[:, 4] * d[:, 2],
                       q[:, 3] * d[:, 0] + q[:, 1] * d[:, 1] + q[:, 5] * d[:, 2],
                       q[:, 4] * d[:, 0] + q[:, 5] * d[:, 1] + q[:, 2] * d[:, 2]], 1)
        qq = (d * sd).sum(1)
        r2 = (d * d).sum(1) + eps2
        phi += (m / np.sqrt(r2) + (3 * qq - r2 * q[:, 6]) / (2 * r2 ** 2.5)).sum()
    return phi


def features(x, com, oct_, idx, extent, link, rad=None):
    """The network's 8 inputs for cells idx seen from point x, and k = m s^3:
    u (3), s/r, (7.5 w - 1.5 T) / (m s^3) (3), u.w / (m s^3), with
    w_i = O_ijk u_j u_k and T_i = O_ijj, from the 10 stored components.
    s is the cell size, or its mass radius under the mass-radius rule (rad)."""
    d = com[idx, :3].astype(np.float64) - x
    r = np.sqrt((d * d).sum(1))
    u = d / r[:, None]
    size = np.ldexp(np.float64(extent), -(link[idx, 1] & 31))
    if rad is not None:
        size = np.asarray(rad, np.float64)[idx]
    m = com[idx, 3].astype(np.float64)
    k = m * size ** 3
    o = oct_[idx].astype(np.float64)
    xxx, yyy, zzz, xxy, xxz, xyy, yyz, xzz, yzz, xyz = (o[:, j] for j in range(10))
    ux, uy, uz = u[:, 0], u[:, 1], u[:, 2]
    xx, yy, zz, xy2, xz2, yz2 = ux * ux, uy * uy, uz * uz, 2 * ux * uy, 2 * ux * uz, 2 * uy * uz
    w = np.stack([xxx * xx + xyy * yy + xzz * zz + xxy * xy2 + xxz * xz2 + xyz * yz2,
                  xxy * xx + yyy * yy + yzz * zz + xyy * xy2 + xyz * xz2 + yyz * yz2,
                  xxz * xx + yyz * yy + zzz * zz + xyz * xy2 + xzz * xz2 + yzz * yz2], 1)
    T = np.stack([xxx + xyy + xzz, xxy + yyy + yzz, xxz + yyz + zzz], 1)
    X = np.concatenate([u, (size / r)[:, None], (7.5 * w - 1.5 * T) / k[:, None],
                        ((u * w).sum(1) / k)[:, None]], 1)
    return X, k


def walk_fp64(x, box, com, link, quad, extent, theta, eps2, mode, net=None, min_ratio=0.0,
              oct_=None, far=0.0, rad=None):
    """Acceleration (without G) at point x for the interaction set of `box`.
    mode 0: all accepted nodes as monopoles; 1: cells as quadrupoles;
    2: as 1 plus the network's residual, evaluated at the box centre, scaled by
    m s^3 / r^5 with r from x (net: a callable (rows, 8) -> (rows, 3); oct_:
    the third moments, (nodes, 10+)).  far > 0 (modes 1 and 2): the cells of
    far_split take their monopole and quadrupole terms from the local
    expansion about the box centre, except, in mode 2, the cells the network
    corrects (their terms stay exact; the network's term is unchanged);
    leaves stay exact.  rad: mass radii for the mass-radius opening rule (then
    also the network's cell measure s)."""
    lo, hi = box
    leaves, cells = accepted(com, link, extent, lo, hi, theta, rad)
    x = np.asarray(x, np.float64)
    if mode == 0:
        return monopole(x, com, np.concatenate([leaves, cells]), eps2), leaves.size + cells.size
    corr = np.zeros(cells.size, bool)
    if mode == 2 and cells.size:
        ctr = 0.5 * (np.asarray(lo, np.float64) + np.asarray(hi, np.float64))
        X, k = features(ctr, com, oct_, cells, extent, link, rad)
        corr = (k > 0) & (X[:, 3] > min_ratio)
    if far > 0:
        fm, g = far_split(cells, com, lo, hi, far)
        fm &= ~corr
        a = monopole(x, com, leaves, eps2) + quadrupole(x, com, quad, cells[~fm], eps2)
        if fm.any():
            g = g.astype(np.float64)
            C = local_coefficients(g, com, quad, cells[fm], eps2)
            a = a + local_eval(C, (x - g)[None, :])[0]
    else:
        a = monopole(x, com, leaves, eps2) + quadrupole(x, com, quad, cells, eps2)
    if corr.any():
        o = net(X[corr])
        d = com[cells[corr], :3].astype(np.float64) - x
        r2 = (d * d).sum(1)
        a = a + ((k[corr] / (r2 * r2 * np.sqrt(r2)))[:, None] * o).sum(0)
    return a, leaves.size + cells.size


def split_ratio(theta):
    """The walk's default split rule: every run (a fraction above 1 means
    "always"; the opening angle does not change it)."""
    return f32(2.0)


def group_slots(keys, pos, ratio):
    """The walk's groups: each run of 64 consecutive bodies (sorted order) is
    split in two at its largest key jump (the adjacent pair whose keys differ
    in the highest bit; the first such pair): always when `ratio` > 1, else
    when the halves' box volumes add up to less than `ratio` of the run's,
    restated in float32 with the kernel's operation order.  Returns (slots, slot_of_body): slots is an int array
    (2 per run) of (first body, bodies), bodies 0 for an empty slot."""
    keys = np.asarray(keys, np.uint64)
    p = np.asarray(pos[:, :3], f32)
    n = keys.shape[0]
    runs = (n + 63) // 64
    slots = np.zeros((2 * runs, 2), np.int64)
    slot_of = np.empty(n, np.int64)
    for g in range(runs):
        s0 = 64 * g
        cnt = min(64, n - s0)
        j = 0
        if cnt > 1 and ratio > 0:
            k = keys[s0:s0 + cnt]
            x = k[:-1] ^ k[1:]
            hb = np.array([int(v).bit_length() for v in x])
            if hb.max() > 0 and ratio > 1:
                j = int(np.argmax(hb)) + 1
            elif hb.max() > 0:
                jj = int(np.argmax(hb)) + 1
                a, b = p[s0:s0 + jj], p[s0 + jj:s0 + cnt]
                amin, amax, bmin, bmax = a.min(0), a.max(0), b.min(0), b.max(0)
                da, db = amax - amin, bmax - bmin
                va = (da[0] * da[1]) * da[2]
                vb = (db[0] * db[1]) * db[2]
                lo, hi = np.minimum(amin, bmin), np.maximum(amax, bmax)
                d = hi - lo
                v = (d[0] * d[1]) * d[2]
                if f32(va + vb) < f32(ratio) * v:
                    j = jj
        if j:
            slots[2 * g] = (s0, j)
            slots[2 * g + 1] = (s0 + j, cnt - j)
            slot_of[s0:s0 + j] = 2 * g
            slot_of[s0 + j:s0 + cnt] = 2 * g + 1
        else:
            slots[2 * g] = (s0, cnt)
            slots[2 * g + 1] = (s0 + cnt, 0)
            slot_of[s0:s0 + cnt] = 2 * g
    return slots, slot_of
