"""CPU restatement of the production walk's interaction sets (test helper).

The production walk accepts node j for a body box [lo, hi] (the 64-body
group's bounding box, or the body itself for the per-body walk) when j is a
leaf or size_j^2 < theta^2 d^2, with d the distance from the box to j's
centre of mass and size_j = extent / 2^level_j.  The test is restated here in
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


def accepted(com, link, extent, lo, hi, theta):
    """Accepted node indices for the box [lo, hi]: (leaves, cells), sorted."""
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


def potential_fp64(x, box, com, link, quad, extent, theta, eps2, mode):
    """Potential (without -G) at point x for the interaction set of `box`:
    sum m / r over monopoles; quadrupole cells (mode 1) add
    (3 d.S d - r^2 tr S) / (2 r^5); r softened, as the walk evaluates it."""
    lo, hi = box
    leaves, cells = accepted(com, link, extent, lo, hi, theta)
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


def features(x, com, oct_, idx, extent, link):
    """The network's 8 inputs for cells idx seen from point x, and k = m s^3:
    u (3), s/r, (7.5 w - 1.5 T) / (m s^3) (3), u.w / (m s^3), with
    w_i = O_ijk u_j u_k and T_i = O_ijj, from the 10 stored components."""
    d = com[idx, :3].astype(np.float64) - x
    r = np.sqrt((d * d).sum(1))
    u = d / r[:, None]
    size = np.ldexp(np.float64(extent), -(link[idx, 1] & 31))
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
              oct_=None):
    """Acceleration (without G) at point x for the interaction set of `box`.
    mode 0: all accepted nodes as monopoles; 1: cells as quadrupoles;
    2: as 1 plus the network's residual, evaluated at the box centre, scaled by
    m s^3 / r^5 with r from x (net: a callable (rows, 8) -> (rows, 3); oct_:
    the third moments, (nodes, 10+))."""
    lo, hi = box
    leaves, cells = accepted(com, link, extent, lo, hi, theta)
    x = np.asarray(x, np.float64)
    if mode == 0:
        return monopole(x, com, np.concatenate([leaves, cells]), eps2), leaves.size + cells.size
    a = monopole(x, com, leaves, eps2) + quadrupole(x, com, quad, cells, eps2)
    if mode == 2 and cells.size:
        ctr = 0.5 * (np.asarray(lo, np.float64) + np.asarray(hi, np.float64))
        X, k = features(ctr, com, oct_, cells, extent, link)
        ok = (k > 0) & (X[:, 3] > min_ratio)
        o = net(X[ok])
        d = com[cells[ok], :3].astype(np.float64) - x
        r2 = (d * d).sum(1)
        a = a + ((k[ok] / (r2 * r2 * np.sqrt(r2)))[:, None] * o).sum(0)
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
