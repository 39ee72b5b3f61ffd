"""Independent restatement of the reference walk's acceptance rule (CPU).

The reference walk (tree.cu:562-676) gives every 32-body group (blocks of
WARPTHREADS = 32 consecutive sorted bodies, basicGravity.cu:396-402, 779) one
set of accepted nodes.  Read as a definition rather than as a stack program:

  * the root's non-empty children are visited (tree.cu:595-601);
  * a visited node is opened iff some valid lane (body index < N,
    tree.cu:573-583) wants to descend, and a lane wants to descend iff the
    node is not a leaf (particleCount <= 1 or no children, tree.cu:627) and
    it fails size^2 < theta^2 d^2 (tree.cu:629-634), with
    size = x extent of the global box / 2^level (tree.cu:501-530) and d the
    offset from the body to the node's centre of mass (tree.cu:610-611);
  * the children of an opened node are visited (tree.cu:640-651);
  * a visited node that is not opened is accepted by every lane of the
    group (tree.cu:660-668).

The accepted set does not depend on the order of the stack, so this module
enumerates it level by level for all groups at once (numpy), instead of
replaying the LIFO stack.  The node size is the x extent scaled by 2^-level
(ldexp; exact, as is the reference's float / double pow), and the squared
distance follows the documented operation order fmaf(dz, dz, fmaf(dy, dy,
dx * dx)) with each fmaf rounded once (fmaf32 below), so that ties and
last-ulp flips of the test come out exactly as in float32.
"""
import numpy as np

GROUP = 32


def fmaf32(a, b, c):
    """float32 fma with one rounding: a * b + c for float32 arrays.

    a * b is exact in float64 (48 significant bits); s = fl64(p + c) plus its
    error e from TwoSum is the exact sum.  Rounding s to odd (if e != 0 and
    s's last bit is even, step s one ulp toward e) keeps the information of
    e, and a float64 rounded to odd rounds correctly to float32 (53 >= 24 + 2;
    Boldo and Melquiond, "Emulation of FMA and correctly rounded sums: proved
    algorithms using rounding to odd", IEEE TC 2008)."""
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    c = np.asarray(c, np.float32).astype(np.float64)
    p = a * b
    s = p + c
    bp = s - c
    e = (p - bp) + (c - (s - bp))              # TwoSum error of s = p + c
    bits = s.view(np.int64).copy()
    fix = (e != 0) & ((bits & 1) == 0) & np.isfinite(s)
    # a step toward e: away from zero if e has s's sign, else toward zero
    away = np.signbit(e) == np.signbit(s)
    bits = np.where(fix, np.where(away, bits + 1, bits - 1), bits)
    return bits.view(np.float64).astype(np.float32)


def leaf_flags(nodes):
    """tree.cu:627 and isLeafCalc (tree.cu:532-545)."""
    return (nodes["count"] <= 1) | (nodes["child"] == -1).all(1)


def group_accepted_sets(pos, nodes, bmin, bmax, theta):
    """Accepted node set of every 32-body group.

    pos: sorted bodies (n, 4) float32; nodes: the node array (only the first
    `total` entries are reached); bmin / bmax: the global box corners.
    Returns (sets, opened_tests, ties): sets[g] a sorted int array of node
    indices; opened_tests the number of (valid lane, opened node) pairs whose
    lane wanted to descend; ties the number of valid-lane tests of non-leaf
    nodes with size^2 == theta^2 d^2 exactly."""
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    groups = (n + GROUP - 1) // GROUP
    ext = np.float32(np.float32(bmax[0]) - np.float32(bmin[0]))
    th2 = np.float32(np.float32(theta) * np.float32(theta))
    leaf = leaf_flags(nodes)
    child = nodes["child"]
    level = nodes["level"]
    com = nodes["mass"][:, :3].astype(np.float32)

    lane = np.arange(GROUP)
    body = np.arange(groups)[:, None] * GROUP + lane[None, :]          # (groups, 32)
    valid = body < n
    padded = np.zeros((groups * GROUP, 3), np.float32)
    padded[:n] = pos[:, :3]                                             # invalid lanes: 0
    lanes_pos = padded.reshape(groups, GROUP, 3)

    root_kids = child[0][child[0] != -1]
    fg = np.repeat(np.arange(groups), root_kids.size)
    fv = np.tile(root_kids, groups)
    accepted_g, accepted_v = [], []
    opened_tests = 0
    ties = 0
    while fg.size:
        p = lanes_pos[fg]                                               # (P, 32, 3)
        c = com[fv][:, None, :]
        d = p - c                                                       # float32
        d2 = fmaf32(d[..., 2], d[..., 2], fmaf32(d[..., 1], d[..., 1], d[..., 0] * d[..., 0]))
        size = np.ldexp(np.full(fv.size, ext, np.float32), -level[fv].astype(np.int32))
        assert size.dtype == np.float32
        s2 = (size * size)[:, None]
        rhs = (th2 * d2).astype(np.float32)
        v = valid[fg]
        inner = ~leaf[fv][:, None]
        want = v & inner & ~(s2 < rhs)
        ties += int((v & inner & (s2 == rhs)).sum())
        opened = want.any(1)
        opened_tests += int(want[opened].sum())
        accepted_g.append(fg[~opened])
        accepted_v.append(fv[~opened])
        og, ov = fg[opened], fv[opened]
        kids = child[ov]                                                # (Q, 8)
        has = kids != -1
        fg = np.repeat(og, has.sum(1))
        fv = kids[has]
    ag = np.concatenate(accepted_g) if accepted_g else np.zeros(0, np.int64)
    av = np.concatenate(accepted_v) if accepted_v else np.zeros(0, np.int64)
    order = np.lexsort((av, ag))
    ag, av = ag[order], av[order]
    bounds = np.searchsorted(ag, np.arange(groups + 1))
    sets = [av[bounds[g]:bounds[g + 1]] for g in range(groups)]
    return sets, opened_tests, ties


def check_partition(sets, nodes, n):
    """Every group's accepted nodes cover the bodies [0, n) exactly once
    (their body ranges [first, first + count) tile [0, n)).  Returns the
    groups that fail."""
    bad = []
    first = nodes["first"]
    count = nodes["count"]
    for g, s in enumerate(sets):
        f = first[s].astype(np.int64)
        e = f + count[s]
        o = np.argsort(f, kind="stable")
        f, e = f[o], e[o]
        ok = f.size > 0 and f[0] == 0 and e[-1] == n and np.array_equal(f[1:], e[:-1])
        if not ok:
            bad.append(g)
    return bad


def compare_with_oracle_lists(sets, lists, n):
    """Bodies whose recorded accepted nodes (the oracle's traced walk) differ
    from their group's set."""
    bad = []
    for i in range(n):
        if not np.array_equal(np.sort(lists[i]), sets[i // GROUP]):
            bad.append(i)
    return bad


def monopole_fp64(pos, nodes, sets, eps2, G=1.0):
    """FP64 softened monopole sum of every body over its group's set, and the
    sum of the terms' magnitudes (the scale of FP32 summation error)."""
    n = pos.shape[0]
    acc = np.zeros((n, 3))
    mag = np.zeros(n)
    cm = nodes["mass"].astype(np.float64)
    p64 = pos[:, :3].astype(np.float64)
    for g, s in enumerate(sets):
        i0, i1 = g * GROUP, min(n, g * GROUP + GROUP)
        d = p64[i0:i1, None, :] - cm[None, s, :3]
        r2 = (d * d).sum(-1) + eps2
        f = -G * cm[None, s, 3] / (r2 * np.sqrt(r2))
        t = f[..., None] * d
        acc[i0:i1] = t.sum(1)
        mag[i0:i1] = np.linalg.norm(t, axis=2).sum(1)
    return acc, mag


def lattice_fixture(drop=13, seed=7):
    """Bodies on a 16 x 16 x 16 lattice, x and z spacing 1 and y spacing 2
    (the x extent, 15, is not the widest: 30), unit masses, so that centres
    of mass are dyadic and exact ties of the opening test occur; `drop`
    bodies removed so the last 32-body group has invalid lanes, and two
    stacks of coincident bodies (multi-body leaves at level 21).  Unsorted."""
    g = np.arange(16, dtype=np.float32)
    x, y, z = np.meshgrid(g, 2 * g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel(), np.ones(4096, np.float32)], 1)
    rng = np.random.default_rng(seed)
    keep = np.sort(rng.choice(4096, 4096 - drop, replace=False))
    p = p[keep]
    dup = np.concatenate([np.repeat(p[100:101], 10, 0), np.repeat(p[2000:2001], 5, 0)])
    return np.ascontiguousarray(np.concatenate([p, dup]).astype(np.float32))
