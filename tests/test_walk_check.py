"""Pins the CPU restatement of the production walk's interaction sets
(tests/walk_check.py) on trees built by the oracle (CPU only).

Checks from what the method fixes: at theta = 0 the set is every leaf and its
monopole sum is the direct sum; for any theta and box the accepted nodes
partition the bodies, every accepted cell passes the opening test and every
ancestor of an accepted node fails it; the quadrupole sum of a cell matches
the brute-force field of its bodies to O((s/r)^3) while the monopole is off
at O((s/r)^2).
"""
import numpy as np
import pytest

from cudagravitydemo_amd import ics
from oracle import binding as ob
from oracle import mlp
from tests.walk_check import (accepted, far_split, features, group_slots, kernel_derivative,
                               local_coefficients, local_eval, mass_radii, monopole, quadrupole,
                               split_ratio, walk_fp64, world_extent)


@pytest.fixture(scope="module")
def tree():
    pos, vel = ics.plummer(3000, seed=4)
    t = ob.build_tree(pos, vel)
    nodes = t["nodes"]
    total = nodes.shape[0]
    com = nodes["mass"].astype(np.float32)
    link = np.zeros((total, 2), np.int32)
    for j in range(total):
        ch = nodes["child"][j]
        ch = np.sort(ch[ch >= 0])
        if ch.size == 0:
            link[j, 0] = -(nodes["first"][j] + 1)
        else:
            # children of a node are consecutive in the level-wise layout
            assert np.array_equal(ch, np.arange(ch[0], ch[0] + ch.size))
            link[j, 0] = ch[0]
        # count: children of an internal node, bodies of a leaf (the walk arrays' layout)
        k = ch.size if ch.size else int(nodes["count"][j])
        link[j, 1] = (k << 5) | nodes["level"][j]
    ext = world_extent(t["bmin"], t["bmax"])
    return dict(t=t, nodes=nodes, com=com, link=link, ext=ext)


def _boxes(ps):
    rng = np.random.default_rng(3)
    out = []
    for i in rng.choice(ps.shape[0], 6, replace=False):
        out.append((ps[i, :3], ps[i, :3]))
    for g in rng.choice(ps.shape[0] // 64, 6, replace=False):
        blk = ps[64 * g:64 * g + 64, :3]
        out.append((blk.min(0), blk.max(0)))
    return out


def test_theta_zero_is_every_leaf_and_the_direct_sum(tree):
    t, com, link = tree["t"], tree["com"], tree["link"]
    ps = t["pos"]
    leaf_nodes = np.flatnonzero(link[:, 0] < 0)
    for lo, hi in _boxes(ps)[:3]:
        lv, cl = accepted(com, link, tree["ext"], lo, hi, 0.0)
        assert cl.size == 0 and np.array_equal(lv, leaf_nodes)
    x = ps[17, :3].astype(np.float64)
    a = monopole(x, com, leaf_nodes, 1e-4)
    want = ob.direct_sum_fp64(ps, np.array([17], np.int32), 1e-4)[0]
    # leaf COMs are float32 (x m) / m: equal to the body position to FP32 rounding
    assert np.linalg.norm(a - want) <= 1e-6 * np.linalg.norm(want)


@pytest.mark.parametrize("theta", [0.3, 0.7, 1.2])
def test_accepted_nodes_partition_bodies_and_pass_the_test(tree, theta):
    t, nodes, com, link, ext = tree["t"], tree["nodes"], tree["com"], tree["link"], tree["ext"]
    ps = t["pos"]
    n = ps.shape[0]
    for lo, hi in _boxes(ps):
        lv, cl = accepted(com, link, ext, lo, hi, theta)
        acc = np.concatenate([lv, cl])
        first, cnt = nodes["first"][acc], nodes["count"][acc]
        cover = np.zeros(n, np.int64)
        for a, k in zip(first, cnt):
            cover[a:a + k] += 1
        assert (cover == 1).all()

        def crit(j):
            c = com[j, :3].astype(np.float64)
            d = np.maximum(np.maximum(lo - c, c - hi), 0.0)
            size = np.ldexp(float(ext), -int(nodes["level"][j]))
            return size * size, theta * theta * (d @ d)

        for j in cl:
            s2, rhs = crit(j)
            assert s2 <= rhs * (1 + 1e-5)
        for j in acc:
            p = int(nodes["parent"][j])
            if p >= 0 and j != 0:
                s2, rhs = crit(p)
                assert s2 >= rhs * (1 - 1e-5), (j, p)


def _exact_radii(nodes, com, ps):
    """Largest distance from each node's centre of mass to its bodies (FP64)."""
    out = np.zeros(nodes.shape[0])
    for j in range(nodes.shape[0]):
        a, k = int(nodes["first"][j]), int(nodes["count"][j])
        d = ps[a:a + k, :3].astype(np.float64) - com[j, :3].astype(np.float64)
        out[j] = np.sqrt((d * d).sum(1)).max() if k else 0.0
    return out


def test_mass_radii_bound_every_body(tree):
    """Each radius bounds the distance from the node's centre of mass to
    every one of its bodies (up to float32 rounding of the coordinates: a
    one-body leaf's centre of mass, (x m) / m, sits an ulp or so off its body,
    and its radius is 0), and a node whose children are one-body leaves gets
    the exact largest distance.  A recursion that drops the children's own
    radii (a plausible slip) breaks the bound."""
    t, nodes, com, link = tree["t"], tree["nodes"], tree["com"], tree["link"]
    ps = t["pos"]
    rad = mass_radii(com, link, ps).astype(np.float64)
    exact = _exact_radii(nodes, com, ps)
    ulps = 4e-7 * (1.0 + np.abs(com[:, :3].astype(np.float64)).max(1))
    assert (rad >= exact * (1 - 1e-6) - ulps).all()
    cnt = link[:, 1] >> 5
    single = (link[:, 0] < 0) & (cnt == 1)
    assert (rad[single] == 0).all()
    internal = np.flatnonzero(link[:, 0] >= 0)
    flat = [j for j in internal
            if (link[link[j, 0]:link[j, 0] + cnt[j], 0] < 0).all()
            and (cnt[link[j, 0]:link[j, 0] + cnt[j]] == 1).all()]
    assert len(flat) > 20
    np.testing.assert_allclose(rad[flat], exact[flat], rtol=2e-6, atol=4e-7 * 30)
    # the bound is not loose everywhere: the root's is within 2x of the truth
    assert rad[0] < 2 * exact[0]
    # planted slip: children's radii left out of the recursion
    slip = np.zeros_like(rad)
    for j in internal[::-1]:
        ch = np.arange(link[j, 0], link[j, 0] + cnt[j])
        d = com[ch, :3].astype(np.float64) - com[j, :3].astype(np.float64)
        slip[j] = np.sqrt((d * d).sum(1)).max()
    assert (slip < exact * (1 - 1e-3) - ulps).any()


@pytest.mark.parametrize("theta", [0.3, 0.6, 1.0])
def test_radius_rule_partitions_bodies_and_passes_the_test(tree, theta):
    """Mass-radius opening rule: the accepted nodes still partition the
    bodies, every accepted cell passes rad^2 < theta^2 d^2 and every ancestor
    of an accepted node fails it; at equal theta it accepts no more leaves
    than the size rule (a radius never exceeds the cell's half diagonal by
    much, and small clumps are accepted sooner)."""
    t, nodes, com, link, ext = tree["t"], tree["nodes"], tree["com"], tree["link"], tree["ext"]
    ps = t["pos"]
    n = ps.shape[0]
    rad = mass_radii(com, link, ps)
    fewer = 0
    for lo, hi in _boxes(ps):
        lv, cl = accepted(com, link, ext, lo, hi, theta, rad)
        acc = np.concatenate([lv, cl])
        cover = np.zeros(n, np.int64)
        for a, k in zip(nodes["first"][acc], nodes["count"][acc]):
            cover[a:a + k] += 1
        assert (cover == 1).all()

        def crit(j):
            c = com[j, :3].astype(np.float64)
            d = np.maximum(np.maximum(lo - c, c - hi), 0.0)
            return float(rad[j]) ** 2, theta * theta * (d @ d)

        for j in cl:
            s2, rhs = crit(j)
            assert s2 <= rhs * (1 + 1e-5)
        for j in acc:
            p = int(nodes["parent"][j])
            if p >= 0 and j != 0:
                s2, rhs = crit(p)
                assert s2 >= rhs * (1 - 1e-5), (j, p)
        lv0, _ = accepted(com, link, ext, lo, hi, theta)
        fewer += lv.size < lv0.size
    assert fewer >= len(_boxes(ps)) // 2


def test_quadrupole_term_against_brute_force():
    rng = np.random.default_rng(2)
    b = np.c_[rng.random((300, 3)) * [1.0, 0.6, 0.3], rng.random(300) + 0.5]
    m = b[:, 3].sum()
    c = (b[:, :3] * b[:, 3:]).sum(0) / m
    d = b[:, :3] - c
    S = (b[:, 3, None, None] * d[:, :, None] * d[:, None, :]).sum(0)
    com = np.r_[c, m][None].astype(np.float64)
    quad = np.array([[S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2], np.trace(S), 0]])
    for dist in (10.0, 20.0):
        x = c + dist * np.array([0.48, -0.6, 0.64])
        r = b[:, :3] - x
        exact = (b[:, 3:4] * r / ((r * r).sum(1) ** 1.5)[:, None]).sum(0)
        q = quadrupole(x, com, quad, np.array([0]), 0.0)
        mono = monopole(x, com, np.array([0]), 0.0)
        eq = np.linalg.norm(q - exact) / np.linalg.norm(exact)
        em = np.linalg.norm(mono - exact) / np.linalg.norm(exact)
        assert eq < 2.0 / dist ** 3 and em > 0.05 / dist ** 2 and eq < 0.2 * em
    # the network's features are the oracle's
    _, O = mlp.moments(b, c)
    oct_ = np.array([[O[0, 0, 0], O[1, 1, 1], O[2, 2, 2], O[0, 0, 1], O[0, 0, 2], O[0, 1, 1],
                      O[1, 1, 2], O[0, 2, 2], O[1, 2, 2], O[0, 1, 2], 0, 0]])
    link = np.array([[5, (3 << 5) | 2]], np.int32)
    X, k = features(x, com, oct_, np.array([0]), 4.0, link)
    want = mlp.features(x, c, m, 1.0, O)
    assert np.allclose(X[0], want, rtol=1e-12, atol=1e-15) and abs(k[0] - m) < 1e-12


def test_kernel_derivatives_are_the_gradients_of_each_other():
    """kernel_derivative(n + 1) is the gradient of kernel_derivative(n)
    (central differences), and n = 1 is -y / rho^3."""
    rng = np.random.default_rng(9)
    y = rng.normal(size=(4, 3))
    eps2 = 0.02
    r2 = (y * y).sum(1) + eps2
    assert np.allclose(kernel_derivative(y, eps2, 1), -y / r2[:, None] ** 1.5, rtol=1e-13)
    h = 1e-5
    for n in range(1, 5):
        up = kernel_derivative(y, eps2, n + 1)
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            fd = (kernel_derivative(y + e, eps2, n) - kernel_derivative(y - e, eps2, n)) / (2 * h)
            assert np.allclose(fd, up[..., j], rtol=1e-6, atol=1e-8), (n, j)


def _one_cell():
    rng = np.random.default_rng(12)
    b = rng.normal(size=(60, 3)) * [0.1, 0.06, 0.03]
    m = rng.uniform(0.5, 1.5, 60)
    c = (m[:, None] * b).sum(0) / m.sum()
    y = b - c
    S = np.einsum("k,ki,kj->ij", m, y, y)
    com = np.array([[c[0], c[1], c[2], m.sum()]])
    quad = np.array([[S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2], np.trace(S), 0]])
    return com, quad


@pytest.mark.parametrize("part,order", [("monopole", 4), ("second moments", 3)])
def test_local_expansion_truncation_order(part, order):
    """The local expansion about g of a cell's field: its error against the
    field evaluated at g + t u falls as t^4 for the monopole (kept to third
    order) and t^3 for the second-moment term (kept to second order), so
    every kept term is right; the errors shrink 2^4 and 2^3 fold per halving."""
    com, quad = _one_cell()
    g = com[0, :3] + np.array([1.1, -0.6, 0.8])
    u = np.array([0.3, -0.5, 0.8])
    eps2 = 1e-4
    idx = np.array([0])
    zero = 0.0 * quad
    Cm = local_coefficients(g, com, zero, idx, eps2)
    Cq = local_coefficients(g, com, quad, idx, eps2)
    errs = []
    for t in (0.08, 0.04, 0.02):
        x = g + t * u
        if part == "monopole":
            got = local_eval(Cm, (x - g)[None])[0]
            want = quadrupole(x, com, zero, idx, eps2)
        else:
            got = local_eval(Cq, (x - g)[None])[0] - local_eval(Cm, (x - g)[None])[0]
            want = quadrupole(x, com, quad, idx, eps2) - quadrupole(x, com, zero, idx, eps2)
        errs.append(np.linalg.norm(got - want))
    for a, b in zip(errs, errs[1:]):
        assert 0.85 * 2 ** order < a / b < 1.15 * 2 ** order, errs
    # at delta = 0 the expansion is the field itself
    assert np.allclose(local_eval(Cq, np.zeros((1, 3)))[0],
                       quadrupole(g, com, quad, idx, eps2), rtol=1e-12)


def test_far_field_walk_stays_near_the_exact_set_sum(tree):
    """walk_fp64 with far cells through the expansion differs from the exact
    sum of the same interaction set by the truncation only: below 1e-3 of the
    force at h / R < 0.3, and the far split follows its definition."""
    t, com, link, ext = tree["t"], tree["com"], tree["link"], tree["ext"]
    ps = t["pos"]
    n = ps.shape[0]
    nodes = tree["nodes"]
    # second moments about each node's centre of mass, from its bodies
    quad = np.zeros((com.shape[0], 8))
    for j in range(com.shape[0]):
        a, k = nodes["first"][j], nodes["count"][j]
        b = ps[a:a + k].astype(np.float64)
        d = b[:, :3] - com[j, :3].astype(np.float64)
        S = np.einsum("k,ki,kj->ij", b[:, 3], d, d)
        quad[j] = [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2], np.trace(S), 0]
    worst, fars = 0.0, 0
    for g0 in range(0, n - 32, 640):
        lo, hi = ps[g0:g0 + 32, :3].min(0), ps[g0:g0 + 32, :3].max(0)
        _, cells = accepted(com, link, ext, lo, hi, 0.5)
        fm, gc = far_split(cells, com, lo, hi, 0.3)
        h = 0.5 * np.linalg.norm(hi.astype(np.float64) - lo)
        R = np.linalg.norm(com[cells, :3].astype(np.float64) - gc, axis=1)
        clear = np.abs(h / R - 0.3) > 1e-5
        assert np.array_equal(fm[clear], (h / R < 0.3)[clear])
        fars += int(fm.sum())
        for i in (g0, g0 + 31):
            x = ps[i, :3].astype(np.float64)
            a0, _ = walk_fp64(x, (lo, hi), com, link, quad, ext, 0.5, 1e-4, 1)
            a1, _ = walk_fp64(x, (lo, hi), com, link, quad, ext, 0.5, 1e-4, 1, far=0.3)
            worst = max(worst, np.linalg.norm(a1 - a0) / np.linalg.norm(a0))
    assert fars > 0 and worst < 1e-3, (fars, worst)


def test_group_slots_partition_and_split_straddling_runs():
    """Slots cover every body once, in order; a run whose two halves sit in
    distant clusters is split at the jump between them; a compact run is not."""
    rng = np.random.default_rng(6)
    n = 64 * 3 + 17
    pos = np.zeros((n, 4), np.float32)
    # run 0: compact; run 1: 40 bodies near (0,0,0) then 24 near (9,9,9); run 2, tail: compact
    pos[:64, :3] = rng.random((64, 3)) * 0.1
    pos[64:104, :3] = rng.random((40, 3)) * 0.1 + 1.0
    pos[104:128, :3] = rng.random((24, 3)) * 0.1 + 9.0
    pos[128:, :3] = rng.random((n - 128, 3)) * 0.1 + 9.5
    keys = np.arange(n, dtype=np.uint64) * np.uint64(3)
    keys[104:] += np.uint64(1) << np.uint64(60)       # the largest jump: between 103 and 104
    slots, slot_of = group_slots(keys, pos, 0.3)
    assert slots.shape == (8, 2)
    covered = np.zeros(n, np.int64)
    for a, c in slots:
        covered[a:a + c] += 1
    assert (covered == 1).all()
    assert list(slots[2]) == [64, 40] and list(slots[3]) == [104, 24]
    assert slots[1, 1] == 0 and slots[0, 1] == 64
    assert slot_of[103] == 2 and slot_of[104] == 3 and slot_of[0] == 0
    assert slots[7, 1] == 0 and slots[6, 1] == n - 192
    whole, _ = group_slots(keys, pos, 0.0)
    assert (whole[1::2, 1] == 0).all()
    # the walk's default: every run with distinct keys splits at its largest jump
    assert split_ratio(0.2) > 1 and split_ratio(0.7) > 1
    always, _ = group_slots(keys, pos, split_ratio(0.5))
    assert list(always[0]) != [0, 64] and list(always[2]) == [64, 40]
    assert (always[1::2, 1] > 0).all()
