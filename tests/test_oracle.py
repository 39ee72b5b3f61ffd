"""Pins the CPU oracle against checks independent of its code (CPU only).

The reference ships no tests, fixtures or golden vectors, so each oracle
function is pinned by brute force, closed forms, known answers and
invariants; tests/golden/ holds fixtures written by the pinned oracle.
"""
import numpy as np
import pytest

from cudagravitydemo_amd import ics
from oracle import binding as ob
from tests import vote_check as vc
from tests.octree_check import check_invariants, reference_octree


def naive_morton(qx, qy, qz):
    """Per-bit interleave: bit k of x -> 3k+2, of y -> 3k+1, of z -> 3k."""
    out = np.zeros(len(qx), dtype=np.uint64)
    for k in range(21):
        for q, off in ((qx, 2), (qy, 1), (qz, 0)):
            bit = (np.asarray(q, np.uint64) >> np.uint64(k)) & np.uint64(1)
            out |= bit << np.uint64(3 * k + off)
    return out


def test_bbox_matches_numpy():
    pos, _ = ics.exponential_disk(5000, seed=3)
    lo, hi = ob.bbox(pos)
    assert np.array_equal(lo, pos[:, :3].min(0))
    assert np.array_equal(hi, pos[:, :3].max(0))


def test_world_box_known_answers():
    b = ob.world_box(np.array([0, 0, 0], np.float32), np.array([2, 1, 0.5], np.float32))
    assert (b.minX, b.minY, b.minZ, b.extent) == (0.0, 0.0, 0.0, 2.0)
    assert b.scale == np.float32(2097151.0) / np.float32(2.0)
    b = ob.world_box(np.array([-1, -3, 0], np.float32), np.array([1, 1, 7], np.float32))
    assert b.extent == 7.0  # z is the widest axis
    assert b.scale == np.float32(np.float32(2097151.0) / np.float32(7.0))


def test_morton_known_answers():
    box = ob.make_box(0.0, 0.0, 0.0, 1.0, 1.0)  # q = (u32) x
    pos = np.array([[1.5, 0, 0, 1], [0, 1.2, 0, 1], [0, 0, 1.9, 1],
                    [2097151, 2097151, 2097151, 1], [0, 0, 0, 1]], np.float32)
    keys, idx = ob.morton(pos, box)
    assert list(keys) == [4, 2, 1, (1 << 63) - 1, 0]
    assert list(idx) == [0, 1, 2, 3, 4]


def test_morton_matches_naive_interleave():
    pos, _ = ics.exponential_disk(20000, seed=11)
    lo, hi = ob.bbox(pos)
    box = ob.world_box(lo, hi)
    keys, _ = ob.morton(pos, box)
    f = np.float32
    q = [((pos[:, a] - f(m)) * f(box.scale)).astype(np.uint32)
         for a, m in zip(range(3), (box.minX, box.minY, box.minZ))]
    assert max(int(x.max()) for x in q) <= (1 << 21) - 1
    assert np.array_equal(keys, naive_morton(*q))
    # the widest axis reaches the top of the grid (2^21-1 or 2^21-2)
    assert max(int(x.max()) for x in q) >= (1 << 21) - 2


def test_morton_degenerate_box_gives_zero_keys():
    # all bodies coincide: extent 0, scale +inf, (x-min)*scale = NaN -> 0
    pos = np.tile(np.array([[1.0, 2.0, 3.0, 0.5]], np.float32), (7, 1))
    lo, hi = ob.bbox(pos)
    box = ob.world_box(lo, hi)
    assert box.extent == 0.0 and np.isinf(box.scale)
    keys, _ = ob.morton(pos, box)
    assert np.all(keys == 0)


@pytest.mark.parametrize("n", [1, 2, 255, 4097, 50000])
def test_sort_is_stable_sort(n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    keys[rng.integers(0, n, n // 3)] = keys[0]  # many ties
    vals = np.arange(n, dtype=np.uint32)
    ks, vs = ob.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ks, keys[order])
    assert np.array_equal(vs, vals[order])


def test_reorder_gathers():
    pos, vel = ics.exponential_disk(1000, seed=2)
    perm = np.random.default_rng(0).permutation(1000).astype(np.uint32)
    p2, v2 = ob.reorder(pos, vel, perm)
    assert np.array_equal(p2, pos[perm]) and np.array_equal(v2, vel[perm])


def _sorted_keys(pos):
    lo, hi = ob.bbox(pos)
    keys, idx = ob.morton(pos, ob.world_box(lo, hi))
    return ob.sort_pairs(keys, idx)[0]


def _compare_topology(keys):
    rc, nodes, lo, lc, d, total = ob.build_topology(keys)
    assert rc == 0
    ref = reference_octree(keys)
    assert total == len(ref)
    got = nodes[:total]
    for f in ("path", "parent", "first", "count", "child", "level"):
        assert np.array_equal(got[f], ref[f]), f
    check_invariants(got, len(keys))
    # level tables: contiguous levels, counts match
    for L in range(d + 1):
        sel = ref["level"] == L
        assert lc[L] == sel.sum()
        assert lo[L] == np.flatnonzero(sel)[0]
    if d < 21:
        assert lc[d + 1] == 0 and lo[d + 1] == total
    return got


@pytest.mark.parametrize("seed", [0, 1])
def test_topology_matches_independent_octree(seed):
    pos, _ = ics.exponential_disk(3000, seed=seed)
    _compare_topology(_sorted_keys(pos))


def test_topology_duplicates_and_edge_sizes():
    # duplicate positions make multi-body leaves at level 21
    pos, _ = ics.uniform_cube(600, seed=4)
    pos[100:110] = pos[100]
    pos[300:303] = pos[7]
    nodes = _compare_topology(_sorted_keys(pos))
    assert (nodes["count"][nodes["level"] == 21] > 1).any()
    for n in (1, 2, 3, 33):
        p, _ = ics.uniform_cube(n, seed=n)
        _compare_topology(_sorted_keys(p))
    # all bodies coincident: one chain down to a level-21 leaf
    p = np.tile(np.array([[0.5, 0.5, 0.5, 1.0]], np.float32), (5, 1))
    keys = np.zeros(5, np.uint64)
    nodes = _compare_topology(keys)
    assert len(nodes) == 22 and nodes[-1]["count"] == 5


def overflow_fixture():
    """512 lattice bodies each with a twin ~2 grid cells away along x: every
    pair shares its cells down to level 20, so the tree needs 9673 nodes,
    more than the 8 N = 8192 the reference allocates (basicGravity.cu:11)."""
    g = np.arange(8, dtype=np.float32)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel(), np.ones(512, np.float32)], 1)
    twin = p.copy()
    twin[:, 0] += np.float32(7.0 * 2.0 ** -20)
    return np.ascontiguousarray(np.concatenate([p, twin]).astype(np.float32))


def test_topology_overflow_at_eight_n():
    pos = overflow_fixture()
    keys = _sorted_keys(pos)
    assert ob.build_topology(keys)[0] == -1
    rc, _, _, _, _, total = ob.build_topology(keys, max_nodes=20 * len(pos))
    assert rc == 0 and total == 9673 > 8 * len(pos)


def test_topology_overflow_reported():
    pos, _ = ics.uniform_cube(256, seed=9)
    rc = ob.build_topology(_sorted_keys(pos), max_nodes=100)[0]
    assert rc == -1


def test_tree_mass_matches_fp64_com():
    pos, vel = ics.exponential_disk(4000, seed=8)
    t = ob.build_tree(pos, vel)
    nodes, ps = t["nodes"], t["pos"].astype(np.float64)
    assert t["zero_mass"] == 0
    m = ps[:, 3]
    assert abs(nodes[0]["mass"][3] - m.sum()) <= 1e-6 * m.sum()
    for i in range(0, len(nodes), 7):
        nd = nodes[i]
        a, c = int(nd["first"]), int(nd["count"])
        mm = m[a:a + c].sum()
        com = (ps[a:a + c, :3] * m[a:a + c, None]).sum(0) / mm
        assert abs(nd["mass"][3] - mm) <= 1e-5 * mm
        scale = np.abs(ps[a:a + c, :3]).max() + 1e-3
        assert np.abs(nd["mass"][:3] - com).max() <= 2e-5 * scale, i
    # every level's masses add up to the total mass of the bodies it holds
    for L in range(1, t["deepest"] + 1):
        sel = nodes["level"] == L
        held = nodes["count"][sel].sum()
        assert held <= len(ps)


def _softened_direct_fp64(targets, sources, eps2, G):
    """FP64 softened sum over sources (self terms give 0), and the sum of |terms|."""
    p = targets.astype(np.float64)
    q = sources.astype(np.float64)
    d = p[:, None, :3] - q[None, :, :3]
    r2 = (d * d).sum(-1) + eps2
    f = -G * q[None, :, 3] / (r2 * np.sqrt(r2))
    terms = f[..., None] * d
    return terms.sum(1), np.linalg.norm(terms, axis=2).sum(1)


def test_walk_theta_zero_is_direct_sum():
    # theta = 0 never accepts an internal node: every leaf acts as its own
    # monopole, so the walk is the softened direct sum over the leaves'
    # (centre of mass, mass), summed in FP32 in traversal order
    pos, vel = ics.uniform_cube(700, seed=12)
    t = ob.build_tree(pos, vel)
    eps2 = np.float32(0.01 ** 2)
    acc, inter, opened = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], eps2, 1.0, 0.0)
    nodes = t["nodes"]
    leaf = (nodes["child"] == -1).all(1)
    assert nodes["count"][leaf].max() == 1
    sources = nodes["mass"][leaf]
    exact, mag = _softened_direct_fp64(t["pos"], sources, float(eps2), 1.0)
    # FP32 summation of 700 terms in another order: a few eps of sum|terms|
    rel = np.linalg.norm(acc - exact, axis=1) / mag
    assert rel.max() < 5e-6
    # against the bodies themselves the leaf centre (x m)/m is off by <= 1 ulp
    exact_b, mag_b = _softened_direct_fp64(t["pos"], t["pos"], float(eps2), 1.0)
    assert (np.linalg.norm(acc - exact_b, axis=1) / mag_b).max() < 1e-4
    assert inter == 700 * 700


def test_walk_error_shrinks_with_theta_and_respects_groups():
    pos, vel = ics.exponential_disk(4096, seed=21)
    t = ob.build_tree(pos, vel)
    eps = 0.01
    sample = ics.subsample(4096, 300, seed=1)
    exact = ob.direct_sum_fp64(t["pos"], sample, eps * eps)
    errs = []
    for th in (0.9, 0.5, 0.25):
        acc, _, _ = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"],
                            np.float32(eps * eps), 1.0, th)
        e = np.linalg.norm(acc[sample] - exact, axis=1) / np.linalg.norm(exact, axis=1)
        errs.append(np.median(e))
    assert errs[0] > errs[1] > errs[2]
    assert errs[1] < 1e-2
    # a group walks only its own bodies: walking groups [3, 5) writes only them
    acc, _, _ = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], np.float32(eps * eps),
                        1.0, 0.5, g0=3, g1=5)
    nz = np.flatnonzero(np.abs(acc).sum(1))
    assert nz.min() >= 96 and nz.max() < 160


def test_walk_planted_mistake_is_caught():
    # an oracle walk with a wrong sign of G must fail the FP64 check used above
    pos, vel = ics.exponential_disk(2048, seed=5)
    t = ob.build_tree(pos, vel)
    sample = ics.subsample(2048, 100, seed=2)
    exact = ob.direct_sum_fp64(t["pos"], sample, 1e-4)
    acc, _, _ = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], np.float32(1e-4), -1.0,
                        0.5)
    e = np.linalg.norm(acc[sample] - exact, axis=1) / np.linalg.norm(exact, axis=1)
    assert np.median(e) > 1.0


# ---- the reference walk's acceptance rule, pinned independently --------------
# tests/vote_check.py enumerates each 32-body group's accepted set from the
# definition (tree.cu:562-676) level by level; the oracle records the nodes
# each body accepted while replaying the stack.  The fixtures are chosen so
# that every plausible misreading of the rule changes some body's set.


def vote_fixture():
    """Lattice fixture with exact ties of the opening test, x not the widest
    axis (x 32, y 45, z 15), 3 valid lanes in the last group and two stacks
    of coincident bodies; plus a far body at x = 32 that makes the x extent a
    power of two (node sizes 32 / 2^level)."""
    g = np.arange(16, dtype=np.float32)
    x, y, z = np.meshgrid(g, 3 * g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel(), np.ones(4096, np.float32)], 1)
    rng = np.random.default_rng(7)
    p = p[np.sort(rng.choice(4096, 4096 - 13, replace=False))]
    extra = np.concatenate([np.repeat(p[100:101], 10, 0), np.repeat(p[2000:2001], 5, 0),
                            np.array([[32, 0, 0, 1]], np.float32)])
    return np.ascontiguousarray(np.concatenate([p, extra]).astype(np.float32))


def disk_fixture():
    """A 3001-body disk stretched 3x along y (x is not the widest axis), 25
    valid lanes in the last group, 7 coincident bodies."""
    p, v = ics.exponential_disk(3001, seed=21)
    p[:, 1] *= 3.0
    p[500:507] = p[500]
    return p, v


def test_fmaf32_rounds_once():
    # against exact rational arithmetic, including heavy cancellation
    from fractions import Fraction
    rng = np.random.default_rng(0)
    a = rng.standard_normal(4000).astype(np.float32)
    b = rng.standard_normal(4000).astype(np.float32)
    c = (-(a.astype(np.float64) * b) * (1 + rng.standard_normal(4000) * 1e-7)).astype(np.float32)
    c[:2000] = rng.standard_normal(2000).astype(np.float32)
    got = vc.fmaf32(a, b, c)
    for i in range(4000):
        ex = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        f = np.float32(float(ex))
        cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
        errs = [abs(Fraction(float(x)) - ex) for x in cands]
        best = [x for x, e in zip(cands, errs) if e == min(errs)]
        assert got[i] in best, (i, a[i], b[i], c[i])
    # where rounding the float64 sum to float32 goes wrong: the exact value
    # 1 + 2^-23 + 2^-24 - 2^-64 lies just below a float32 tie; float64 rounds
    # it onto the tie and ties-to-even would then round up to 1 + 2^-22
    a = np.float32(2.0 ** -24 * (1 + 2.0 ** -20))
    b = np.float32(1 - 2.0 ** -20)
    c = np.float32(1 + 2.0 ** -23)
    assert np.float32(a.astype(np.float64) * b + np.float64(c)) == np.float32(1 + 2.0 ** -22)
    assert vc.fmaf32(a, b, c) == c


def _vote_case(pos, vel, theta):
    t = ob.build_tree(pos, vel)
    sets, opened, ties = vc.group_accepted_sets(t["pos"], t["nodes"], t["bmin"], t["bmax"], theta)
    return t, sets, opened, ties


@pytest.mark.parametrize("fixture,theta", [("lattice", 0.5), ("lattice", 1.0), ("disk", 0.5),
                                           ("disk", 0.7)])
def test_walk_accepted_sets_match_group_vote_definition(fixture, theta):
    if fixture == "lattice":
        pos = vote_fixture()
        vel = np.zeros((pos.shape[0], 3), np.float32)
    else:
        pos, vel = disk_fixture()
    n = pos.shape[0]
    t, sets, opened, ties = _vote_case(pos, vel, theta)
    if fixture == "lattice":
        assert ties > 0            # the strict '<' is exercised
    eps2 = np.float32(1e-4)
    acc, inter, op, lists = ob.walk_traced(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], eps2,
                                           1.0, theta)
    # the oracle's accepted nodes, body by body, equal the group's set
    assert vc.compare_with_oracle_lists(sets, lists, n) == []
    valid = [min(32, n - 32 * g) for g in range(len(sets))]
    assert inter == sum(len(s) * k for s, k in zip(sets, valid))
    assert op == opened
    # or_walk itself takes the same path as the traced replay
    acc0, inter0, op0 = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], eps2, 1.0, theta)
    assert np.array_equal(acc0, acc) and (inter0, op0) == (inter, op)
    # every group's accepted nodes tile the bodies
    assert vc.check_partition(sets, t["nodes"], n) == []
    # forces: the oracle's FP32 sums against FP64 sums over the same sets
    a64, mag = vc.monopole_fp64(t["pos"], t["nodes"], sets, float(eps2))
    assert (np.linalg.norm(acc - a64, axis=1) / mag).max() < 2e-6


def test_walk_accepted_sets_satisfy_the_rule_node_by_node():
    # the rule restated per node, by parent pointers instead of levels:
    # an accepted node is a leaf or passes for every valid lane; each of its
    # proper ancestors below the root is not a leaf and fails for some lane
    pos, vel = disk_fixture()
    theta = 0.6
    t, sets, _, _ = _vote_case(pos, vel, theta)
    nodes = t["nodes"]
    n = pos.shape[0]
    ext = np.float64(t["bmax"][0]) - np.float64(t["bmin"][0])
    leaf = vc.leaf_flags(nodes)
    P = t["pos"][:, :3].astype(np.float64)
    for g in range(0, len(sets), 5):
        bodies = P[32 * g:min(n, 32 * g + 32)]
        for v in sets[g]:
            chain = [int(v)]
            while nodes["parent"][chain[-1]] > 0:
                chain.append(int(nodes["parent"][chain[-1]]))
            for k, u in enumerate(chain):
                size = ext / 2.0 ** int(nodes["level"][u])
                d2 = ((bodies - nodes["mass"][u, :3].astype(np.float64)) ** 2).sum(1)
                margin = size * size - theta * theta * d2     # < 0: the lane passes
                if k == 0:
                    assert leaf[u] or (margin < 1e-5 * size * size).all(), (g, u)
                else:
                    assert not leaf[u] and (margin > -1e-5 * size * size).any(), (g, u)


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_walk_check_rejects_planted_mistakes(variant):
    # 1 per-lane acceptance, 2 widest-axis node size, 3 '<=', 4 invalid lanes
    # vote, 5 leaf rule without its no-children clause: each planted in the
    # oracle, each must change some body's accepted set on a fixture
    caught = []
    for fixture, theta in (("lattice", 0.5), ("lattice", 1.0), ("disk", 0.5)):
        if fixture == "lattice":
            pos = vote_fixture()
            vel = np.zeros((pos.shape[0], 3), np.float32)
        else:
            pos, vel = disk_fixture()
        t, sets, _, _ = _vote_case(pos, vel, theta)
        _, _, _, lists = ob.walk_traced(t["pos"], t["nodes_full"], t["bmin"], t["bmax"],
                                        np.float32(1e-4), 1.0, theta, variant=variant)
        caught.append(len(vc.compare_with_oracle_lists(sets, lists, pos.shape[0])) > 0)
    assert caught[0] and caught[1], caught   # the lattice catches every variant


def test_integrate_kick_then_drift():
    rng = np.random.default_rng(3)
    pos = rng.standard_normal((500, 4)).astype(np.float32)
    vel = rng.standard_normal((500, 3)).astype(np.float32)
    acc = rng.standard_normal((500, 3)).astype(np.float32)
    dt = np.float32(1e-3)
    p2, v2 = ob.integrate(pos, vel, acc, dt)
    v_exp = vel.astype(np.float64) + acc.astype(np.float64) * float(dt)
    x_exp = pos[:, :3].astype(np.float64) + v2.astype(np.float64) * float(dt)
    assert np.allclose(v2, v_exp, rtol=0, atol=2e-7 * np.abs(v_exp).max())
    assert np.allclose(p2[:, :3], x_exp, rtol=0, atol=2e-7 * np.abs(x_exp).max())
    assert np.array_equal(p2[:, 3], pos[:, 3])


def test_direct_sum_closed_form_and_momentum():
    pos = np.array([[0, 0, 0, 2.0], [3, 4, 0, 1.0]], np.float32)
    a = ob.direct_sum_fp64(pos, np.array([0, 1], np.int32), 0.0, 1.0)
    # |a_0| = m_1 / r^2 = 1/25 toward body 1; |a_1| = 2/25 toward body 0
    assert np.allclose(a[0], np.array([3, 4, 0]) / 5 * (1 / 25), rtol=1e-12)
    assert np.allclose(a[1], -np.array([3, 4, 0]) / 5 * (2 / 25), rtol=1e-12)
    p, _ = ics.plummer(800, seed=1)
    a = ob.direct_sum_fp64(p, np.arange(800, dtype=np.int32), 0.05 ** 2)
    mom = (p[:, 3:4].astype(np.float64) * a).sum(0)
    assert np.abs(mom).max() < 1e-12 * np.abs(a).max()


def test_golden_fixtures_match_oracle():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "disk1024.npz")
    g = np.load(path)
    t = ob.build_tree(g["pos"], g["vel"])
    assert np.array_equal(t["keys"], g["keys"])
    assert np.array_equal(t["perm"], g["perm"])
    raw = t["nodes"].view(np.uint8).reshape(-1, 80)
    assert np.array_equal(raw[:, :56], g["nodes"][:, :56])
    assert np.array_equal(raw[:, 64:], g["nodes"][:, 64:])
    acc, _, _ = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], np.float32(1e-4), 1.0, 0.5)
    assert np.array_equal(acc, g["acc"])
