"""GPU tests of the learned corrector: the BF16-MFMA kernels against the CPU
oracle (oracle/mlp.py, pinned by finite differences), the sample generator
against direct sums, and what the trained network gains over the walk without
it."""
import numpy as np
import pytest

from cudagravitydemo_amd import Simulation, ics
from cudagravitydemo_amd import corrector as cr
from oracle import mlp
from tests.walk_check import accepted, group_slots, split_ratio, world_extent

pytestmark = pytest.mark.gpu


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, 8)).astype(np.float32)
    X[:, 3] = rng.random(B)
    Y = (rng.standard_normal((B, 3)) * 0.3).astype(np.float32)
    return X, Y


def test_forward_matches_oracle():
    p = mlp.init_params(4)
    net = cr.Corrector(p)
    try:
        X, _ = _batch(1000, 1)
        got = net.forward(X)
        want, _ = mlp.forward(p, X, bf16_ops=True)
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"forward max error {err:.2e}", flush=True)
        assert err < 5e-3
        assert np.array_equal(net.params(), p)
    finally:
        net.close()


def test_gradient_matches_oracle():
    p = mlp.init_params(6)
    net = cr.Corrector(p)
    try:
        for B in (16, 1000, 4099):
            X, Y = _batch(B, B)
            loss, g = net.gradient(X, Y)
            loss_ref, g_ref = mlp.backward(p, X, Y, bf16_ops=True)
            rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
            print(f"B={B} loss {loss:.6e} vs {loss_ref:.6e}; gradient rel err {rel:.2e}",
                  flush=True)
            assert abs(loss - loss_ref) <= 1e-3 * loss_ref
            assert rel < 1e-2
            # every layer's block is right, not just the total
            k = 0
            for o, i in mlp.SIZES:
                for n_ in (o * i, o):
                    a, b = g[k:k + n_], g_ref[k:k + n_]
                    assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-12, (k, n_)
                    k += n_
    finally:
        net.close()


def test_adam_step_matches_closed_form():
    p = mlp.init_params(8)
    net = cr.Corrector(p)
    lib = net.lib
    try:
        X, Y = _batch(2048, 3)
        _, g = net.gradient(X, Y)
        from cudagravitydemo_amd.engine import to_device
        xb, yb = to_device(lib, X), to_device(lib, Y)
        ob = to_device(lib, np.arange(2048, dtype=np.int32))
        lib.gd_mlp_train_step(net.ptr, xb.ptr, yb.ptr, ob.ptr, 2048, 1e-3)
        want, _, _ = mlp.adam(p, np.zeros(p.size), np.zeros(p.size), g.astype(np.float64), 1,
                              lr=1e-3)
        got = net.params()
        assert np.abs(got - want).max() < 1e-6
        for b in (xb, yb, ob):
            b.close()
    finally:
        net.close()


def test_canonical_sample_order_is_reproducible():
    """The sample kernel claims output rows with an atomic counter; after
    canonical_order() two fills from the same snapshot hold the same rows in
    the same order (so the seeded shuffle trains the same network), and the
    rows are a permutation of what the kernel wrote."""
    pos, vel = ics.exponential_disk(8192, seed=13)
    with Simulation(pos, vel, eps=0.01) as sim:
        sim.set_quadrupole(2)
        sim.build_tree_async()
        targets = ics.subsample(8192, 256, seed=3)
        rows = []
        for _ in range(2):
            s = cr.SampleSet(sim.lib, 400000)
            try:
                n, dropped = s.add_from(sim, targets, 0.8, 0.5)
                assert n > 1000 and dropped == 0
                X0, Y0 = s.host()
                s.canonical_order()
                X, Y = s.host()
                # the same multiset of rows
                a = np.concatenate([X0, Y0], 1)
                b = np.concatenate([X, Y], 1)
                assert np.array_equal(a[np.lexsort(a.T[::-1])], b[np.lexsort(b.T[::-1])])
                rows.append((X, Y))
            finally:
                s.close()
        assert np.array_equal(rows[0][0], rows[1][0])
        assert np.array_equal(rows[0][1], rows[1][1])


@pytest.mark.parametrize("ratio,min_ratio,rule", [(-1.0, 0.5, "size"), (0.0, 0.0, "size"),
                                                  (-1.0, 0.4, "radius")])
def test_samples_match_direct_sums(ratio, min_ratio, rule):
    """ratio -1: the learned walk's default groups at the sampling angle 0.8
    (every run split at its largest key jump), cells with s/r > 0.5 only; 0: whole
    64-body runs, as gd_walk_set_group_split(0) makes the walk keep them, and
    every accepted cell.  rule "radius": the mass-radius opening rule, whose
    cell measure s (features, target scale) is the mass radius."""
    pos, vel = ics.exponential_disk(8192, seed=13)
    with Simulation(pos, vel, eps=0.01) as sim:
        sim.set_quadrupole(2)
        sim.build_tree_async()
        targets = ics.subsample(8192, 64, seed=2)
        s = cr.SampleSet(sim.lib, 200000)
        prev = sim.lib.gd_walk_set_group_split(ratio)
        prev_rule = sim.set_opening_rule(rule)
        try:
            n, dropped = s.add_from(sim, targets, 0.8, min_ratio)
            assert n > 64 and dropped == 0
            X, Y = s.host()
        finally:
            sim.lib.gd_walk_set_group_split(prev)
            sim.set_opening_rule(prev_rule)
            s.close()
        assert (X[:, 3] > min_ratio * (1 - 1e-6)).all()
        if min_ratio == 0:
            assert (X[:, 3] < 0.5).mean() > 0.3   # the far cells are there too
        # recompute a few rows from the tree with the oracle's definitions
        total = sim.tree_summary()[0]
        com, link, quad = sim.walk_arrays(total)
        oct_ = sim.third_moments(total)
        rad = sim.mass_radii(total)
        from oracle.binding import NODE_DTYPE
        nodes = sim.nodes_raw(total).view(NODE_DTYPE).reshape(-1)
        ps = sim.positions()
        ext = float(np.max(ps[:, :3].max(0) - ps[:, :3].min(0)))
        found = 0
        rng = np.random.default_rng(0)
        cand = np.flatnonzero(link[:, 0] >= 0)
        slots, slot_of = group_slots(sim.keys_sorted(), ps,
                                     split_ratio(0.8) if ratio < 0 else ratio)
        assert (slots[1::2, 1] > 0).any() == (ratio < 0)
        for row in rng.choice(n, 40, replace=False):
            # find the (target, cell) of this row among all internal cells; targets in
            # one 64-body group share the features, and a node and its only child share
            # the centre of mass, so keep every candidate that matches all features
            cands = []
            for t in targets:
                x = ps[t, :3].astype(np.float64)
                # features are taken at the centre of the target's group
                a0, c0 = slots[slot_of[t]]
                blk = ps[a0:a0 + c0, :3]
                g = (np.float32(0.5) * (blk.min(0) + blk.max(0))).astype(np.float64)
                d = com[cand, :3] - g
                u = d / np.linalg.norm(d, axis=1)[:, None]
                for idx in cand[np.all(np.abs(u - X[row, :3]) < 1e-6, axis=1)]:
                    size = ext * 2.0 ** -int(nodes[idx]["level"])
                    if rule == "radius":
                        size = float(rad[idx])
                    S = np.array([[quad[idx, 0], quad[idx, 3], quad[idx, 4]],
                                  [quad[idx, 3], quad[idx, 1], quad[idx, 5]],
                                  [quad[idx, 4], quad[idx, 5], quad[idx, 2]]])
                    O = mlp.third_tensor(oct_[idx])
                    f = mlp.features(g, com[idx, :3], com[idx, 3], size, O)
                    if np.abs(f - X[row]).max() < 1e-4 * max(1, np.abs(f).max()):
                        cands.append((x, idx, size, S))
            assert cands, row
            ok = False
            for x, idx, size, S in cands:
                a, c = int(nodes[idx]["first"]), int(nodes[idx]["count"])
                y = mlp.residual_target(x, ps[a:a + c], com[idx, :3], com[idx, 3], size, S,
                                        0.01 ** 2)
                ok = ok or np.abs(y - Y[row]).max() < 1e-3 * max(1.0, np.abs(y).max())
            assert ok, row
            found += 1
        assert found >= 30


@pytest.mark.parametrize("rule", ["size", "radius"])
def test_sample_count_follows_the_opening_rule(rule):
    """The sampler's traversal opens cells by the walks' rule: under either
    rule the number of rows equals, target by target, the accepted internal
    cells of the target's group (tests/walk_check.accepted under the same rule)
    with s^2 > minRatio^2 |c - g|^2 from the group centre g, s the rule's cell
    measure (size, or mass radius), restated in float32 -- and the two rules
    give different sets."""
    pos, vel = ics.exponential_disk(8192, seed=13)
    theta, min_ratio = 0.8, 0.5
    f32 = np.float32
    with Simulation(pos, vel, eps=0.01) as sim:
        sim.set_quadrupole(2)
        sim.build_tree_async()
        targets = ics.subsample(8192, 64, seed=2)
        total = sim.tree_summary()[0]
        com, link, _ = sim.walk_arrays(total)
        rad = sim.mass_radii(total)
        ps = sim.positions()
        ext = world_extent(*sim.bbox())
        slots, slot_of = group_slots(sim.keys_sorted(), ps, split_ratio(theta))
        want = {}
        for r in ("size", "radius"):
            cnt = 0
            for t in targets:
                a0, c0 = slots[slot_of[t]]
                blk = ps[a0:a0 + c0, :3].astype(f32)
                lo, hi = blk.min(0), blk.max(0)
                _, cells = accepted(com, link, ext, lo, hi, theta, rad if r == "radius" else None)
                g = (lo + hi) * f32(0.5)
                size = np.ldexp(f32(ext), -(link[cells, 1] & 31)).astype(f32)
                if r == "radius":
                    size = rad[cells].astype(f32)
                d = com[cells, :3].astype(f32) - g
                d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
                mr2 = f32(min_ratio) * f32(min_ratio)
                cnt += int((size * size > mr2 * d2).sum())
            want[r] = cnt
        assert want["size"] != want["radius"]
        s = cr.SampleSet(sim.lib, 400000)
        prev = sim.set_opening_rule(rule)
        try:
            n, dropped = s.add_from(sim, targets, theta, min_ratio)
        finally:
            sim.set_opening_rule(prev)
            s.close()
        assert dropped == 0
        assert n == want[rule], (n, want)


def test_trained_corrector_gains_over_quadrupole_walk():
    """Train on snapshots of a disk, then compare at theta 0.8 against FP64:
    the learned walk (quadrupole + learned residual) beats the monopole walk by
    far and the quadrupole walk it refines, in the median and in the tail."""
    pos, vel = ics.exponential_disk(1 << 16, seed=21)
    n = pos.shape[0]
    eps = 0.01
    with Simulation(pos, vel, eps=eps) as sim:
        sim.set_quadrupole(2)
        s = cr.SampleSet(sim.lib, 4_000_000)
        net = cr.Corrector(seed=1)
        try:
            for snap in range(3):
                sim.build_tree_async()
                got, _ = s.add_from(sim, cr.default_targets(n, 2048, seed=snap), 0.9)
                print(f"snapshot {snap}: {got} samples", flush=True)
                for _ in range(20):
                    sim.step(1e-3, 0.5)
            losses = cr.train(net, s, epochs=6)
            print("losses", ["%.3e" % v for v in losses], flush=True)
            assert losses[-1] < 0.9 * losses[0]
            sim.build_tree_async()
            sample = ics.subsample(n, 2048, seed=77)
            exact = sim.direct_sum_fp64(sample)

            def err():
                a = sim.accelerations()[sample]
                e = np.linalg.norm(a - exact, axis=1) / np.linalg.norm(exact, axis=1)
                return float(np.median(e)), float(np.quantile(e, 0.99))

            sim.set_quadrupole(0)   # moments of this tree stay; the walk is monopole
            sim.walk(0.8)
            plain = err()
            sim.set_quadrupole(2)
            sim.walk(0.8)
            quad = err()
            cr.walk_learned(sim, net, 0.8)
            learned = err()
            print(f"theta 0.8 median / p99 error: plain {plain[0]:.3e} / {plain[1]:.3e}, "
                  f"quadrupole {quad[0]:.3e} / {quad[1]:.3e}, "
                  f"learned {learned[0]:.3e} / {learned[1]:.3e}", flush=True)
            assert learned[0] < 0.3 * plain[0]
            assert learned[0] < 0.95 * quad[0]
            assert learned[1] < 0.9 * quad[1]
        finally:
            net.close()
            s.close()
