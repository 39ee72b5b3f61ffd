"""Training and use of the learned far-field corrector (host orchestration).

All arithmetic runs in libgravity.so: samples come from gd_corrector_samples
(exact per-cell sums on the GPU), training from gd_mlp_train_step (BF16 MFMA
forward/backward + Adam), and the corrected walk from gd_walk_learned.
The FP64 direct-sum judge decides whether the corrector may be used:
JudgedRun checks the walk every k steps and, when the measured median force
error exceeds the budget, switches the corrector off and continues with the
walk without it at a smaller opening angle.
"""
import ctypes as C

import numpy as np

from ._lib import WalkStats, check, load_library
from .engine import DeviceBuffer, from_device, to_device
from . import ics

N_PARAMS = 9091
SIZES = [(64, 8), (64, 64), (64, 64), (3, 64)]


def init_params(seed=0):
    """He-normal hidden weights, zero biases and a zero output layer (the
    untrained network predicts no correction), float32 in the library's
    layout."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for li, (o, i) in enumerate(SIZES):
        w = rng.standard_normal(o * i) * np.sqrt(2.0 / i)
        parts.append(w if li < len(SIZES) - 1 else np.zeros(o * i))
        parts.append(np.zeros(o))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == N_PARAMS
    return p


def scale_output(p, factor):
    """Parameters whose outputs are `factor` times those of p (last layer scaled)."""
    q = np.array(p, np.float32, copy=True)
    q[-3 - 3 * 64:] *= np.float32(factor)
    return q


class Corrector:
    """A network on the GPU plus its Adam state."""

    def __init__(self, params=None, seed=0):
        self.lib = load_library()
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        self.ptr = self.lib.gd_mlp_create(p.ctypes.data)
        check(self.lib)

    def params(self):
        out = np.empty(N_PARAMS, np.float32)
        self.lib.gd_mlp_get_params(self.ptr, out.ctypes.data)
        check(self.lib)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.lib.gd_mlp_set_params(self.ptr, p.ctypes.data)
        check(self.lib)

    def forward(self, X):
        """X: host (rows, 8) float32 -> host (rows, 3)."""
        X = np.ascontiguousarray(X, np.float32)
        xb = to_device(self.lib, X)
        yb = DeviceBuffer(self.lib, 12 * X.shape[0])
        try:
            self.lib.gd_mlp_forward(self.ptr, xb.ptr, yb.ptr, X.shape[0])
            check(self.lib)
            return from_device(self.lib, yb.ptr, np.float32, (X.shape[0], 3))
        finally:
            xb.close()
            yb.close()

    def gradient(self, X, Y):
        """(loss, gradient) of the mean squared error over all rows."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        B = X.shape[0]
        xb, yb = to_device(self.lib, X), to_device(self.lib, Y)
        ob = to_device(self.lib, np.arange(B, dtype=np.int32))
        g = np.empty(N_PARAMS, np.float32)
        try:
            loss = self.lib.gd_mlp_gradient(self.ptr, xb.ptr, yb.ptr, ob.ptr, B, g.ctypes.data)
            check(self.lib)
            return loss, g
        finally:
            for b in (xb, yb, ob):
                b.close()

    def close(self):
        if self.ptr:
            self.lib.gd_mlp_destroy(self.ptr)
            self.ptr = None


class SampleSet:
    """Device-resident training rows (X: n x 8, Y: n x 3)."""

    def __init__(self, lib, cap):
        self.lib = lib
        self.cap = int(cap)
        self.X = DeviceBuffer(lib, 32 * self.cap)
        self.Y = DeviceBuffer(lib, 12 * self.cap)
        self.n = 0

    def add_from(self, sim, targets, theta, min_ratio=0.55):
        """Append the (body, cell) samples of one snapshot: the cells whose
        size / distance from the group's centre exceeds min_ratio, the ones
        walk_learned corrects with the same min_ratio; returns (rows added,
        targets dropped).  Raises ValueError, adding nothing, when the rows do
        not fit the set's remaining room: the kernel would keep an arbitrary
        subset of them."""
        lib = self.lib
        t = to_device(lib, np.ascontiguousarray(targets, np.int32))
        dropped = C.c_int(0)
        room = self.cap - self.n
        try:
            got = lib.gd_corrector_samples(C.byref(sim.state), sim.pos, t.ptr, len(targets),
                                           theta, min_ratio, sim.eps * sim.eps,
                                           self.X.ptr + 32 * self.n,
                                           self.Y.ptr + 12 * self.n, room, C.byref(dropped))
            check(lib)
        finally:
            t.close()
        if got > room:
            raise ValueError(f"sample set full: the snapshot yields {got} rows, "
                             f"{room} of {self.cap} left")
        self.n += int(got)
        return int(got), dropped.value

    def canonical_order(self):
        """Put the rows in an order that depends only on their values.  The
        sample kernel claims output rows with an atomic counter, so the order
        of a snapshot's rows (and with it the minibatches of a seeded
        shuffle) varies from run to run; sorted by a hash of each row's bytes,
        the same rows train the same network every time."""
        if self.n <= 1:
            return
        X, Y = self.host()
        w = np.concatenate([X.view(np.uint32), Y.view(np.uint32)], 1).astype(np.uint64)
        h = np.zeros(self.n, np.uint64)
        with np.errstate(over="ignore"):
            for j in range(w.shape[1]):
                h = (h ^ w[:, j]) * np.uint64(0x100000001B3)
                h ^= h >> np.uint64(29)
        order = np.argsort(h, kind="stable")
        self.X.upload(np.ascontiguousarray(X[order]))
        self.Y.upload(np.ascontiguousarray(Y[order]))

    def host(self, limit=None):
        n = self.n if limit is None else min(limit, self.n)
        return (from_device(self.lib, self.X.ptr, np.float32, (n, 8)),
                from_device(self.lib, self.Y.ptr, np.float32, (n, 3)))

    def close(self):
        self.X.close()
        self.Y.close()


def train(net, samples, epochs=1, batch=8192, lr=2e-3, seed=0, log=None, decay=True):
    """Adam on shuffled minibatches of the sample set, with the learning
    rate cosine-decayed from lr to lr / 20 (decay) or constant.  The targets are trained at unit RMS (divided
    by their RMS over the set) and the factor is folded back into the output
    layer afterwards.  Returns per-epoch mean loss (in the unit-RMS scale)."""
    lib = net.lib
    rng = np.random.Generator(np.random.PCG64(seed))
    n = samples.n
    if n == 0:
        return []
    Y = from_device(lib, samples.Y.ptr, np.float32, (n, 3))
    rms = float(np.sqrt(np.mean(np.sum(Y.astype(np.float64) ** 2, axis=1))))
    rms = rms if rms > 0 else 1.0
    ys = DeviceBuffer(lib, 12 * n)
    ys.upload(np.ascontiguousarray(Y / np.float32(rms)))
    net.set_params(scale_output(net.params(), 1.0 / rms))
    losses = []
    order_buf = DeviceBuffer(lib, 4 * n)
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    step = 0
    try:
        for ep in range(epochs):
            order = rng.permutation(n).astype(np.int32)
            order_buf.upload(order)
            tot, cnt = 0.0, 0
            for b0 in range(0, n - batch + 1 if n >= batch else 1, batch):
                B = min(batch, n - b0)
                f = 0.5 * (1.0 + np.cos(np.pi * step / max(total, 1))) if decay else 1.0
                rate = lr * (0.05 + 0.95 * f)
                loss = lib.gd_mlp_train_step(net.ptr, samples.X.ptr, ys.ptr,
                                             order_buf.ptr + 4 * b0, B, rate)
                check(lib)
                tot += loss
                cnt += 1
                step += 1
            losses.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.4e}")
    finally:
        order_buf.close()
        ys.close()
        net.set_params(scale_output(net.params(), rms))
    return losses


def walk_learned(sim, net, theta, min_ratio=0.55, stats=False):
    """The learned walk on sim's current tree (built with gd_set_quadrupole
    level 2).  Cells with size / distance <= min_ratio (from the group's
    centre) keep the analytic quadrupole alone: their residual is small and
    the network is not evaluated for them."""
    st = WalkStats() if stats else None
    sim.lib.gd_walk_learned(C.byref(sim.state), sim.pos, sim.acc, None, sim.eps * sim.eps,
                            sim.G, theta, sim.n, net.ptr, min_ratio,
                            C.byref(st) if st is not None else None)
    check(sim.lib)
    if st is not None:
        return {"interactions": st.interactions, "opened": st.opened,
                "overflow_groups": st.overflowGroups, "cell_interactions": st.cellInteractions,
                "body_groups": st.bodyGroups, "network_evals": st.networkEvals,
                "far_cells": st.farCells,
                "far_interactions": st.farInteractions}
    return None


def median_error(sim, sample):
    """Median relative force error of sim.acc on sample vs the FP64 judge."""
    exact = sim.direct_sum_fp64(sample)
    acc = sim.accelerations()[sample]
    e = np.linalg.norm(acc - exact, axis=1) / np.linalg.norm(exact, axis=1)
    return float(np.median(e)), e


def default_targets(n, k=8192, seed=11):
    return ics.subsample(n, k, seed=seed)


def train_on_snapshots(sim, net, advance, snapshots=4, targets=8192, subsamples=1,
                       sample_theta=0.9, min_ratio=0.5875, epochs=10, batch=8192, lr=2e-3,
                       cap=24_000_000, seed=0, log=None, decay=True):
    """Train `net` on (features -> residual) samples from snapshots of the
    run in `sim`: at each snapshot the tree is built with second and third
    moments and `subsamples` seeded sets of `targets` bodies give their
    samples (gd_corrector_samples at opening angle sample_theta, cells with
    size / distance above min_ratio); `advance()` moves the run on to the
    next snapshot.  Returns a dict with the sample count, per-epoch losses and
    the seconds spent."""
    import time
    t0 = time.perf_counter()
    samples = SampleSet(sim.lib, cap)
    try:
        for snap in range(snapshots):
            sim.set_quadrupole(2)   # advance() may have changed the moment level
            sim.build_tree_async()
            for k in range(subsamples):
                got, _ = samples.add_from(
                    sim, default_targets(sim.n, targets, seed=seed + 1000 * snap + k),
                    sample_theta, min_ratio)
                if log:
                    log(f"snapshot {snap} subsample {k}: {got} samples")
            if snap + 1 < snapshots:
                advance()
        samples.canonical_order()
        n = samples.n
        losses = train(net, samples, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log,
                       decay=decay)
    finally:
        samples.close()
    return {"samples": n, "losses": losses, "seconds": time.perf_counter() - t0}


class JudgedRun:
    """The learned walk with its FP64 judge and automatic fallback.

    Each step is the reference's (basicGravity.cu:774-783): build the tree,
    walk, kick-drift.  Every `every`-th step, counting from the first, the
    walk's accelerations on a seeded sample of `sample_size` bodies are
    compared with the FP64 direct sum (gd_direct_sum_fp64, then
    gd_force_errors) before the kick-drift.  When the median relative error
    exceeds `budget` the corrector is switched off for the rest of the run:
    the step's walk is done again with the fallback walk (by default the
    quadrupole walk at theta 0.4, the corrector's own method without the
    network) and every later step uses it.  The fallback's accelerations are
    judged on the same step as well, so every judged step records the error
    of the accelerations that were integrated.

    checks: one dict per judged walk (step, walk, theta, median, p99, within).
    judge_ms: time spent in the judge (HIP events), for rates with and without it.
    """

    def __init__(self, sim, net, theta, dt, min_ratio=0.5875, budget=1e-3, every=20,
                 sample_size=4096, sample_seed=4242, fallback_theta=0.4, fallback="quad"):
        if fallback not in ("quad", "plain"):
            raise ValueError("fallback is 'quad' or 'plain'")
        self.sim = sim
        self.lib = sim.lib
        self.net = net
        self.theta = float(theta)
        self.dt = float(dt)
        self.min_ratio = float(min_ratio)
        self.budget = float(budget)
        self.every = int(every)
        self.fallback = fallback
        self.fallback_theta = float(fallback_theta)
        self.corrector_on = net is not None
        self.steps = 0
        self.checks = []
        self.judge_ms = 0.0
        self.sample = ics.subsample(sim.n, sample_size, seed=sample_seed)
        ns = self.sample.shape[0]
        self._sample = to_device(self.lib, self.sample)
        self._exact = DeviceBuffer(self.lib, 24 * ns)
        self._err = DeviceBuffer(self.lib, 8 * ns)
        self._ev = (self.lib.gd_event_create(), self.lib.gd_event_create())
        sim.set_quadrupole(2 if self.corrector_on else (1 if fallback == "quad" else 0))

    def _walk(self):
        if self.corrector_on:
            walk_learned(self.sim, self.net, self.theta, self.min_ratio)
        else:
            self.sim.walk(self.fallback_theta)

    def _judge(self, exact_ready):
        """Median / p99 error of sim.acc on the sample; computes the direct sum
        unless this step's is already in place."""
        lib, ns = self.lib, self.sample.shape[0]
        lib.gd_event_record(self._ev[0])
        if not exact_ready:
            lib.gd_direct_sum_fp64(self.sim.pos, self.sim.lo if self.sim.hilo else None,
                                   self.sim.n, self._sample.ptr, ns, self.sim.eps ** 2,
                                   self.sim.G, self._exact.ptr)
        lib.gd_force_errors(self.sim.acc, self._sample.ptr, ns, self._exact.ptr, self._err.ptr)
        lib.gd_event_record(self._ev[1])
        err = from_device(lib, self._err.ptr, np.float64, (ns,))
        self.judge_ms += lib.gd_event_elapsed_ms(self._ev[0], self._ev[1])
        check(lib)
        return float(np.median(err)), float(np.quantile(err, 0.99))

    def step(self):
        sim = self.sim
        sim.build_tree_async()
        self._walk()
        if self.steps % self.every == 0:
            med, p99 = self._judge(False)
            rec = {"step": self.steps, "walk": "learned" if self.corrector_on else self.fallback,
                   "theta": self.theta if self.corrector_on else self.fallback_theta,
                   "median": med, "p99": p99, "within": med <= self.budget}
            self.checks.append(rec)
            if self.corrector_on and not rec["within"]:
                # switch the corrector off and redo this step's walk without it
                # (gd_walk takes quadrupole cells when the moment level is >= 1;
                # this step's tree holds the second moments either way)
                self.corrector_on = False
                sim.set_quadrupole(1 if self.fallback == "quad" else 0)
                self._walk()
                med, p99 = self._judge(True)
                self.checks.append({"step": self.steps, "walk": self.fallback,
                                    "theta": self.fallback_theta, "median": med, "p99": p99,
                                    "within": med <= self.budget})
        sim.integrate(self.dt)
        self.steps += 1

    def run(self, steps):
        for _ in range(steps):
            self.step()

    def close(self):
        for b in (self._sample, self._exact, self._err):
            b.close()
        for e in self._ev:
            self.lib.gd_event_destroy(e)
        self._ev = ()
