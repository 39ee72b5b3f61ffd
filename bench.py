"""Benchmark: Barnes-Hut body-steps/s on one MI355X at <= 1e-3 median force error.

Workload (N=1): the 16M-body merger -- two exponential disks of 2^23 bodies
each, 30 degree inclination, 20 scale lengths apart on a parabolic approach --
with FP32 hi/lo positions, 63-bit keys, eps = 0.002, dt = 2.5e-4.  One step =
the full pipeline with every input resident in HBM: bbox -> Morton keys ->
radix sort -> reorder -> octree topology + centre of mass (+ the moments the
walk uses) -> walk -> hi/lo kick-drift.

The headline compares two walks at equal accuracy, each at the largest
opening angle whose median force error against the FP64 direct sum (seeded
4096-body subsample) is <= 1e-3:
  plain    monopole cells, the reference's force law
  learned  quadrupole cells plus the trained network's residual correction
           (trained first, on this run's own snapshots, with the BF16-MFMA
           kernels); when no opening angle meets the budget the corrector is
           switched off and `value` falls back to the plain walk.  After its
           timed steps the judge checks the learned walk again on the evolved
           state; over budget, the corrector is switched off the same way.
The analytic quadrupole walk (the corrector's control) is reported too.
`value` is the learned walk's body-steps/s (or the fallback's).

Also on the line:
  roofline      the dominant kernel, the walk's evaluation kernel in the
                chosen mode: FP32 FLOP per launch (20 per monopole
                interaction, 54 per quadrupole cell, 15 per learned
                correction; matrix-core FLOP of the network reported apart)
                over its duration from HIP events on the launch stream;
  cpu_baseline  the oracle (a serial C restatement of the reference's step)
                on one host core over a bounded sample of the same workload;
  reference_walk  the reference's 32-body warp-vote walk on the same tree.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-steps/s at ≤1e-3 median force err vs FP64 sum: learned vs plain walk"
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per launch of the evaluation kernel, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) scaled by what they read for a 1 GiB copy; keyed by
# (walk, theta, bodies) of the workload they were measured on.
MEASURED_TRAFFIC = {
    ("learned", 0.75, 1 << 21): {"bytes": 630689294,
                                 "profile": "profiles/pmc_walk_eval_learned_theta0.75_merger2M.md"},
    ("learned", 0.675, 1 << 24): {"bytes": 6329801814,
                                  "profile": "profiles/pmc_walk_eval_learned_theta0.675_merger16M.md"},
    ("learned", 0.7, 1 << 24): {"bytes": 5994216230,
                                "profile": "profiles/pmc_walk_eval_learned_theta0.7_merger16M.md"},
    ("learned", 0.6875, 1 << 24): {"bytes": 6084437655,
                                   "profile": "profiles/pmc_walk_eval_learned_theta0.6875_merger16M.md"},
}
BUDGET = 1e-3
# Each walk's opening angle is calibrated against this share of the budget:
# the learned walk is judged again after its timed steps, on the evolved
# state, and falls back to the plain walk if it is then over budget.  On the
# 2^24-body merger the evolved state's error was 0.2-1.7 % below the
# calibration's in every run seen, so 2 % of the budget is left as margin.
CALIBRATION_SHARE = 0.98
FLOP_MONO = 20
FLOP_QUAD = 54
FLOP_CORR = 15
FLOP_MLP = 17792
FLOP_OPEN = 9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--per-disk", type=int, default=1 << 23)
    ap.add_argument("--eps", type=float, default=0.002)
    ap.add_argument("--dt", type=float, default=2.5e-4)
    ap.add_argument("--no-hilo", action="store_true", help="plain FP32 positions")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-reference-walk", action="store_true")
    ap.add_argument("--cpu-group-stride", type=int, default=0,
                    help="walk every k-th 32-body group on the CPU (0: about 8192 groups)")
    ap.add_argument("--train-epochs", type=int, default=10)
    ap.add_argument("--min-ratio", type=float, default=0.5875,
                    help="the corrector acts on cells with size / distance above this")
    ap.add_argument("--train-targets", type=int, default=8192,
                    help="target bodies per snapshot for the corrector's samples")
    return ap.parse_args()


def cpu_baseline(pos, vel, eps, theta, dt, stride):
    """Oracle step on one host core: full build, every stride-th 32-body walk
    group (scaled by stride), full integrate."""
    from oracle import binding as ob
    n = pos.shape[0]
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    groups = (n + 31) // 32
    stride = stride or max(1, groups // 8192)
    t0 = time.perf_counter()
    t = ob.build_tree(pos, vel)
    t1 = time.perf_counter()
    acc, _, _ = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], np.float32(eps * eps),
                        1.0, theta, 0, groups, stride)
    t2 = time.perf_counter()
    ob.integrate(t["pos"], t["vel"], acc, dt)
    t3 = time.perf_counter()
    walked = len(range(0, groups, stride))
    step_s = (t1 - t0) + (t2 - t1) * groups / walked + (t3 - t2)
    return {
        "value": n / step_s,
        "unit": "body-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"one step of the same {n}-body workload at theta={theta}: full build "
                   f"({t1 - t0:.2f} s) and kick-drift ({t3 - t2:.2f} s); reference walk on "
                   f"{walked} of {groups} 32-body groups (every {stride}th, {t2 - t1:.2f} s) "
                   f"scaled by {groups / walked:.1f}"),
        "host_cores": os.cpu_count(),
    }


class Walks:
    """The three walks on one simulation."""

    def __init__(self, sim, net, min_ratio):
        self.sim = sim
        self.net = net
        self.min_ratio = min_ratio

    def prepare(self, mode):
        # moments are built only for the walks that use them
        self.sim.set_quadrupole({"plain": 0, "quad": 1, "learned": 2}[mode])

    def walk(self, mode, theta, stats=False):
        from cudagravitydemo_amd import corrector as cr
        if mode == "learned":
            return cr.walk_learned(self.sim, self.net, theta, self.min_ratio, stats=stats)
        return self.sim.walk(theta, stats=stats)

    def step(self, mode, theta, dt):
        self.sim.build_tree_async()
        self.walk(mode, theta)
        self.sim.integrate(dt)


def median_err(sim, sample, exact):
    a = sim.accelerations()[sample]
    e = np.linalg.norm(a - exact, axis=1) / np.linalg.norm(exact, axis=1)
    return float(np.median(e)), float(np.quantile(e, 0.99))


def calibrate(walks, mode, thetas, sample):
    """Largest theta (thetas descending) whose median error on the current
    state is within CALIBRATION_SHARE of the budget."""
    sim = walks.sim
    walks.prepare(mode)
    sim.build_tree_async()
    exact = sim.direct_sum_fp64(sample)
    tried = []
    for th in thetas:
        walks.walk(mode, th)
        med, p99 = median_err(sim, sample, exact)
        tried.append({"theta": th, "median": med, "p99": p99})
        if med <= CALIBRATION_SHARE * BUDGET:
            return th, med, p99, tried
    return None, None, None, tried


def timed_steps(walks, mode, theta, dt, steps, warmup):
    sim = walks.sim
    walks.prepare(mode)
    for _ in range(warmup):
        walks.step(mode, theta, dt)
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        walks.step(mode, theta, dt)
    sim.sync()
    return time.perf_counter() - t0


def eval_roofline(walks, mode, theta, n):
    lib = walks.sim.lib
    walks.prepare(mode)
    walks.sim.build_tree_async()
    lib.gd_walk_timing(1)
    ph = (C.c_float * 3)()
    phases = []
    st = None
    for _ in range(5):
        st = walks.walk(mode, theta, stats=True)
        lib.gd_walk_phase_ms(ph)
        phases.append(list(ph))
    lib.gd_walk_timing(0)
    phases = np.array(phases)
    ev_ms = float(phases[:, 1].mean())
    cells = st["cell_interactions"]
    mono = st["interactions"] - cells
    if mode == "plain":
        valu = FLOP_MONO * st["interactions"]
    elif mode == "quad":
        valu = FLOP_MONO * mono + FLOP_QUAD * cells
    else:
        valu = FLOP_MONO * mono + (FLOP_QUAD + FLOP_CORR) * cells
    mfma = FLOP_MLP * st.get("network_evals", 0) if mode == "learned" else 0
    achieved = valu / (ev_ms * 1e-3) / 1e12
    return {
        "bound": "mfma",
        "compute": ("FP32 vector ALU (peak = MI355X FP32; packed FP32 is not used, so 78.6 is "
                    "the practical ceiling)"),
        "kernel": f"k_walk_eval_pair ({mode} walk, theta={theta})",
        "achieved": achieved,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS,
        "traffic": (MEASURED_TRAFFIC.get((mode, theta, n)) or {}).get("bytes"),
        "traffic_unit": "bytes per launch (HBM read + write)",
        "traffic_source": (MEASURED_TRAFFIC.get((mode, theta, n)) or {}).get("profile"),
        "kernel_ms": ev_ms,
        "traversal_ms": float(phases[:, 0].mean()),
        "fallback_ms": float(phases[:, 2].mean()),
        "flop_per_launch": int(valu),
        "mfma_flop_per_launch": int(mfma),
        "interactions_per_body": st["interactions"] / n,
        "cell_interactions_per_body": cells / n,
        "overflow_groups": st["overflow_groups"],
    }


def log(msg):
    """Progress of the run, on stderr (stdout carries the one JSON line)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(no multi-GPU path is built)", file=sys.stderr)
        sys.exit(1)
    from cudagravitydemo_amd import Simulation, ics
    from cudagravitydemo_amd import corrector as cr

    eps, dt = args.eps, args.dt
    pos, vel = ics.disk_merger(args.per_disk, seed=2024)
    n = pos.shape[0]
    sim = Simulation(pos, vel, eps=eps, hilo=not args.no_hilo)
    net = cr.Corrector(seed=1)
    walks = Walks(sim, net, args.min_ratio)
    sample = ics.subsample(n, 4096, seed=99)
    for _ in range(args.warmup):
        walks.step("plain", 0.5, dt)

    # plain walk
    th_p, med_p, p99_p, tried_p = calibrate(walks, "plain", (0.3, 0.25, 0.2, 0.15, 0.1), sample)
    th_pt = th_p if th_p is not None else tried_p[-1]["theta"]
    secs_p = timed_steps(walks, "plain", th_pt, dt, args.steps, 1)
    log(f"plain walk: theta {th_pt}, {1e3 * secs_p / args.steps:.2f} ms/step")

    # corrector: samples from this run's own snapshots, BF16-MFMA training
    t0 = time.perf_counter()
    sim.set_quadrupole(2)
    samples = cr.SampleSet(sim.lib, 24_000_000)
    for snap in range(4):
        sim.build_tree_async()
        samples.add_from(sim, cr.default_targets(n, args.train_targets, seed=snap), 0.9,
                         args.min_ratio)
        for _ in range(5):
            walks.step("quad", 0.6, dt)
    samples.canonical_order()   # the same network from the same samples, run after run
    log(f"corrector: {samples.n} samples")
    losses = cr.train(net, samples, epochs=args.train_epochs)
    train_s = time.perf_counter() - t0
    log(f"corrector trained in {train_s:.1f} s")
    n_samples = samples.n
    samples.close()

    thetas = (1.0, 0.9, 0.85, 0.8, 0.75, 0.725, 0.7, 0.6875, 0.675, 0.65, 0.625, 0.6, 0.5)
    th_l, med_l, p99_l, tried_l = calibrate(walks, "learned", thetas, sample)
    corrector_on = th_l is not None
    secs_l = timed_steps(walks, "learned", th_l, dt, args.steps, 1) if corrector_on else None
    if corrector_on:
        log(f"learned walk: theta {th_l}, {1e3 * secs_l / args.steps:.2f} ms/step")
    after = None
    if corrector_on:
        # the judge again, on the state the timed steps left
        walks.prepare("learned")
        sim.build_tree_async()
        exact = sim.direct_sum_fp64(sample)
        walks.walk("learned", th_l)
        med_a, p99_a = median_err(sim, sample, exact)
        after = {"median_force_err": med_a, "p99_force_err": p99_a,
                 "within_budget": med_a <= BUDGET}
        if med_a > BUDGET:
            corrector_on = False

    # analytic quadrupole control
    th_q, med_q, p99_q, tried_q = calibrate(walks, "quad", thetas, sample)
    secs_q = timed_steps(walks, "quad", th_q if th_q is not None else 0.5, dt, args.steps, 1)
    log(f"quadrupole walk: theta {th_q}, {1e3 * secs_q / args.steps:.2f} ms/step")

    mode = "learned" if corrector_on else "plain"
    theta = th_l if corrector_on else th_pt
    secs = secs_l if corrector_on else secs_p
    roof = eval_roofline(walks, mode, theta, n)
    log(f"roofline: {roof['achieved']:.1f} {roof['unit']} on {roof['kernel']}")

    ref = None
    if not args.no_reference_walk:
        walks.prepare("plain")
        sim.build_tree()
        exact = sim.direct_sum_fp64(sample)
        lib = sim.lib
        e0, e1 = lib.gd_event_create(), lib.gd_event_create()
        lib.gd_event_record(e0)
        sim.walk_reference(0.5)
        lib.gd_event_record(e1)
        ref_ms = lib.gd_event_elapsed_ms(e0, e1)
        med_r, _ = median_err(sim, sample, exact)
        ref = {"walk_ms": ref_ms, "median_force_err": med_r, "theta": 0.5,
               "note": "32-body warp-vote walk (tree.cu:562-676), same tree"}
    final_pos, final_vel = sim.positions(), sim.velocities()
    net.close()
    sim.close()

    def leg(th, med, p99, secs_, tried):
        if th is None:
            return {"theta": None, "within_budget": False, "tried": tried}
        return {"theta": th, "median_force_err": med, "p99_force_err": p99,
                "body_steps_per_s": n * args.steps / secs_, "ms_per_step": 1e3 * secs_ / args.steps,
                "tried": tried}

    out = {
        "metric": METRIC,
        "value": n * args.steps / secs,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"two-disk merger, {n} bodies, "
                         f"{'plain FP32' if args.no_hilo else 'FP32 hi/lo'} positions, "
                         f"eps={eps}, dt={dt}, full step, {mode} walk at theta={theta}"),
            "n_bodies": n,
            "eps": eps,
            "dt": dt,
            "walk": mode,
            "theta": theta,
            "force_error_budget": BUDGET,
            "calibration_budget": CALIBRATION_SHARE * BUDGET,
            "judge": "FP64 direct sum, 4096-body seeded subsample",
        },
        "plain": leg(th_p, med_p, p99_p, secs_p, tried_p),
        "learned": leg(th_l, med_l, p99_l, secs_l, tried_l),
        "quadrupole_control": leg(th_q, med_q, p99_q, secs_q, tried_q),
        "learned_vs_plain": (secs_p / secs_l) if corrector_on else None,
        "judge_after_timed_steps": after,
        "corrector": {"on": corrector_on, "train_s": train_s, "samples": n_samples,
                      "epochs": args.train_epochs, "final_loss": losses[-1] if losses else None,
                      "first_loss": losses[0] if losses else None},
        "roofline": roof,
        "reference_walk": ref,
    }
    if not args.no_cpu:
        log("CPU baseline (host only)")
        out["cpu_baseline"] = cpu_baseline(final_pos, final_vel, eps, 0.5, dt,
                                           args.cpu_group_stride)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
