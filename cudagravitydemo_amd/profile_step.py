"""Profiling driver: a few engine steps of the two-disk merger, for rocprofv3.

Runs the same workload as bench.py (2^21 bodies by default) for --steps
steps with the chosen walk, then (with --copy-mib) one stream copy of a known
size, so that hardware byte counters of the step's kernels can be scaled by
what the same counters read for the copy.  Prints one line per phase.
--walk learned uses an untrained network (the kernels' work does not depend
on the weights); --params loads trained ones (float32 .npy, 9091 values).

Examples (from the repository root):
  rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- python3 cudagravitydemo_amd/profile_step.py
  rocprofv3 --pmc FETCH_SIZE -d OUT -o fetch -- python3 cudagravitydemo_amd/profile_step.py --steps 2 --copy-mib 256
"""
import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

from cudagravitydemo_amd import Simulation, ics  # noqa: E402
from cudagravitydemo_amd import corrector as cr  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-disk", type=int, default=1 << 20)
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--hilo", action="store_true", help="FP32 hi/lo positions")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--walk", choices=["plain", "quad", "learned", "reference"], default="plain")
    ap.add_argument("--params", default=None)
    ap.add_argument("--min-ratio", type=float, default=0.55,
                    help="learned walk: the corrector acts on cells with size / distance above this")
    ap.add_argument("--list-cap", type=int, default=0,
                    help="fixed first-pass list capacity per group (0: adaptive)")
    ap.add_argument("--group-split", type=float, default=-1.0,
                    help="volume fraction for splitting 64-body groups (< 0: by theta)")
    ap.add_argument("--far", type=float, default=0.0,
                    help="quadrupole and learned walks: far-field ratio (gd_walk_set_far_field)")
    ap.add_argument("--opening", choices=["size", "radius"], default="size",
                    help="opening rule of the production walks (gd_walk_set_opening_rule)")
    ap.add_argument("--phases", action="store_true",
                    help="time the walk's phases with HIP events (last step's walk)")
    ap.add_argument("--copy-mib", type=int, default=0)
    a = ap.parse_args()
    pos, vel = ics.disk_merger(a.per_disk, seed=2024)
    n = pos.shape[0]
    eps, dt = a.eps, a.dt
    net = None
    if a.walk == "learned":
        net = cr.Corrector(np.load(a.params) if a.params else None, seed=1)
    with Simulation(pos, vel, eps=eps, hilo=a.hilo) as sim:
        if a.walk != "reference":
            sim.set_node_records(False)   # as bench.py's engine step
        if a.list_cap:
            sim.lib.gd_walk_set_list_capacity(a.list_cap, 0)
        sim.lib.gd_walk_set_group_split(a.group_split)
        sim.set_far_field(a.far)
        sim.set_opening_rule(a.opening)
        if a.walk in ("quad", "learned"):
            sim.set_quadrupole(2 if a.walk == "learned" else 1)
        sim.sync()
        t0 = time.perf_counter()
        for it in range(a.warmup + a.steps):
            if it == a.warmup:
                sim.sync()
                t0 = time.perf_counter()
            if a.walk == "reference":
                sim.step_reference(dt, a.theta)
            elif a.walk == "learned":
                sim.build_tree_async()
                cr.walk_learned(sim, net, a.theta, a.min_ratio)
                sim.integrate(dt)
            else:
                sim.step(dt, a.theta)
        sim.sync()
        t1 = time.perf_counter()
        print(f"profile_step: {n} bodies, {a.steps} steps, walk={a.walk}, theta={a.theta}, "
              f"far field {a.far}, opening rule {a.opening}: "
              f"{1e3 * (t1 - t0) / a.steps:.3f} ms/step", flush=True)
        if a.phases and a.walk != "reference":
            import ctypes as C
            sim.lib.gd_walk_timing(1)
            sim.build_tree_async()
            if a.walk == "learned":
                cr.walk_learned(sim, net, a.theta, a.min_ratio)
            else:
                sim.walk(a.theta)
            ms = (C.c_float * 3)()
            sim.lib.gd_walk_phase_ms(C.cast(ms, C.c_void_p))
            sim.lib.gd_walk_timing(0)
            print(f"profile_step: walk phases: traversal {ms[0]:.3f} ms, evaluation {ms[1]:.3f} "
                  f"ms, second pass and per-body walk {ms[2]:.3f} ms", flush=True)
        if a.copy_mib:
            lib = sim.lib
            nbytes = a.copy_mib << 20
            src = lib.gd_malloc(nbytes)
            dst = lib.gd_malloc(nbytes)
            lib.gd_memset(src, 1, nbytes)
            lib.gd_synchronize()
            e0, e1 = lib.gd_event_create(), lib.gd_event_create()
            lib.gd_event_record(e0)
            lib.gd_stream_copy(dst, src, nbytes)
            lib.gd_event_record(e1)
            ms = lib.gd_event_elapsed_ms(e0, e1)
            print(f"profile_step: stream copy {a.copy_mib} MiB in {ms:.3f} ms = "
                  f"{2 * nbytes / ms / 1e6:.1f} GB/s (read + write)", flush=True)
            lib.gd_free(src)
            lib.gd_free(dst)
    if net is not None:
        net.close()


if __name__ == "__main__":
    main()
