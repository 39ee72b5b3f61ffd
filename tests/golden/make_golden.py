"""Writes tests/golden/disk1024.npz from the CPU oracle.

The reference ships no fixtures, so these vectors come from the oracle after
tests/test_oracle.py has pinned it (naive interleave, stable sort, an
independent octree, FP64 sums).  They freeze the oracle's exact outputs for
one small seeded disk: inputs, keys, permutation, node array, accelerations
of the reference walk at theta 0.5, eps 0.01.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cudagravitydemo_amd import ics  # noqa: E402
from oracle import binding as ob  # noqa: E402


def main():
    pos, vel = ics.exponential_disk(1024, seed=77)
    t = ob.build_tree(pos, vel)
    acc, _, _ = ob.walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], np.float32(1e-4), 1.0,
                        0.5)
    raw = t["nodes"].view(np.uint8).reshape(-1, 80).copy()
    raw[:, 56:64] = 0
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "disk1024.npz")
    np.savez_compressed(out, pos=pos, vel=vel, keys=t["keys"], perm=t["perm"], nodes=raw, acc=acc)
    print("wrote", out, raw.shape)


if __name__ == "__main__":
    main()
