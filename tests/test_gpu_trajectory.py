"""The C driver (gravity_demo.c, the reference's main loop against the C-ABI)
and its trajectory file, replayed by the oracle.

The driver runs inside this process (libgravity_demo.so: the same program
with main renamed) on the reference's own scenario -- an exponential disk
plus a 20 % halo (initialConditions.c:246-382) -- and writes the reference's
format (basicGravity.cu:614-635, 793-801; inputOutput.c:41-58): int32 N,
int32 frames, RGB[N], then float32[3 N] per frame.  The oracle steps the same
bodies (buildTree, calculateAcceleration, positionUpdater per step) and
carries each step's sort permutation, so every frame can be compared body by
body in the original order.
"""
import ctypes as C
import os

import numpy as np
import pytest

from cudagravitydemo_amd import ics, load_library
from oracle import binding as ob

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "cudagravitydemo_amd", "libgravity_demo.so")


def run_demo(args):
    load_library()                        # the driver shares this libgravity.so
    demo = C.CDLL(DEMO)
    demo.gd_demo_main.restype = C.c_int
    demo.gd_demo_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
    argv = (C.c_char_p * (len(args) + 1))(b"gravity_demo", *[a.encode() for a in args])
    return demo.gd_demo_main(len(args) + 1, argv)


@pytest.mark.parametrize("n,steps,every", [(2000, 20, 5), (777, 12, 3)])
def test_trajectory_frames_in_original_order(tmp_path, n, steps, every):
    theta, eps, dt = 0.5, 0.02, 0.08          # the reference run's eps and dt
    pos, vel = ics.reference_scenario(n, seed=11)
    icf, out = str(tmp_path / "ic.bin"), str(tmp_path / "outputDump.bin")
    ics.write_ic_file(icf, pos, vel)
    print(f"case gravity_demo N={n} steps={steps} frames every {every}", flush=True)
    rc = run_demo([str(n), str(steps), str(theta), str(eps), str(dt), str(every), out, icf])
    assert rc == 0
    got_n, colours, frames = ics.read_trajectory(out)
    assert got_n == n and frames.shape == (steps // every, n, 3)
    assert np.all(colours == 100)

    # the oracle's run, identities carried through every step's permutation
    ident = np.arange(n)
    p, v = pos.copy(), vel.copy()
    worst = 0.0
    for step in range(steps):
        t, _, p, v = ob.step(p, v, eps, 1.0, theta, dt)
        ident = ident[t["perm"]]
        if step % every == 0 and step // every < frames.shape[0]:
            want = np.empty((n, 3), np.float32)
            want[ident] = p[:, :3]
            diff = np.abs(frames[step // every] - want).max()
            worst = max(worst, float(diff))
    print(f"largest difference from the oracle's trajectory: {worst:.3e}", flush=True)
    # the per-step parity bar is bit-exact arithmetic; the frames in the
    # wrong order would differ by the size of the system
    assert worst <= 1e-5 * np.abs(pos[:, :3]).max()
