"""Host-side mirror of the reference's main loop over the HIP engine.

Simulation owns one treeState and the body arrays on the device and offers
the reference's step (basicGravity.cu:767-783: buildTree, then
calculateAcceleration, then positionUpdater) plus the engine's own step
(gd_build_tree_async, the wavefront walk gd_walk, kick-drift with optional
hi/lo positions).  Every call goes through libgravity.so; nothing here
computes physics on the CPU.
"""
import ctypes as C

import numpy as np

from ._lib import GravityError, TreeState, WalkStats, WorldBox, check, load_library


class DeviceBuffer:
    """A device allocation from gd_malloc, freed on close()."""

    def __init__(self, lib, nbytes):
        self.lib = lib
        self.nbytes = int(nbytes)
        self.ptr = lib.gd_malloc(max(self.nbytes, 1))
        check(lib)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, (arr.nbytes, self.nbytes)
        self.lib.gd_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes)
        check(self.lib)

    def close(self):
        if self.ptr:
            self.lib.gd_free(self.ptr)
            self.ptr = None


def to_device(lib, arr):
    buf = DeviceBuffer(lib, np.ascontiguousarray(arr).nbytes)
    buf.upload(arr)
    return buf


def from_device(lib, ptr, dtype, shape):
    out = np.empty(shape, dtype=dtype)
    if out.nbytes:
        lib.gd_memcpy_d2h(out.ctypes.data, ptr, out.nbytes)
        check(lib)
    return out


class Simulation:
    """Bodies + octree state on one GPU.

    pos_mass: (N, 4) float32 (x, y, z, m); vel: (N, 3) float32.
    hilo: keep a float3 low part of the positions (precision form).
    """

    def __init__(self, pos_mass, vel, eps=0.01, G=1.0, hilo=False, exit_on_error=False):
        self.lib = lib = load_library()
        lib.gd_set_exit_on_error(1 if exit_on_error else 0)
        pos_mass = np.ascontiguousarray(pos_mass, dtype=np.float32)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        n = pos_mass.shape[0]
        self.n = n
        self.eps = float(eps)
        self.G = float(G)
        self.padded = lib.roundNumberToMultiple(n, 32)
        integrate_blocks = self.padded // 256 + (1 if n % 256 else 0)
        self.state = TreeState()
        lib.initialiseTree(C.byref(self.state), n, self.padded, integrate_blocks)
        check(lib)
        pad_pos = np.zeros((self.padded, 4), np.float32)
        pad_pos[:n] = pos_mass
        pad_vel = np.zeros((self.padded, 3), np.float32)
        pad_vel[:n] = vel
        self._bufs = []
        self._pos0 = to_device(lib, pad_pos)
        self._vel0 = to_device(lib, pad_vel)
        self._acc = to_device(lib, np.zeros((self.padded, 3), np.float32))
        self._phi = DeviceBuffer(lib, 4 * self.padded)
        self.pos = C.c_void_p(self._pos0.ptr)
        self.vel = C.c_void_p(self._vel0.ptr)
        self.acc = self._acc.ptr
        self.hilo = bool(hilo)
        if self.hilo:
            self._lo0 = to_device(lib, np.zeros((self.padded, 3), np.float32))
            self._lo1 = to_device(lib, np.zeros((self.padded, 3), np.float32))
            self.lo = C.c_void_p(self._lo0.ptr)
            self.lo_spare = C.c_void_p(self._lo1.ptr)
        else:
            self.lo = C.c_void_p(None)
            self.lo_spare = C.c_void_p(None)
        self._records = True          # engine builds write the node array too
        self._records_stale = False   # the node array lags the last build
        self.closed = False

    # ---- the reference's step -------------------------------------------
    def build_tree(self):
        """buildTree (basicGravity.cu:116-325): sorts the bodies, swaps pointers."""
        self.lib.buildTree(C.byref(self.state), C.byref(self.pos), C.byref(self.vel))
        check(self.lib)
        self._records_stale = False

    def walk_reference(self, theta):
        """calculateAcceleration (tree.cu:562-676) into self.acc."""
        self._records_now()
        self.lib.calculateAcceleration(self.pos, self.state, self.acc, self.eps * self.eps,
                                       self.G, theta, self.n)
        check(self.lib)

    def integrate(self, dt):
        """positionUpdater (tree.cu:678-703) or its hi/lo form."""
        if self.hilo:
            self.lib.gd_integrate_hilo(self.pos, self.lo, self.vel, self.acc, self.n, dt)
        else:
            self.lib.positionUpdater(self.pos, self.vel, self.acc, self.n, dt)
        check(self.lib)

    def step_reference(self, dt, theta):
        """One step of the reference's main loop (basicGravity.cu:774-783)."""
        self.build_tree()
        self.walk_reference(theta)
        self.integrate(dt)

    # ---- the engine's step ----------------------------------------------
    def build_tree_async(self):
        lo = C.byref(self.lo) if self.hilo else None
        spare = C.byref(self.lo_spare) if self.hilo else None
        self.lib.gd_build_tree_async(C.byref(self.state), C.byref(self.pos), C.byref(self.vel),
                                     lo, spare)
        check(self.lib)
        self._records_stale = not self._records

    def set_node_records(self, enable):
        """Whether engine builds also write the reference's 80-byte node array
        (gd_set_node_records; on by default).  Off, the engine step writes
        only the walk arrays, and the node array is written from them when
        something reads it (nodes_raw, walk_reference)."""
        self._records = bool(enable)
        self.lib.gd_set_node_records(C.byref(self.state), 1 if enable else 0)
        check(self.lib)

    def _records_now(self):
        if self._records_stale:
            self.lib.gd_write_node_records(C.byref(self.state))
            check(self.lib)
            self._records_stale = False

    def set_quadrupole(self, enable):
        """Compute second moments in later builds (the walk then uses them);
        enable=2 also computes the third moments (the learned walk's inputs)."""
        level = 2 if enable == 2 else (1 if enable else 0)
        self.lib.gd_set_quadrupole(C.byref(self.state), level)
        check(self.lib)

    def walk(self, theta, potential=False, stats=False):
        st = WalkStats() if stats else None
        self.lib.gd_walk(C.byref(self.state), self.pos, self.acc,
                         self._phi.ptr if potential else None, self.eps * self.eps, self.G,
                         theta, self.n, C.byref(st) if st is not None else None)
        check(self.lib)
        if st is not None:
            return {"interactions": st.interactions, "opened": st.opened,
                    "overflow_groups": st.overflowGroups,
                    "cell_interactions": st.cellInteractions,
                    "body_groups": st.bodyGroups, "network_evals": st.networkEvals,
                "far_cells": st.farCells,
                "far_interactions": st.farInteractions}
        return None

    def set_far_field(self, ratio):
        """Far cells of the quadrupole and learned list walks through one local
        expansion per group (gd_walk_set_far_field): cells with h / R below
        `ratio` (h: the group box's half diagonal, R: distance from its centre);
        0 evaluates every cell body by body.  Returns the previous setting.  The
        setting belongs to the library, not to this state."""
        prev = self.lib.gd_walk_set_far_field(float(ratio))
        check(self.lib)
        return float(prev)

    def set_opening_rule(self, rule):
        """Opening rule of the production walks (gd_walk_set_opening_rule):
        "size" (cell size, the default) or "radius" (the cells' mass radii).
        Returns the previous rule.  The setting belongs to the library."""
        code = {"size": 0, "radius": 1}[rule]
        prev = self.lib.gd_walk_set_opening_rule(code)
        check(self.lib)
        return ("size", "radius")[prev]

    def step(self, dt, theta):
        self.build_tree_async()
        self.walk(theta)
        self.integrate(dt)

    def set_walk_async(self, enable):
        """Walks without host synchronisation (gd_walk_set_async); with it on,
        step() queues its kernels and returns at once.  Returns the previous
        setting.  The setting belongs to the library, not to this state."""
        prev = self.lib.gd_walk_set_async(1 if enable else 0)
        check(self.lib)
        return bool(prev)

    def capture_steps(self, dt, theta, steps=2):
        """A HIP graph of `steps` engine steps (an even number: the graph must
        leave the ping-pong buffers where it found them).  Walks must be
        asynchronous (set_walk_async) and every buffer allocated by eager steps
        first.  Returns the graph; replay() runs it, close_graph() frees it."""
        if steps <= 0 or steps % 2:
            raise ValueError("capture an even number of steps")
        self.lib.gd_graph_begin()
        check(self.lib)
        try:
            for _ in range(steps):
                self.build_tree_async()
                self.walk(theta)
                self.integrate(dt)
        finally:
            graph = self.lib.gd_graph_end()
        check(self.lib)
        return graph

    def replay(self, graph):
        self.lib.gd_graph_launch(graph)
        check(self.lib)

    def close_graph(self, graph):
        self.lib.gd_graph_destroy(graph)
        check(self.lib)

    def poll(self):
        """Waits for the last asynchronous walk's status; raises GravityError
        for a deferred error (a tree that did not fit its node array)."""
        self.lib.gd_walk_poll()
        check(self.lib)

    # ---- inspection -------------------------------------------------------
    def sync(self):
        self.lib.gd_synchronize()
        check(self.lib)

    def positions(self):
        return from_device(self.lib, self.pos.value, np.float32, (self.n, 4))

    def positions_lo(self):
        if not self.hilo:
            return np.zeros((self.n, 3), np.float32)
        return from_device(self.lib, self.lo.value, np.float32, (self.n, 3))

    def velocities(self):
        return from_device(self.lib, self.vel.value, np.float32, (self.n, 3))

    def accelerations(self):
        return from_device(self.lib, self.acc, np.float32, (self.n, 3))

    def potentials(self):
        return from_device(self.lib, self._phi.ptr, np.float32, (self.n,))

    def set_state(self, pos_mass, vel, lo=None):
        """Overwrite the current body arrays (unsorted input is fine); lo, the
        low part of hi/lo positions in the same order, is zeroed when not given."""
        pos_mass = np.ascontiguousarray(pos_mass, dtype=np.float32)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        assert pos_mass.shape == (self.n, 4) and vel.shape == (self.n, 3)
        self.lib.gd_memcpy_h2d(self.pos.value, pos_mass.ctypes.data, pos_mass.nbytes)
        self.lib.gd_memcpy_h2d(self.vel.value, vel.ctypes.data, vel.nbytes)
        if self.hilo:
            lo = np.zeros((self.n, 3), np.float32) if lo is None else \
                np.ascontiguousarray(lo, dtype=np.float32)
            assert lo.shape == (self.n, 3)
            self.lib.gd_memcpy_h2d(self.lo.value, lo.ctypes.data, lo.nbytes)
        check(self.lib)

    def snapshot(self):
        """Host copies (pos_mass, vel, lo or None) of the current state, for set_state."""
        return (self.positions(), self.velocities(),
                self.positions_lo() if self.hilo else None)

    def keys_sorted(self):
        return from_device(self.lib, self.state.mortonOut, np.uint64, (self.n,))

    def keys_unsorted(self):
        return from_device(self.lib, self.state.mortonIn, np.uint64, (self.n,))

    def permutation(self):
        return from_device(self.lib, self.state.indexOut, np.uint32, (self.n,))

    def bbox(self):
        lo = from_device(self.lib, self.state.boundingData[2], np.float32, (3,))
        hi = from_device(self.lib, self.state.boundingData[3], np.float32, (3,))
        return lo, hi

    def tree_summary(self):
        total, overflow, zero = C.c_int(), C.c_int(), C.c_int()
        self.lib.gd_tree_summary(C.byref(self.state), C.byref(total), C.byref(overflow),
                                 C.byref(zero))
        check(self.lib)
        return total.value, overflow.value, zero.value

    def nodes_raw(self, count=None):
        """The node array as raw bytes, shape (count, 80)."""
        self._records_now()
        if count is None:
            count = self.tree_summary()[0]
        return from_device(self.lib, self.state.nodes, np.uint8, (count, 80))

    def walk_arrays(self, count):
        """Compact walk records of the first `count` nodes: (COM+mass (count,4),
        link (count,2), second moments (count,8) or None)."""
        com = np.empty((count, 4), np.float32)
        link = np.empty((count, 2), np.int32)
        quad = np.empty((count, 8), np.float32)
        self.lib.gd_tree_walk_arrays(C.byref(self.state), com.ctypes.data, link.ctypes.data,
                                     quad.ctypes.data, count)
        check(self.lib)
        return com, link, quad

    def mass_radii(self, count):
        """Mass radii (count,) of the first `count` nodes (gd_tree_mass_radii)."""
        out = np.empty(count, np.float32)
        self.lib.gd_tree_mass_radii(C.byref(self.state), out.ctypes.data, count)
        check(self.lib)
        return out

    def third_moments(self, count):
        """Third moments (count, 12) of the first `count` nodes: Oxxx Oyyy Ozzz
        Oxxy Oxxz Oxyy Oyyz Oxzz Oyzz Oxyz 0 0."""
        out = np.empty((count, 12), np.float32)
        self.lib.gd_tree_third_moments(C.byref(self.state), out.ctypes.data, count)
        check(self.lib)
        return out

    def root_fp64(self):
        """The last build's root-level centre of mass and mass, summed in FP64
        over all bodies (hi + lo): array (cx, cy, cz, M)."""
        out = (C.c_double * 4)()
        self.lib.gd_tree_root_fp64(C.byref(self.state), out)
        check(self.lib)
        return np.array(list(out))

    def levels(self):
        """Host copy of levelOffsets / levelCount / deepestLevel."""
        self.lib.gd_tree_levels(C.byref(self.state))
        check(self.lib)
        d = self.state.deepestLevel
        return list(self.state.levelOffsets), list(self.state.levelCount), d

    def direct_sum_fp64(self, sample, with_lo=True):
        """FP64 direct-sum accelerations (ns, 3) of the bodies at sorted indices
        `sample` (positions hi + lo unless with_lo is False)."""
        sample = np.ascontiguousarray(sample, dtype=np.int32)
        ns = sample.shape[0]
        sbuf = to_device(self.lib, sample)
        obuf = DeviceBuffer(self.lib, 8 * 3 * ns)
        try:
            self.lib.gd_direct_sum_fp64(self.pos, self.lo if (self.hilo and with_lo) else None,
                                        self.n,
                                        sbuf.ptr, ns, self.eps * self.eps, self.G, obuf.ptr)
            check(self.lib)
            return from_device(self.lib, obuf.ptr, np.float64, (ns, 3))
        finally:
            sbuf.close()
            obuf.close()

    def diagnostics(self, with_potential=False):
        out = (C.c_double * 7)()
        self.lib.gd_diagnostics(self.pos, self.vel, self._phi.ptr if with_potential else None,
                                self.n, out)
        check(self.lib)
        keys = ["kinetic", "px", "py", "pz", "mass", "potential", "total"]
        return dict(zip(keys, list(out)))

    def close(self):
        if self.closed:
            return
        lib = self.lib
        st = self.state
        for p in (st.nodes, st.mortonIn, st.mortonOut, st.indexIn, st.indexOut,
                  st.boundingData[0], st.boundingData[1], st.boundingData[2],
                  st.boundingData[3], st.sortedPosMassVals, st.sortedVelVals,
                  self.pos.value, self.vel.value):
            if p:
                lib.gd_free(p)
        self._acc.close()
        self._phi.close()
        if self.hilo:
            lib.gd_free(self.lo.value)
            lib.gd_free(self.lo_spare.value)
        if st.sorter:
            lib.sorterDestroyer(st.sorter)
        if st.builder:
            lib.sumDestroyer(st.builder)
        self.closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ = ["Simulation", "DeviceBuffer", "to_device", "from_device", "GravityError", "WorldBox"]
