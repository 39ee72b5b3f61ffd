"""ctypes binding of oracle/liboracle.so (test infrastructure only).

Each wrapper takes and returns numpy arrays; the C functions are the serial
restatement of the reference documented in oracle/oracle.h.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

NODE_DTYPE = np.dtype([("path", "<u8"), ("parent", "<i4"), ("first", "<i4"), ("count", "<i4"),
                       ("child", "<i4", (8,)), ("level", "<i4"), ("pad", "<i4", (2,)),
                       ("mass", "<f4", (4,))])
assert NODE_DTYPE.itemsize == 80

# fields the reference defines for every node (the padding is never written)
NODE_FIELDS = ("path", "parent", "first", "count", "child", "level", "mass")


class _WorldBox(C.Structure):
    _fields_ = [("minX", C.c_float), ("minY", C.c_float), ("minZ", C.c_float),
                ("extent", C.c_float), ("scale", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        i32 = C.c_int
        f32 = C.c_float
        L.or_bbox.argtypes = [vp, i32, vp, vp]
        L.or_world_box.argtypes = [vp, vp, C.POINTER(_WorldBox)]
        L.or_morton.argtypes = [vp, i32, C.POINTER(_WorldBox), vp, vp]
        L.or_sort_pairs.argtypes = [vp, vp, vp, vp, i32]
        L.or_reorder.argtypes = [vp, vp, vp, vp, vp, i32]
        L.or_build_topology.argtypes = [vp, i32, vp, i32, vp, vp, vp]
        L.or_build_topology.restype = i32
        L.or_tree_mass.argtypes = [vp, vp, vp, vp, i32]
        L.or_tree_mass.restype = i32
        L.or_walk.argtypes = [vp, i32, vp, vp, vp, f32, f32, f32, vp, i32, i32, i32,
                              C.POINTER(C.c_longlong)]
        L.or_walk.restype = C.c_longlong
        L.or_walk_traced.argtypes = [vp, i32, vp, vp, vp, f32, f32, f32, vp,
                                     C.POINTER(C.c_longlong), i32, vp, vp, i32]
        L.or_walk_traced.restype = C.c_longlong
        L.or_integrate.argtypes = [vp, vp, vp, i32, f32]
        L.or_direct_sum_fp64.argtypes = [vp, i32, vp, i32, C.c_double, C.c_double, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def bbox(pos):
    pos = np.ascontiguousarray(pos, np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().or_bbox(_p(pos), pos.shape[0], _p(lo), _p(hi))
    return lo, hi


def world_box(lo, hi):
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    b = _WorldBox()
    lib().or_world_box(_p(lo), _p(hi), C.byref(b))
    return b


def make_box(minx, miny, minz, extent, scale):
    return _WorldBox(minx, miny, minz, extent, scale)


def morton(pos, box):
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    keys = np.zeros(n, np.uint64)
    idx = np.zeros(n, np.uint32)
    lib().or_morton(_p(pos), n, C.byref(box), _p(keys), _p(idx))
    return keys, idx


def sort_pairs(keys, vals):
    keys = np.ascontiguousarray(keys, np.uint64)
    vals = np.ascontiguousarray(vals, np.uint32)
    ko = np.zeros_like(keys)
    vo = np.zeros_like(vals)
    lib().or_sort_pairs(_p(keys), _p(ko), _p(vals), _p(vo), keys.shape[0])
    return ko, vo


def reorder(pos, vel, idx):
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    po = np.zeros_like(pos)
    vo = np.zeros_like(vel)
    lib().or_reorder(_p(pos), _p(vel), _p(po), _p(vo), _p(idx), pos.shape[0])
    return po, vo


def build_topology(sorted_keys, max_nodes=None):
    k = np.ascontiguousarray(sorted_keys, np.uint64)
    n = k.shape[0]
    if max_nodes is None:
        max_nodes = 8 * n
    nodes = np.zeros(max(max_nodes, 1), NODE_DTYPE)
    lo = np.zeros(22, np.int32)
    lc = np.zeros(22, np.int32)
    deep = np.zeros(1, np.int32)
    rc = lib().or_build_topology(_p(k), n, _p(nodes), max_nodes, _p(lo), _p(lc), _p(deep))
    d = int(deep[0])
    total = int(lo[d] + lc[d]) if rc == 0 else -1
    return rc, nodes, lo, lc, d, total


def tree_mass(nodes, pos, lo, lc, deepest):
    pos = np.ascontiguousarray(pos, np.float32)
    return lib().or_tree_mass(_p(nodes), _p(pos), _p(lo), _p(lc), int(deepest))


def walk(pos, nodes, bmin, bmax, eps2, G, theta, g0=0, g1=None, gstride=1):
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    if g1 is None:
        g1 = (n + 31) // 32
    acc = np.zeros((n, 3), np.float32)
    bmin = np.ascontiguousarray(bmin, np.float32)
    bmax = np.ascontiguousarray(bmax, np.float32)
    opened = C.c_longlong(0)
    inter = lib().or_walk(_p(pos), n, _p(nodes), _p(bmin), _p(bmax), eps2, G, theta, _p(acc),
                          g0, g1, gstride, C.byref(opened))
    return acc, int(inter), int(opened.value)


def walk_traced(pos, nodes, bmin, bmax, eps2, G, theta, variant=0, per_body=4096):
    """or_walk over all groups plus each body's accepted node indices.
    variant > 0 plants one of the oracle's deliberate mistakes (oracle.h).
    Returns (acc, interactions, opened, lists) with lists[i] an int32 array."""
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    acc = np.zeros((n, 3), np.float32)
    bmin = np.ascontiguousarray(bmin, np.float32)
    bmax = np.ascontiguousarray(bmax, np.float32)
    opened = C.c_longlong(0)
    counts = np.zeros(n, np.int32)
    lists = np.full((n, per_body), -1, np.int32)
    inter = lib().or_walk_traced(_p(pos), n, _p(nodes), _p(bmin), _p(bmax), eps2, G, theta,
                                 _p(acc), C.byref(opened), int(variant), _p(counts), _p(lists),
                                 per_body)
    if inter < 0:
        raise ValueError(f"or_walk_traced failed (variant {variant})")
    if counts.max(initial=0) > per_body:
        raise ValueError(f"a body accepted {counts.max()} nodes, more than per_body={per_body}")
    return acc, int(inter), int(opened.value), [lists[i, :counts[i]] for i in range(n)]


def integrate(pos, vel, acc, dt):
    """Returns updated copies."""
    pos = np.array(pos, np.float32, copy=True)
    vel = np.array(vel, np.float32, copy=True)
    acc = np.ascontiguousarray(acc, np.float32)
    lib().or_integrate(_p(pos), _p(vel), _p(acc), pos.shape[0], dt)
    return pos, vel


def direct_sum_fp64(pos, sample, eps2, G=1.0):
    pos = np.ascontiguousarray(pos, np.float32)
    sample = np.ascontiguousarray(sample, np.int32)
    out = np.zeros((sample.shape[0], 3), np.float64)
    lib().or_direct_sum_fp64(_p(pos), pos.shape[0], _p(sample), sample.shape[0], eps2, G,
                             _p(out))
    return out


def build_tree(pos, vel):
    """The reference's buildTree on host arrays: returns a dict with the sorted
    state, keys, permutation, nodes and level tables (basicGravity.cu:116-325)."""
    pos = np.ascontiguousarray(pos, np.float32)
    lo3, hi3 = bbox(pos)
    box = world_box(lo3, hi3)
    keys, idx = morton(pos, box)
    ks, perm = sort_pairs(keys, idx)
    ps, vs = reorder(pos, vel, perm)
    rc, nodes, lo, lc, d, total = build_topology(ks)
    if rc != 0:
        raise RuntimeError("oracle node overflow")
    zero = tree_mass(nodes, ps, lo, lc, d)
    return dict(bmin=lo3, bmax=hi3, box=box, keys=keys, keys_sorted=ks, perm=perm, pos=ps,
                vel=vs, nodes=nodes[:total], nodes_full=nodes, level_offsets=lo,
                level_count=lc, deepest=d, total=total, zero_mass=zero)


def step(pos, vel, eps, G, theta, dt):
    """One reference step on host arrays; returns (tree dict, acc, new pos, new vel)."""
    t = build_tree(pos, vel)
    acc, _, _ = walk(t["pos"], t["nodes_full"], t["bmin"], t["bmax"], np.float32(eps * eps), G,
                     theta)
    p2, v2 = integrate(t["pos"], t["vel"], acc, dt)
    return t, acc, p2, v2
