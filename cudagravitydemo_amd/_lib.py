"""ctypes binding of libgravity.so (the HIP engine's C-ABI).

The structs below mirror include/gravity_types.h, which keeps the byte
layout of the reference's types.h (types.h:7-105).  Loading fails loudly when
the library has not been built: there is no CPU fallback anywhere in the
product path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgravity.so")


class LibraryMissing(RuntimeError):
    """libgravity.so is not built (run __graft_entry__.build())."""


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Float4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class WorldBox(C.Structure):
    """types.h:25-34."""
    _fields_ = [("minX", C.c_float), ("minY", C.c_float), ("minZ", C.c_float),
                ("extent", C.c_float), ("scale", C.c_float)]


class Node(C.Structure):
    """types.h:52-62 (80 bytes; 8 bytes of padding before massData)."""
    _fields_ = [("nodePathFromRoot", C.c_uint64), ("parentIndex", C.c_int),
                ("firstParticleIndex", C.c_int), ("particleCount", C.c_int),
                ("child", C.c_int * 8), ("treeLevel", C.c_int), ("_pad", C.c_int * 2),
                ("massData", C.c_float * 4)]


class TreeState(C.Structure):
    """types.h:75-90 (304 bytes)."""
    _fields_ = [("nodes", C.c_void_p), ("maxNodes", C.c_int),
                ("mortonIn", C.c_void_p), ("mortonOut", C.c_void_p),
                ("indexIn", C.c_void_p), ("indexOut", C.c_void_p),
                ("boundingData", C.c_void_p * 4), ("sorter", C.c_void_p),
                ("builder", C.c_void_p), ("levelOffsets", C.c_int * 22),
                ("levelCount", C.c_int * 22), ("deepestLevel", C.c_int),
                ("treeBlocks", C.c_int), ("starCount", C.c_int),
                ("sortedPosMassVals", C.c_void_p), ("sortedVelVals", C.c_void_p)]


class WalkStats(C.Structure):
    _fields_ = [("interactions", C.c_ulonglong), ("opened", C.c_ulonglong),
                ("overflowGroups", C.c_ulonglong), ("cellInteractions", C.c_ulonglong),
                ("bodyGroups", C.c_ulonglong), ("networkEvals", C.c_ulonglong),
                ("farCells", C.c_ulonglong), ("farInteractions", C.c_ulonglong)]


assert C.sizeof(Node) == 80 and C.sizeof(TreeState) == 304 and C.sizeof(WorldBox) == 20

vp = C.c_void_p
vpp = C.POINTER(C.c_void_p)
i32 = C.c_int
f32 = C.c_float
f64 = C.c_double

# name: (restype, argtypes)
_PROTOS = {
    # reference API: sorting.h
    "sorterCreate": (vp, [i32]),
    "radixSortPairs": (None, [vp, vp, vp, vp, vp, i32]),
    "sorterDestroyer": (None, [vp]),
    "reorderParticles": (None, [vp, vp, vp, vp, vp, i32]),
    # reference API: tree.h
    "mortonEncode": (None, [vp, vp, vp, i32, WorldBox]),
    "localMinMaxFinder": (None, [vp, C.c_uint, vp, vp]),
    "globalMinMaxReducer": (None, [vp, vp, i32, vp, vp]),
    "identifyNodesAtLevel": (None, [vp, i32, i32, vp]),
    "createNodes": (None, [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32]),
    "updateTreeParticles": (None, [vp, i32, i32, i32]),
    "updateTreeMass": (None, [vp, vp, i32, i32, i32]),
    "calculateAcceleration": (None, [vp, TreeState, vp, f32, f32, f32, i32]),
    "positionUpdater": (None, [vp, vp, vp, i32, f32]),
    "sumCreate": (vp, [i32]),
    "prefixSum": (None, [vp, vp, vp, i32]),
    "sumDestroyer": (None, [vp]),
    # reference API: basicGravity.cu host functions
    "initialiseTree": (None, [C.POINTER(TreeState), i32, i32, i32]),
    "createWorldBox": (None, [vp, i32, vp, vp, vp, vp, C.POINTER(WorldBox)]),
    "buildTree": (None, [C.POINTER(TreeState), vpp, vpp]),
    "roundNumberToMultiple": (i32, [i32, i32]),
    "blockNumCalculator": (None, [C.POINTER(i32)] * 6 + [i32]),
    # engine extensions: gravity_ext.h
    "gd_set_exit_on_error": (None, [i32]),
    "gd_last_error": (C.c_char_p, []),
    "gd_clear_error": (None, []),
    "gd_malloc": (vp, [C.c_size_t]),
    "gd_free": (None, [vp]),
    "gd_memcpy_h2d": (None, [vp, vp, C.c_size_t]),
    "gd_memcpy_d2h": (None, [vp, vp, C.c_size_t]),
    "gd_memcpy_d2d": (None, [vp, vp, C.c_size_t]),
    "gd_memset": (None, [vp, i32, C.c_size_t]),
    "gd_synchronize": (None, []),
    "gd_device_count": (i32, []),
    "gd_device_name": (None, [C.c_char_p, i32]),
    "gd_tree_summary": (None, [C.POINTER(TreeState), C.POINTER(i32), C.POINTER(i32),
                                C.POINTER(i32)]),
    "gd_build_tree_async": (None, [C.POINTER(TreeState), vpp, vpp, vpp, vpp]),
    "gd_tree_levels": (None, [C.POINTER(TreeState)]),
    "gd_tree_root_fp64": (None, [C.POINTER(TreeState), C.POINTER(C.c_double)]),
    "gd_set_quadrupole": (None, [C.POINTER(TreeState), i32]),
    "gd_set_node_records": (None, [C.POINTER(TreeState), i32]),
    "gd_write_node_records": (None, [C.POINTER(TreeState)]),
    "gd_build_timing": (None, [i32]),
    "gd_build_phase_ms": (i32, [vp]),
    "gd_tree_walk_arrays": (None, [C.POINTER(TreeState), vp, vp, vp, i32]),
    "gd_tree_mass_radii": (None, [C.POINTER(TreeState), vp, i32]),
    "gd_tree_third_moments": (None, [C.POINTER(TreeState), vp, i32]),
    "gd_walk": (None, [C.POINTER(TreeState), vp, vp, vp, f32, f32, f32, i32,
                       C.POINTER(WalkStats)]),
    "gd_walk_timing": (None, [i32]),
    "gd_walk_set_list_capacity": (i32, [i32, i32]),
    "gd_walk_set_group_split": (f32, [f32]),
    "gd_walk_set_far_field": (f32, [f32]),
    "gd_walk_set_opening_rule": (i32, [i32]),
    "gd_walk_set_async": (i32, [i32]),
    "gd_walk_poll": (i32, []),
    "gd_walk_kernel_ms": (f32, []),
    "gd_walk_phase_ms": (i32, [vp]),
    "gd_mlp_create": (vp, [vp]),
    "gd_mlp_destroy": (None, [vp]),
    "gd_mlp_get_params": (None, [vp, vp]),
    "gd_mlp_set_params": (None, [vp, vp]),
    "gd_mlp_forward": (None, [vp, vp, vp, C.c_longlong]),
    "gd_mlp_gradient": (f32, [vp, vp, vp, vp, C.c_longlong, vp]),
    "gd_mlp_train_step": (f32, [vp, vp, vp, vp, C.c_longlong, f32]),
    "gd_mlp_adam_state": (None, [vp, vp, vp]),
    "gd_corrector_samples": (C.c_longlong, [C.POINTER(TreeState), vp, vp, i32, f32, f32, f32,
                                             vp, vp,
                                             C.c_longlong, C.POINTER(i32)]),
    "gd_walk_learned": (None, [C.POINTER(TreeState), vp, vp, vp, f32, f32, f32, i32, vp, f32,
                               C.POINTER(WalkStats)]),
    "gd_integrate_hilo": (None, [vp, vp, vp, vp, i32, f32]),
    "gd_direct_sum_fp64": (None, [vp, vp, i32, vp, i32, f64, f64, vp]),
    "gd_direct_sum_fp64_noalloc": (None, [vp, vp, i32, vp, i32, f64, f64, vp]),
    "gd_force_errors": (None, [vp, vp, i32, vp, vp]),
    "gd_diagnostics": (None, [vp, vp, vp, i32, C.POINTER(C.c_double)]),
    "gd_stream_copy": (None, [vp, vp, C.c_size_t]),
    "gd_graph_begin": (None, []),
    "gd_graph_end": (vp, []),
    "gd_graph_launch": (None, [vp]),
    "gd_graph_destroy": (None, [vp]),
    "gd_event_create": (vp, []),
    "gd_event_record": (None, [vp]),
    "gd_event_elapsed_ms": (f32, [vp, vp]),
    "gd_event_destroy": (None, [vp]),
}

# Optional symbols (present once the corresponding engine part is built).
_OPTIONAL = {}

_lib = None


def exported_symbols():
    """Names every header in include/ declares (checked by the CPU tests)."""
    return sorted(_PROTOS)


def load_library(path=LIB_PATH):
    """Load libgravity.so and attach prototypes; raises LibraryMissing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in _OPTIONAL.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    return lib


class GravityError(RuntimeError):
    pass


def check(lib=None):
    """Raise GravityError if the library recorded an error (exit-on-error off)."""
    lib = lib or load_library()
    msg = lib.gd_last_error()
    if msg:
        lib.gd_clear_error()
        raise GravityError(msg.decode())
