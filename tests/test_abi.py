"""The C-ABI library loads and exports what include/*.h declares (CPU only).

No compute call is made here: only host-side helpers and the argument
checks that fail before any GPU work (the reference's error paths,
sorting.cu:16-20,69-76 and tree.cu:242-246,287-295).
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from cudagravitydemo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, re.M):
            name = m.group(1)
            if name not in ("sizeof", "offsetof"):
                names.add(name)
    return names


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_every_declared_function_is_exported():
    decl = declared_functions()
    assert {"sorterCreate", "radixSortPairs", "buildTree", "calculateAcceleration",
            "positionUpdater", "gd_walk"} <= decl
    missing = decl - exported()
    assert not missing, missing
    # the units.c globals are exported too
    assert {"G", "theta"} <= exported()


def test_binding_covers_every_declaration():
    assert declared_functions() <= set(_lib.exported_symbols())


def test_library_loads_with_prototypes():
    lib = _lib.load_library()
    assert lib.roundNumberToMultiple(100, 32) == 128
    assert lib.roundNumberToMultiple(96, 32) == 96
    assert lib.roundNumberToMultiple(5, 0) == 5
    assert C.c_float.in_dll(lib, "G").value == 1.0
    assert abs(C.c_float.in_dll(lib, "theta").value - 0.7) < 1e-7


def test_block_num_calculator_matches_reference_geometry():
    lib = _lib.load_library()
    vals = [C.c_int() for _ in range(6)]
    lib.blockNumCalculator(*[C.byref(v) for v in vals], 100000)
    acc, tree, integ, padded, ithr, tthr = (v.value for v in vals)
    # basicGravity.cu:389-413: 100000 % 32 == 0, 100000 % 256 != 0
    assert padded == 100000 and acc == 3125
    assert integ == 100000 // 256 + 1 and tree == integ and ithr == 256 and tthr == 256


def test_argument_errors_are_reported_without_gpu_work():
    lib = _lib.load_library()
    lib.gd_set_exit_on_error(0)
    try:
        assert lib.sorterCreate(0) is None
        assert b"<= 0" in lib.gd_last_error()
        lib.gd_clear_error()
        assert lib.sumCreate(-5) is None
        assert b"<= 0" in lib.gd_last_error()
        lib.gd_clear_error()
        lib.radixSortPairs(None, None, None, None, None, 10)
        assert b"uninitialised" in lib.gd_last_error()
        lib.gd_clear_error()
        lib.prefixSum(None, None, None, 10)
        assert b"uninitialised" in lib.gd_last_error()
        lib.gd_clear_error()
        lib.sorterDestroyer(None)
        assert b"non-existant" in lib.gd_last_error()
        lib.gd_clear_error()
        with pytest.raises(_lib.GravityError):
            lib.sumDestroyer(None)
            _lib.check(lib)
    finally:
        lib.gd_clear_error()
        lib.gd_set_exit_on_error(1)


def test_exit_on_error_is_the_default():
    # a child process: the reference convention prints and exits with -1
    code = ("import sys; sys.path.insert(0, %r); from cudagravitydemo_amd import _lib; "
            "lib = _lib.load_library(); lib.sorterCreate(0)") % ROOT
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 255
    assert "Attempted to create a radix sort" in r.stdout


def test_public_headers_compile_as_c_and_cxx(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "gravity_host.h"\n#include "gravity_ext.h"\n'
                   "int main(void) { treeState s; node n; (void)s; (void)n; return 0; }\n")
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", inc, "-c", str(src), "-o",
                    str(tmp_path / "t.o")], check=True)
    subprocess.run(["g++", "-x", "c++", "-std=c++17", "-Wall", "-Werror", "-I", inc, "-c",
                    str(src), "-o", str(tmp_path / "t2.o")], check=True)


def test_walk_settings_are_checked_on_the_host():
    """gd_walk_set_group_split, gd_walk_set_list_capacity,
    gd_walk_set_far_field and gd_walk_set_opening_rule return the previous
    setting and reject values out of range without touching the GPU."""
    lib = _lib.load_library()
    lib.gd_set_exit_on_error(0)
    try:
        assert lib.gd_walk_set_group_split(0.3) == -1.0          # default: the theta rule
        assert abs(lib.gd_walk_set_group_split(0.0) - 0.3) < 1e-7
        assert lib.gd_walk_set_group_split(-5.0) == 0.0          # any negative: the rule again
        assert lib.gd_walk_set_group_split(1.5) == -1.0
        assert b"above 1" in lib.gd_last_error()
        lib.gd_clear_error()
        lib.gd_walk_set_group_split(float("nan"))
        assert b"above 1" in lib.gd_last_error()
        lib.gd_clear_error()
        assert lib.gd_walk_set_group_split(-1.0) == -1.0         # the rejected values left no trace
        assert lib.gd_walk_set_list_capacity(-1, 0) == 0
        assert b"out of range" in lib.gd_last_error()
        lib.gd_clear_error()
        assert lib.gd_walk_set_list_capacity(0, 1 << 27) == 0
        assert b"out of range" in lib.gd_last_error()
        lib.gd_clear_error()
        assert lib.gd_walk_set_far_field(0.25) == 0.0           # default: off
        assert lib.gd_walk_set_far_field(1.0) == 0.25
        assert b"outside [0, 1)" in lib.gd_last_error()
        lib.gd_clear_error()
        lib.gd_walk_set_far_field(float("nan"))
        assert b"outside [0, 1)" in lib.gd_last_error()
        lib.gd_clear_error()
        assert lib.gd_walk_set_far_field(-0.1) == 0.25
        assert lib.gd_walk_set_far_field(0.0) == 0.25            # the rejected values left no trace
        lib.gd_clear_error()
        assert lib.gd_walk_set_opening_rule(1) == 0              # default: the cell-size rule
        assert lib.gd_walk_set_opening_rule(2) == -1
        assert b"neither 0" in lib.gd_last_error()
        lib.gd_clear_error()
        assert lib.gd_walk_set_opening_rule(-1) == -1
        lib.gd_clear_error()
        assert lib.gd_walk_set_opening_rule(0) == 1              # the rejected values left no trace
    finally:
        lib.gd_clear_error()
        lib.gd_set_exit_on_error(1)
