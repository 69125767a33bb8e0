"""The op_profile_* timing tree (op2/include/op_profile.h:15-38,197-212;
op2/src/externlib/op_profile.cpp:214-572) and the aligned host allocators
(op_lib_core.cpp:1388-1420), through the C ABI without a GPU: before any loop
the runtime has no device to synchronise, so the tree runs on the host alone.
The tree is a process-wide singleton, so each case runs in a fresh
interpreter."""
import ctypes
import json
import os
import subprocess
import sys
import textwrap

import pytest

from conftest import REPO

LIB = REPO / "op2-common_amd" / "lib" / "libop2hip.so"

SCRIPT = textwrap.dedent("""
    import ctypes, sys, time
    lib = ctypes.CDLL(sys.argv[1])
    s = lambda x: x.encode()
    lib.op_profile_start(s("App"))
    for it in range(3):
        lib.op_profile_enter(s("Outer"))
        lib.op_profile_enter_kernel(s("res_calc"), s("hip"), s("owner"))
        lib.op_profile_enter(s("Computation"))
        time.sleep(0.002)
        lib.op_profile_exit()
        lib.op_profile_exit()
        lib.op_profile_next(s("Second"))
        time.sleep(0.001)
        lib.op_profile_exit()
    lib.op_profile_enter(s("Left open"))
    lib.op_profile_end()
    lib.op_profile_output()
    lib.op_profile_output_json(s(sys.argv[2]))
""")


def run_tree(tmp_path, level):
    env = dict(os.environ)
    env.pop("OP_PROFILE_JSON_OUTPUT", None)
    if level is None:
        env.pop("OP_PROFILE_LEVEL", None)
    else:
        env["OP_PROFILE_LEVEL"] = str(level)
    out = tmp_path / "tree.json"
    r = subprocess.run([sys.executable, "-c", SCRIPT, str(LIB), str(out)], env=env,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    tree = json.loads(out.read_text()) if out.exists() else None
    return r.stdout, tree


def child(node, name):
    hits = [c for c in node["children"] if c["name"] == name]
    assert len(hits) == 1, (name, [c["name"] for c in node["children"]])
    return hits[0]


def check_clock(node):
    c = node["clock"]
    assert c["n"] >= 1 and 0 <= c["min"] <= c["max"] <= c["total"]
    assert c["total"] <= c["n"] * c["max"]
    for ch in node["children"]:
        check_clock(ch)
        # a section's children ran inside it
        assert ch["clock"]["total"] <= c["total"] + 1000


def test_default_level_times_sections_only(tmp_path):
    out, tree = run_tree(tmp_path, None)
    assert tree["name"] == "App" and tree["type"] == 0 and tree["num_ranks"] == 1
    assert tree["clock"]["n"] == 1
    outer = child(tree, "Outer")
    assert outer["clock"]["n"] == 3 and outer["type"] == 0
    # level 1: the loop node and everything under it are not timed
    assert outer["children"] == []
    second = child(tree, "Second")
    assert second["clock"]["n"] == 3
    assert second["clock"]["min"] >= 1_000_000  # slept 1 ms, nanoseconds
    assert child(tree, "Left open")["clock"]["n"] == 1
    check_clock(tree)
    assert "Outer *3 total:" in out


def test_level_two_adds_loop_nodes_without_their_sections(tmp_path):
    out, tree = run_tree(tmp_path, 2)
    loop = child(child(tree, "Outer"), "res_calc/hip/owner")
    assert loop["type"] == 1 and loop["clock"]["n"] == 3
    assert loop["children"] == []  # sections inside a loop need level 3
    assert loop["clock"]["min"] >= 2_000_000
    check_clock(tree)
    # the loop table under its section, with the reference's columns
    assert "App/Outer" in out and "num    total      avg      min      max" in out
    assert "%kern" not in out


def test_level_three_times_sections_inside_loops(tmp_path):
    out, tree = run_tree(tmp_path, 3)
    loop = child(child(tree, "Outer"), "res_calc/hip/owner")
    comp = child(loop, "Computation")
    assert comp["clock"]["n"] == 3 and comp["clock"]["min"] >= 2_000_000
    check_clock(tree)
    assert "%kern" in out


def test_level_zero_disables(tmp_path):
    out, tree = run_tree(tmp_path, 0)
    assert tree is None and out == ""


def test_unsupported_level_warns_and_keeps_default(tmp_path):
    out, tree = run_tree(tmp_path, "7")
    assert "warning: OP_PROFILE_LEVEL set to unsupported value: 7" in out
    assert child(tree, "Outer")["children"] == []


def test_json_output_variable(tmp_path):
    env = dict(os.environ, OP_PROFILE_LEVEL="2", OP_PROFILE_JSON_OUTPUT=str(tmp_path / "v.json"))
    r = subprocess.run([sys.executable, "-c", SCRIPT, str(LIB), str(tmp_path / "x.json")],
                       env=env, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert f"Writing timing JSON to: {tmp_path / 'v.json'}" in r.stdout
    assert json.loads((tmp_path / "v.json").read_text())["name"] == "App"


@pytest.mark.parametrize("size", [0, 1, 63, 64, 65, 1000, 1 << 20])
def test_host_allocators_alignment(size):
    lib = ctypes.CDLL(str(LIB))
    lib.op_malloc.restype = lib.op_calloc.restype = lib.op_realloc.restype = ctypes.c_void_p
    lib.op_malloc.argtypes = [ctypes.c_size_t]
    lib.op_calloc.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
    lib.op_realloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    lib.op_free.argtypes = [ctypes.c_void_p]
    p = lib.op_malloc(size)
    if size:
        assert p and p % 64 == 0
        ctypes.memset(p, 0xAB, size)
    lib.op_free(p)
    c = lib.op_calloc(size + 3, 8)
    assert c and c % 64 == 0
    assert ctypes.string_at(c, (size + 3) * 8) == b"\0" * ((size + 3) * 8)
    ctypes.memset(c, 7, (size + 3) * 8)
    r = lib.op_realloc(c, (size + 3) * 8 * 3)
    assert r and r % 64 == 0
    assert ctypes.string_at(r, (size + 3) * 8) == b"\7" * ((size + 3) * 8)
    lib.op_free(r)
