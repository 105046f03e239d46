"""CPU tests of the drop-in boundary: the C ABI library loads and exports every
entry point include/curtree/curtree_c.h declares, the ctypes mirror's structs
have the C layout, and the C++ drop-in header compiles against the reference's
template API (no compute calls: no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "curtree", "curtree_c.h")


def declared_entry_points():
    with open(HDR) as f:
        text = f.read()
    names = re.findall(r"CURTREE_API\s+[\w\s\*]+?\b(curtree_\w+)\s*\(", text)
    assert len(names) > 20
    return names


def test_library_exports_every_declared_symbol():
    import curtree_amd

    missing = [n for n in declared_entry_points() if not hasattr(curtree_amd.lib, n)]
    assert not missing, missing


def test_struct_layouts_match_c_header(tmp_path):
    import curtree_amd as ct

    src = tmp_path / "lay.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "curtree/curtree_c.h"\n'
        "int main(void){\n"
        'printf("%zu %zu %zu %zu\\n", sizeof(curtree_bbox_f32), offsetof(curtree_bbox_f32,id),'
        " sizeof(curtree_rtree_f32), offsetof(curtree_rtree_f32,len));\n"
        'printf("%zu %zu\\n", sizeof(curtree_idpair64), offsetof(curtree_idpair64,tid));\n'
        "return 0;}\n")
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = [ctypes.sizeof(ct.BBox), ct.BBox.id.offset, ctypes.sizeof(ct.RTreeStruct),
           ct.RTreeStruct.len.offset, ctypes.sizeof(ct.IdPair64), ct.IdPair64.tid.offset]
    assert [int(x) for x in out] == got


def test_cpp_dropin_header_compiles(tmp_path):
    # code written against the reference's template API compiles unchanged
    src = tmp_path / "dropin.cpp"
    src.write_text(
        '#include "curtree/curtree.hpp"\n'
        "int run(BBOX<float>& d_r_box, BBOX<float>& d_q_box, RTREE<float>& d_rt){\n"
        "  IDPAIR d_dq, h_dq; int r = build_rtree(d_r_box, d_rt);\n"
        "  bfs_frontier(d_rt, d_r_box, d_q_box, d_dq); idpair_d2h(h_dq, d_dq);\n"
        "  bfs_thrust(d_rt, d_r_box, d_q_box, d_dq);\n"
        "  idpair_d_free(d_dq); idpair_h_free(h_dq); rt_d_free(d_rt); box_d_free(d_r_box);\n"
        "  return r; }\n")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I",
                    os.path.join(ROOT, "include"), str(src)], check=True)


def test_error_path_reports_message_without_gpu_work():
    import curtree_amd as ct

    # an invalid fanout is rejected before any device call
    rt = ct.RTree(1)
    b = ct.BBox()
    b.sz = 10
    rc = ct.lib.curtree_build_f32(ctypes.byref(b), ctypes.byref(rt.s), 0, None)
    assert rc == ct.EINVAL
    assert b"fanout" in ct.lib.curtree_last_error()
    rc = ct.lib.curtree_build_f32(ctypes.byref(b), ctypes.byref(rt.s), 7, None)
    assert rc == ct.EINVAL


def test_product_does_not_reference_oracle():
    # the shipped library and package never load the CPU oracle
    pkg = os.path.join(ROOT, "curtree_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                with open(os.path.join(dirpath, fn)) as f:
                    text = f.read()
                assert "liboracle" not in text and "oracle_py" not in text, fn
    so = os.path.join(pkg, "libcurtree_amd.so")
    if os.path.exists(so):
        out = subprocess.run(["nm", "-D", so], capture_output=True, text=True).stdout
        assert " or_" not in out


@pytest.mark.parametrize("bad", [-1])
def test_negative_sizes_rejected(bad):
    import curtree_amd as ct

    rt = ct.RTree(16)
    b = ct.BBox()
    b.sz = bad
    assert ct.lib.curtree_build_f32(ctypes.byref(b), ctypes.byref(rt.s), 0, None) == ct.EINVAL
