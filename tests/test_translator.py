"""op2amd.translator on the OP2 programs under tests/apps: the program model
it reads (loops, argument kinds, constants), the rewritten application and the
generated HIP code, which must compile for gfx950."""
import os
import subprocess

import pytest

from conftest import REPO
from op2amd.translator import op, source, translate
from op2amd.translator.op import ArgDat, ArgGbl, ArgIdx, ArgInfo

APPS = REPO / "tests" / "apps"


def test_source_helpers():
    text = 'f(a, g(b, "x,y"), {1, 2}) // f(z)\n/* f(q) */ f(c)'
    clean = source.strip_comments(text)
    assert len(clean) == len(text)
    calls = source.find_calls(clean, "f")
    assert [c[2] for c in calls] == [["a", 'g(b, "x,y")', "{1, 2}"], ["c"]]
    assert source.int_constants("#define NN 12\nconst int M = 3;\n") == {"NN": 12, "M": 3}
    fn = "static inline void k(const double *a,\n    double *b) {\n  b[0] = a[0];\n}\n"
    s, e = source.find_function("int x;\n" + fn + "k(p, q);", "k")
    assert ("int x;\n" + fn)[s:e].startswith("static inline void k(")


def test_airfoil_program_model():
    prog = op.parse_program(str(APPS / "airfoil" / "airfoil_tr.cpp"))
    assert [l.kernel for l in prog.loops] == ["save_soln", "adt_calc", "res_calc",
                                              "bres_calc", "update"]
    assert [l.stub for l in prog.loops][2] == "op_par_loop_airfoil_tr_3_res_calc"
    res = prog.loops[2]
    assert [a.acc for a in res.args] == ["OP_READ"] * 6 + ["OP_INC"] * 2
    assert all(isinstance(a, ArgDat) and a.map == "pecell" for a in res.args[2:])
    upd = prog.loops[4]
    assert isinstance(upd.args[4], ArgGbl) and upd.args[4].acc == "OP_INC"
    assert isinstance(upd.args[6], ArgIdx) and upd.args[6].map is None
    assert isinstance(upd.args[7], ArgInfo) and upd.args[7].ref == 5
    assert [(c.ptr, c.dim) for c in prog.consts] == [
        ("gam", 1), ("gm1", 1), ("cfl", 1), ("eps", 1), ("mach", 1), ("alpha", 1), ("qinf", 4)]
    assert set(prog.kernel_text) == {"save_soln", "adt_calc", "res_calc", "bres_calc", "update"}


def test_rewritten_application(tmp_path):
    paths = translate(str(APPS / "functional" / "consts.cpp"), str(tmp_path))
    app = open(paths["app"]).read()
    assert '#include "op2_hip.h"' in app and '#include "consts_op2gen.h"' in app
    assert "op_par_loop(" not in source.strip_comments(app)
    assert app.count("op_par_loop_consts_1_use_scalar(") == 1
    assert app.count("op_par_loop_consts_3_use_scalar(") == 1
    assert app.count('op_decl_const2("scale", ') == 2
    hdr = open(paths["header"]).read()
    assert "void op_par_loop_consts_2_use_array(char const *, op_set, op_arg);" in hdr


def test_unsupported_forms_are_refused(tmp_path):
    bad = tmp_path / "bad.cpp"
    bad.write_text('#include "op_seq.h"\nvoid k(double *a) {}\nint main() {\n'
                   '  op_par_loop(k, "k", s, op_arg_dat(d, -1, OP_ID, n, "double", OP_RW));\n}\n')
    with pytest.raises(op.OpError, match="dim must be an integer constant"):
        op.parse_program(str(bad))
    bad.write_text('#include "op_seq.h"\nint main() {\n'
                   '  op_par_loop(k, "k", s, op_arg_dat(d, -1, OP_ID, 1, "double", OP_RW));\n}\n')
    with pytest.raises(op.OpError, match="user kernel k not found"):
        op.parse_program(str(bad))


def test_called_functions_go_to_the_device_callees_first():
    """Functions a kernel calls (from the file or a quoted include, at any
    depth) are copied to the device once each, before their callers; calls
    of functions the program does not define (std:: maths, members) are
    not followed."""
    prog = op.parse_program(str(APPS / "functional" / "helpers.cpp"))
    assert list(prog.device_text) == ["square", "cube", "poly", "k_poly", "k_squares"]
    assert op._called_names("void f(S s) { s.g(1); p->h(2); std::sqrt(3.0); k(4); if (x) {} }") \
        == ["k"]


@pytest.mark.parametrize("name", ["gbl", "airfoil_tr", "helpers"])
def test_generated_kernels_compile_for_gfx950(name, tmp_path):
    src = {"gbl": APPS / "functional" / "gbl.cpp", "airfoil_tr": APPS / "airfoil" / "airfoil_tr.cpp",
           "helpers": APPS / "functional" / "helpers.cpp"}[name]
    paths = translate(str(src), str(tmp_path))
    cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "--offload-arch=gfx950", "-w",
           "-I" + str(REPO / "include"), "-I" + str(src.parent), "-I" + str(tmp_path),
           "-x", "hip", "-c", paths["kernels"], "-o", str(tmp_path / "k.o")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_generated_airfoil_kernel_shapes(tmp_path):
    """The shapes the generated kernels rely on: args on one map read its rows
    through the first such arg (one pointer, so a row loads at once); the
    owner kernel walks its instances in chunks of blockDim with LDS slots per
    chunk and reads instance rows through op2gen_row (16-bit where a block
    allows); every element kernel takes XCD-ordered blocks; update's partial
    records come from one barrier per block."""
    paths = translate(str(APPS / "airfoil" / "airfoil_tr.cpp"), str(tmp_path))
    k = open(paths["kernels"]).read()
    # adt_calc: the four x args through pcell share A.m0
    for i in range(4):
        assert f"const int t{i} = A.m0[(size_t)n * A.md0 + {i}];" in k
    # res_calc owner kernel: pedge rows through arg 0, pecell rows through arg 2
    owner = k[k.index("op2gen_airfoil_tr_3_owner(") :]
    owner = owner[: owner.index("\n}\n")]
    assert "op2gen_row(A.r0, A.h0, hb0, (size_t)(e0 + j) * A.md0 + 1)" in owner
    for i in (3, 4, 5, 6, 7):
        assert f"const int t{i} = op2gen_row(A.r2, A.h2, hb2," in owner
    assert "for (int base = 0; base < ne; base += blockDim.x)" in owner
    assert "lds[(size_t)k * lds_s + threadIdx.x * 2 + 1] = b7[k];" in owner
    assert owner.count("__syncthreads();") == 2 and "op2gen_xcd_block()" in owner
    upd = k[k.index("op2gen_airfoil_tr_5_kernel(") :]
    upd = upd[: upd.index("\n}\n")]
    assert upd.count("__syncthreads();") == 1 and "op2gen_xcd_block()" in upd
    stub = open(paths["kernels"]).read()
    assert "op2gen_rows16, op->nblocks, 256" in stub
