"""The C-ABI libraries load and export every function their headers declare
(no device calls: this runs without a GPU)."""
import ctypes
import pathlib
import re

import pytest

from conftest import REPO

HEADERS = ["op2_hip.h", "op2_hip_airfoil.h", "op_profile.h", "op2_hip_rt.h", "op2amd_mesh.h", "op2amd_airfoil.h",
           "op2amd_ensemble.h", "op2amd_synth.h", "op2amd_costmodel.h"]


def declared_functions(header):
    text = (REPO / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    # C declarations inside extern "C": "<type> <name>(...);"
    c_part = text.split('extern "C" {', 1)[1] if 'extern "C" {' in text else text
    c_part = c_part.split("#ifdef __cplusplus\n}", 1)[0]
    c_part = c_part.split("} /* extern", 1)[0]
    c_part = re.sub(r"typedef[^;]*;", "", c_part)  # function-pointer typedefs
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", c_part)
    return sorted(set(n for n in names if n not in ("if", "return", "sizeof")))


@pytest.mark.parametrize("header", HEADERS)
def test_library_exports_header(header):
    lib = ctypes.CDLL(str(REPO / "op2-common_amd" / "lib" / "libop2hip.so"))
    names = declared_functions(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_drop_in_symbols_present():
    """The symbols a translated airfoil app and its generated stubs bind
    (op_lib_c.h:75-121, the five op_par_loop_airfoil_* stubs)."""
    lib = ctypes.CDLL(str(REPO / "op2-common_amd" / "lib" / "libop2hip.so"))
    for n in ["op_init", "op_exit", "op_decl_set", "op_decl_map", "op_decl_dat_char",
              "op_decl_const_char", "op_arg_dat", "op_arg_gbl_char", "op_arg_idx",
              "op_arg_info_char", "op_fetch_data_char", "op_get_size",
              "op_print_dat_to_binfile", "op_timing_output", "op_mpi_halo_exchanges",
              "op_malloc", "op_calloc", "op_realloc", "op_free", "op_fetch_data_idx_char",
              "op_profile_start", "op_profile_end", "op_profile_output",
              "op_par_loop_airfoil_1_save_soln", "op_par_loop_airfoil_2_adt_calc",
              "op_par_loop_airfoil_3_res_calc", "op_par_loop_airfoil_4_bres_calc",
              "op_par_loop_airfoil_5_update"]:
        assert hasattr(lib, n), n
    for n in ["OP_diags", "OP_block_size", "OP_part_size", "OP_kernels", "OP_auto_soa"]:
        ctypes.c_int.in_dll(lib, n) if n != "OP_kernels" else ctypes.c_void_p.in_dll(lib, n)


def test_struct_layouts_match_reference():
    """op_arg is passed by value between app, stubs and runtime: its size and
    field offsets follow op_lib_core.h:134-199 on x86-64 (sizes and offsets
    measured with g++ offsetof on the reference header)."""
    import op2amd.api as api
    assert ctypes.sizeof(api.op_arg) == 96
    assert api.op_arg.dat.offset == 8 and api.op_arg.acc.offset == 80
    assert api.op_arg.opt.offset == 92
    assert ctypes.sizeof(api.op_set_core) == 32
    assert ctypes.sizeof(api.op_dat_core) == 96 and api.op_dat_core.dirty_hd.offset == 76
    assert ctypes.sizeof(api.op_map_core) == 72


def test_oracle_library_loads():
    from oracle import oracle
    c = oracle.consts(1)
    assert c.gam == float.fromhex("0x1.6666660000000p+0")  # 1.4f as double
