"""ctypes mirror of the OP2 C API (include/op2_hip.h).

Struct layouts follow op2/include/op_lib_core.h:134-199 of the reference; the
function names, argument meaning and error behaviour are the reference's
(op2/include/op_lib_c.h:75-121).
"""
import ctypes as C

from ._lib import lib

OP_READ, OP_WRITE, OP_RW, OP_INC, OP_MIN, OP_MAX = 0, 1, 2, 3, 4, 5
OP_ARG_GBL, OP_ARG_DAT, OP_ARG_IDX, OP_ARG_INFO = 0, 1, 2, 3

PLAN_AUTO, PLAN_ATOMIC, PLAN_COLOR, PLAN_STAGED, PLAN_GATHER, PLAN_OWNER = 0, 1, 2, 3, 4, 5
PLAN_NAMES = {"auto": 0, "atomic": 1, "color": 2, "staged": 3, "gather": 4, "owner": 5}
SYNC_EVERY_LOOP, SYNC_GBL_ONLY = 0, 1


class op_set_core(C.Structure):
    _fields_ = [("index", C.c_int), ("size", C.c_int), ("name", C.c_char_p),
                ("core_size", C.c_int), ("exec_size", C.c_int),
                ("nonexec_size", C.c_int), ("stride", C.c_int)]


op_set = C.POINTER(op_set_core)


class op_map_core(C.Structure):
    _fields_ = [("index", C.c_int), ("from_", op_set), ("to", op_set),
                ("dim", C.c_int), ("map", C.POINTER(C.c_int)),
                ("map_gbl", C.POINTER(C.c_longlong)),
                ("map_d", C.POINTER(C.c_int)), ("name", C.c_char_p),
                ("user_managed", C.c_int), ("force_part", C.c_bool)]


op_map = C.POINTER(op_map_core)


class op_dat_core(C.Structure):
    _fields_ = [("index", C.c_int), ("set", op_set), ("dim", C.c_int),
                ("size", C.c_int), ("data", C.c_void_p),
                ("data_d", C.c_void_p), ("type", C.c_char_p),
                ("name", C.c_char_p), ("buffer_d", C.c_void_p),
                ("buffer_d_r", C.c_void_p), ("dirtybit", C.c_int),
                ("dirty_hd", C.c_int), ("user_managed", C.c_int),
                ("mpi_buffer", C.c_void_p)]


op_dat = C.POINTER(op_dat_core)


class op_arg(C.Structure):
    _fields_ = [("index", C.c_int), ("dat", op_dat), ("map", op_map),
                ("dim", C.c_int), ("idx", C.c_int), ("size", C.c_int),
                ("data", C.c_void_p), ("data_d", C.c_void_p),
                ("map_data", C.POINTER(C.c_int)),
                ("map_data_d", C.POINTER(C.c_int)), ("type", C.c_char_p),
                ("acc", C.c_int), ("argtype", C.c_int), ("sent", C.c_int),
                ("opt", C.c_int)]


OP_ID = C.cast(None, op_map)


def _fn(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_cstr = C.c_char_p
op_init = _fn("op_init", None, C.c_int, C.POINTER(C.c_char_p), C.c_int)
op_exit = _fn("op_exit", None)
op_decl_set = _fn("op_decl_set", op_set, C.c_int, _cstr)
op_decl_map = _fn("op_decl_map", op_map, op_set, op_set, C.c_int,
                  C.POINTER(C.c_int), _cstr)
op_decl_dat_char = _fn("op_decl_dat_char", op_dat, op_set, C.c_int, _cstr,
                       C.c_int, C.c_void_p, _cstr)
op_decl_const_char = _fn("op_decl_const_char", None, C.c_int, _cstr, C.c_int,
                         C.c_void_p, _cstr)
op_arg_dat = _fn("op_arg_dat", op_arg, op_dat, C.c_int, op_map, C.c_int,
                 _cstr, C.c_int)
op_arg_gbl_char = _fn("op_arg_gbl_char", op_arg, C.c_void_p, C.c_int, _cstr,
                      C.c_int, C.c_int)
op_arg_info_char = _fn("op_arg_info_char", op_arg, C.c_void_p, C.c_int, _cstr,
                       C.c_int, C.c_int)
op_arg_idx = _fn("op_arg_idx", op_arg, C.c_int, op_map)
op_fetch_data_char = _fn("op_fetch_data_char", None, op_dat, C.c_void_p)
op_get_size = _fn("op_get_size", C.c_longlong, op_set)
op2h_set_loop_plan = _fn("op2h_set_loop_plan", C.c_int, _cstr, C.c_int, C.c_int)
op2h_set_sync_mode = _fn("op2h_set_sync_mode", None, C.c_int)
op2h_set_plan_builder = _fn("op2h_set_plan_builder", None, C.c_int)
op2h_device_sync = _fn("op2h_device_sync", None)
op2h_version = _fn("op2h_version", C.c_char_p)
op2h_algorithmic_bytes = _fn("op2h_algorithmic_bytes", C.c_double, op_set,
                             C.c_int, C.POINTER(op_arg))

_loop_args = {
    "op_par_loop_airfoil_1_save_soln": 2,
    "op_par_loop_airfoil_2_adt_calc": 6,
    "op_par_loop_airfoil_3_res_calc": 8,
    "op_par_loop_airfoil_4_bres_calc": 6,
    "op_par_loop_airfoil_5_update": 8,
}
loops = {name: _fn(name, None, _cstr, op_set, *([op_arg] * n))
         for name, n in _loop_args.items()}


class op_staged_plan(C.Structure):
    _fields_ = [("set_size", C.c_int), ("part_size", C.c_int),
                ("nblocks", C.c_int), ("ncolors", C.c_int), ("nmaps", C.c_int),
                ("ncolblk_h", C.POINTER(C.c_int)),
                ("colstart_h", C.POINTER(C.c_int)),
                ("max_nthrcol_h", C.POINTER(C.c_int)),
                ("max_ind_h", C.c_int * 32),
                ("blkmap", C.c_void_p), ("offset", C.c_void_p),
                ("nelems", C.c_void_p), ("nthrcol", C.c_void_p),
                ("thrcol", C.c_void_p), ("ind_offs", C.c_void_p * 32),
                ("ind_sizes", C.c_void_p * 32), ("ind_map", C.c_void_p * 32),
                ("loc_map", C.c_void_p * 32), ("map_of_arg", C.c_int * 32),
                ("loc_of_arg", C.c_int * 32), ("built_on_device", C.c_int),
                ("build_ms", C.c_double)]


class op_color_plan(C.Structure):
    _fields_ = [("set_size", C.c_int), ("ncolors", C.c_int),
                ("col_offsets_h", C.POINTER(C.c_int)),
                ("col_reord", C.c_void_p), ("built_on_device", C.c_int),
                ("build_ms", C.c_double)]


class op_owner_plan(C.Structure):
    _fields_ = [("set_size", C.c_int), ("to_size", C.c_int), ("block_cells", C.c_int),
                ("nblocks", C.c_int), ("ninc", C.c_int), ("total", C.c_int),
                ("max_elems", C.c_int), ("max_owned", C.c_int),
                ("owned_offs", C.c_void_p), ("owned", C.c_void_p),
                ("elem_offs", C.c_void_p), ("elems", C.c_void_p),
                ("csr_offs", C.c_void_p), ("csr", C.c_void_p),
                ("inc_of_arg", C.c_int * 32),
                ("redundancy", C.c_double), ("build_ms", C.c_double),
                ("n_owned", C.c_int), ("max_incs", C.c_int), ("csr4", C.c_void_p)]


op2h_plan_owner_for = _fn("op2h_plan_owner_for", C.POINTER(op_owner_plan),
                          _cstr, op_set, C.c_int, C.c_int, C.POINTER(op_arg))
op2h_owner_plan_download = _fn("op2h_owner_plan_download", C.c_int,
                               C.POINTER(op_owner_plan), C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)
op2h_owner_partition = _fn("op2h_owner_partition", C.c_int, C.c_int, C.c_int, C.c_int,
                           C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int,
                           C.POINTER(C.c_int))


def owner_partition(nto, table, inc_idx, block_cells):
    """Host-only owner partition of targets [0, nto) of a map table; returns
    (number of blocks, block of each target, -1 where unreferenced)."""
    import numpy as np
    t = np.ascontiguousarray(table, dtype=np.int32)
    idx = (C.c_int * len(inc_idx))(*inc_idx)
    out = np.empty(nto, np.int32)
    nb = op2h_owner_partition(nto, t.shape[0], t.shape[1],
                              t.ctypes.data_as(C.POINTER(C.c_int)), len(inc_idx), idx,
                              block_cells, out.ctypes.data_as(C.POINTER(C.c_int)))
    if nb < 0:
        raise ValueError("bad owner partition arguments")
    return nb, out


class op2h_check_access(C.Structure):
    _fields_ = [("table", C.POINTER(C.c_int)), ("dim", C.c_int), ("idx", C.c_int),
                ("to_size", C.c_int), ("dat", C.c_int), ("writes", C.c_int)]


_ip = C.POINTER(C.c_int)


class op2h_staged_plan_host(C.Structure):
    _fields_ = [("set_size", C.c_int), ("nblocks", C.c_int), ("ncolors", C.c_int),
                ("colstart", _ip), ("blkmap", _ip), ("offset", _ip), ("nelems", _ip),
                ("nthrcol", _ip), ("thrcol", C.POINTER(C.c_ubyte)), ("nmaps", C.c_int),
                ("map_table", _ip * 32), ("map_dim", C.c_int * 32),
                ("map_to_size", C.c_int * 32), ("ind_offs", _ip * 32),
                ("ind_sizes", _ip * 32), ("ind_map", _ip * 32), ("nlocs", C.c_int),
                ("loc_map_slot", C.c_int * 32), ("loc_idx", C.c_int * 32),
                ("loc_map", C.POINTER(C.c_ushort) * 32)]


class op2h_owner_plan_host(C.Structure):
    _fields_ = [("set_size", C.c_int), ("to_size", C.c_int), ("nblocks", C.c_int),
                ("ninc", C.c_int), ("map_table", _ip), ("map_dim", C.c_int),
                ("inc_idx", C.c_int * 32), ("owned_offs", _ip), ("owned", _ip),
                ("elem_offs", _ip), ("elems", _ip), ("csr_offs", _ip),
                ("csr", C.POINTER(C.c_ushort)), ("csr4", C.POINTER(C.c_ushort)),
                ("elem_key", _ip), ("instances_by_key", C.c_int)]


op2h_check_staged_plan = _fn("op2h_check_staged_plan", C.c_int,
                             C.POINTER(op2h_staged_plan_host), C.c_int,
                             C.POINTER(op2h_check_access))
op2h_check_color_plan = _fn("op2h_check_color_plan", C.c_int, C.c_int, C.c_int, _ip, _ip,
                            C.c_int, C.POINTER(op2h_check_access))
op2h_check_owner_plan = _fn("op2h_check_owner_plan", C.c_int,
                            C.POINTER(op2h_owner_plan_host))
op2h_check_gather_plan = _fn("op2h_check_gather_plan", C.c_int, C.c_int, C.c_int, C.c_int,
                             _ip, C.c_int, _ip, _ip, _ip, _ip)
op2h_set_old_index = _fn("op2h_set_old_index", C.c_void_p, op_set)
op2h_set_fusion = _fn("op2h_set_fusion", None, C.c_int)
op2h_fused_count = _fn("op2h_fused_count", C.c_ulonglong)
op2h_set_streaming = _fn("op2h_set_streaming", None, C.c_int)
op2h_streaming = _fn("op2h_streaming", C.c_int, C.c_size_t)


op2h_rcm_order = _fn("op2h_rcm_order", C.c_int, C.c_int, C.POINTER(C.c_int),
                     C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int))


def rcm_order(off, adj, on_gpu):
    """Reverse Cuthill-McKee order (new number of each vertex) of a CSR graph
    with strictly ascending neighbour lists, built on the host (on_gpu=0) or
    the device (1; 2 forces the device's two-pass level sort)."""
    import numpy as np
    o = np.ascontiguousarray(off, dtype=np.int32)
    a = np.ascontiguousarray(adj, dtype=np.int32)
    n = len(o) - 1
    out = np.empty(max(n, 0), np.int32)
    ip = C.POINTER(C.c_int)
    rc = op2h_rcm_order(n, o.ctypes.data_as(ip), a.ctypes.data_as(ip) if len(a) else None,
                        int(on_gpu), out.ctypes.data_as(ip))
    if rc != 0:
        raise ValueError("bad RCM arguments")
    return out


op2h_plan_staged_for = _fn("op2h_plan_staged_for", C.POINTER(op_staged_plan),
                           _cstr, op_set, C.c_int, C.c_int, C.POINTER(op_arg))
op2h_plan_color_for = _fn("op2h_plan_color_for", C.POINTER(op_color_plan),
                          _cstr, op_set, C.c_int, C.POINTER(op_arg))
op2h_staged_plan_download = _fn("op2h_staged_plan_download", C.c_int,
                                C.POINTER(op_staged_plan), C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_void_p)
op2h_color_plan_download = _fn("op2h_color_plan_download", C.c_int,
                               C.POINTER(op_color_plan), C.c_void_p)
op2h_plan_gather = _fn("op2h_plan_gather", C.c_void_p, op_map)


def init(*tokens, diags=0):
    """op_init with argv tokens such as 'OP_AUTO_SOA' or 'OP_PART_SIZE=128'."""
    argv = [b"op2amd"] + [t.encode() for t in tokens]
    arr = (C.c_char_p * len(argv))(*argv)
    op_init(len(argv), arr, diags)


def decl_dat(set_, dim, typ, array, name):
    """op_decl_dat from a numpy array (copied by the library)."""
    import numpy as np
    a = np.ascontiguousarray(array)
    return op_decl_dat_char(set_, dim, typ.encode(), a.dtype.itemsize,
                            a.ctypes.data_as(C.c_void_p), name.encode())


def decl_map(from_, to, dim, table, name):
    import numpy as np
    t = np.ascontiguousarray(table, dtype=np.int32)
    return op_decl_map(from_, to, dim, t.ctypes.data_as(C.POINTER(C.c_int)),
                       name.encode())


def arg_dat(dat, idx, map_, dim, typ, acc):
    return op_arg_dat(dat, idx, map_ if map_ is not None else OP_ID, dim,
                      typ.encode(), acc)


def fetch(dat, shape, dtype):
    import numpy as np
    out = np.empty(shape, dtype=dtype)
    op_fetch_data_char(dat, out.ctypes.data_as(C.c_void_p))
    return out
