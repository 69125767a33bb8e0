"""The airfoil application entry point (include/op2amd_airfoil.h)."""
import ctypes as C

import numpy as np

from ._lib import lib
from .api import PLAN_NAMES
from .mesh import op2amd_mesh

PLAIN, TUNED, GRAPHED = 0, 1, 2
MODES = {"plain": PLAIN, "tuned": TUNED, "graphed": GRAPHED}
LOOPS = ("save_soln", "adt_calc", "res_calc", "bres_calc", "update")


class Opts(C.Structure):
    _fields_ = [("niter", C.c_int), ("warmup", C.c_int), ("alpha_mode", C.c_int),
                ("mode", C.c_int), ("soa", C.c_int), ("plan_res", C.c_int),
                ("block_res", C.c_int), ("plan_bres", C.c_int),
                ("block_bres", C.c_int), ("block_direct", C.c_int),
                ("block_adt", C.c_int), ("kernel_timing", C.c_int),
                ("plan_on_device", C.c_int), ("renumber", C.c_int),
                ("mixed", C.c_int), ("autotune", C.c_int), ("graph_steps", C.c_int),
                ("dump", C.c_char_p)]


class Result(C.Structure):
    _fields_ = [("rms", C.POINTER(C.c_double)), ("rms_k0", C.POINTER(C.c_double)),
                ("maxerr", C.POINTER(C.c_double)), ("errloc", C.POINTER(C.c_int)),
                ("q_out", C.POINTER(C.c_double)), ("seconds", C.c_double),
                ("setup_seconds", C.c_double), ("loop_kernel_ms", C.c_double * 5),
                ("loop_calls", C.c_int * 5), ("loop_bytes", C.c_double * 5),
                ("chosen_plan", C.c_int * 5), ("chosen_block", C.c_int * 5),
                ("eager_iterations", C.c_int), ("graph_steps", C.c_int)]


lib.op2amd_airfoil_run.restype = C.c_int
lib.op2amd_airfoil_run.argtypes = [C.POINTER(op2amd_mesh), C.POINTER(Opts),
                                   C.POINTER(Result)]


def run(mesh, niter, mode="plain", warmup=0, alpha_mode=1, soa=False,
        plan_res="auto", block_res=0, plan_bres="auto", block_bres=0,
        block_direct=0, block_adt=0, kernel_timing=False, plan_on_device=False,
        renumber=False, mixed=False, autotune=0, graph_steps=1, dump=None):
    """Run the airfoil app; returns a dict of numpy arrays and timings.  With
    `dump` (a path prefix) the final q is also written to <dump>.dat and
    <dump>.bin, and a renumbered run's cell permutation to <dump>.perm.
    renumber: False; True (op_renumber(pecell), the application sees the new
    numbering); "transparent" (the device alone works in that numbering)."""
    o = Opts(niter, warmup, alpha_mode, MODES[mode], int(soa),
             PLAN_NAMES[plan_res], block_res, PLAN_NAMES[plan_bres], block_bres,
             block_direct, block_adt, int(kernel_timing), int(plan_on_device),
             {False: 0, True: 1, "transparent": 2}[renumber], int(mixed), int(autotune),
             int(graph_steps),
             None if dump is None else str(dump).encode())
    rms = np.zeros(niter)
    rms_k0 = np.zeros(niter)
    maxerr = np.zeros(niter)
    errloc = np.zeros(niter, np.int32)
    q = np.zeros((mesh.ncell, 4))
    r = Result()
    r.rms = rms.ctypes.data_as(C.POINTER(C.c_double))
    r.rms_k0 = rms_k0.ctypes.data_as(C.POINTER(C.c_double))
    r.maxerr = maxerr.ctypes.data_as(C.POINTER(C.c_double))
    r.errloc = errloc.ctypes.data_as(C.POINTER(C.c_int))
    r.q_out = q.ctypes.data_as(C.POINTER(C.c_double))
    rc = lib.op2amd_airfoil_run(C.byref(mesh.c()), C.byref(o), C.byref(r))
    if rc:
        raise RuntimeError(f"op2amd_airfoil_run failed with code {rc}")
    return {
        "rms": rms, "rms_k0": rms_k0, "maxerr": maxerr, "errloc": errloc, "q": q,
        "seconds": r.seconds, "setup_seconds": r.setup_seconds,
        "loop_kernel_ms": dict(zip(LOOPS, list(r.loop_kernel_ms))),
        "loop_calls": dict(zip(LOOPS, list(r.loop_calls))),
        "loop_bytes": dict(zip(LOOPS, list(r.loop_bytes))),
        "chosen": {lp: (r.chosen_plan[i], r.chosen_block[i]) for i, lp in enumerate(LOOPS)},
        "eager_iterations": r.eager_iterations, "graph_steps": r.graph_steps,
    }
