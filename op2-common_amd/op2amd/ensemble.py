"""Batched airfoil ensembles (include/op2amd_ensemble.h)."""
import ctypes as C

import numpy as np

from ._lib import lib
from .mesh import ORDERINGS


class Spec(C.Structure):
    _fields_ = [("nmesh", C.c_int), ("min_I", C.c_int), ("min_J", C.c_int),
                ("max_I", C.c_int), ("max_J", C.c_int), ("seed", C.c_uint64),
                ("mach_lo", C.c_double), ("mach_hi", C.c_double),
                ("alpha_lo_deg", C.c_double), ("alpha_hi_deg", C.c_double),
                ("ordering", C.c_int)]


class Member(C.Structure):
    _fields_ = [("I", C.c_int), ("J", C.c_int), ("ncell", C.c_int),
                ("mach", C.c_double), ("alpha_deg", C.c_double),
                ("qinf", C.c_double * 4)]


class Opts(C.Structure):
    _fields_ = [("niter", C.c_int), ("warmup", C.c_int), ("batched", C.c_int),
                ("graphed", C.c_int), ("member_timing", C.c_int), ("autotune", C.c_int),
                ("max_workgroups", C.c_int), ("patch_cells", C.c_int),
                ("separate_bres", C.c_int)]


class Result(C.Structure):
    _fields_ = [("rms", C.POINTER(C.c_double)), ("maxerr", C.POINTER(C.c_double)),
                ("errloc", C.POINTER(C.c_int)), ("q_out", C.POINTER(C.c_double)),
                ("member_step_ms", C.POINTER(C.c_double)),
                ("member_busy_ms", C.POINTER(C.c_double)), ("seconds", C.c_double),
                ("setup_seconds", C.c_double), ("total_cells", C.c_longlong),
                ("total_edges", C.c_longlong), ("blocks", C.c_int * 4),
                ("grids", C.c_int * 4), ("dead_res", C.c_int), ("fused_bres", C.c_int)]

LOOPS = ("update_adt", "res_calc", "bres_calc", "update_save_adt")


lib.op2amd_ensemble_describe.restype = C.c_int
lib.op2amd_ensemble_describe.argtypes = [C.POINTER(Spec), C.POINTER(Member)]
lib.op2amd_ensemble_run.restype = C.c_int
lib.op2amd_ensemble_run.argtypes = [C.POINTER(Spec), C.POINTER(Opts), C.POINTER(Result)]


def spec(nmesh=256, min_IJ=(40, 20), max_IJ=(640, 320), seed=0, mach=(0.3, 0.6),
         alpha_deg=(0.0, 5.0), ordering="rcm"):
    return Spec(nmesh, min_IJ[0], min_IJ[1], max_IJ[0], max_IJ[1], seed, mach[0],
                mach[1], alpha_deg[0], alpha_deg[1], ORDERINGS[ordering])


def describe(sp):
    out = (Member * sp.nmesh)()
    if lib.op2amd_ensemble_describe(C.byref(sp), out):
        raise ValueError("bad ensemble spec")
    return [dict(I=m.I, J=m.J, ncell=m.ncell, mach=m.mach, alpha_deg=m.alpha_deg,
                 qinf=list(m.qinf)) for m in out]


def run(sp, niter, warmup=0, batched=True, graphed=True, member_timing=False,
        want_q=False, autotune=False, max_workgroups=0, patch_cells=0, separate_bres=False):
    """Run `niter` time steps of the ensemble `sp` (op2amd_ensemble_run).

    batched: each loop one persistent launch over every member's work items
    (graphed: the step replayed from a HIP graph), else one launch per loop
    per member.  autotune: workgroup sizes picked over the first ten steps.
    member_timing: per-member step times (batched: device-clock stamps of
    every work item; looped: HIP events), in member_step_ms (latency inside
    the step) and member_busy_ms (the member's items' spans summed).
    max_workgroups caps the persistent grids, patch_cells sets res_calc's
    owner patch size, separate_bres keeps bres_calc out of res_calc's launch;
    none of them changes a bit of the results.  Returns per-iteration,
    per-member rms/maxerr/errloc ([niter][nmesh]), q when want_q, timings
    and the launch shapes used."""
    members = describe(sp)
    M = sp.nmesh
    total = sum(m["ncell"] for m in members)
    rms = np.zeros((niter, M))
    maxerr = np.zeros((niter, M))
    errloc = np.zeros((niter, M), np.int32)
    step_ms = np.zeros(M)
    busy_ms = np.zeros(M)
    q = np.zeros((total, 4)) if want_q else None
    r = Result()
    r.rms = rms.ctypes.data_as(C.POINTER(C.c_double))
    r.maxerr = maxerr.ctypes.data_as(C.POINTER(C.c_double))
    r.errloc = errloc.ctypes.data_as(C.POINTER(C.c_int))
    r.q_out = q.ctypes.data_as(C.POINTER(C.c_double)) if want_q else C.POINTER(C.c_double)()
    r.member_step_ms = step_ms.ctypes.data_as(C.POINTER(C.c_double))
    r.member_busy_ms = busy_ms.ctypes.data_as(C.POINTER(C.c_double))
    o = Opts(niter, warmup, int(batched), int(graphed), int(member_timing), int(autotune),
             int(max_workgroups), int(patch_cells), int(separate_bres))
    rc = lib.op2amd_ensemble_run(C.byref(sp), C.byref(o), C.byref(r))
    if rc:
        raise RuntimeError(f"ensemble run failed ({rc})")
    return {"members": members, "rms": rms, "maxerr": maxerr, "errloc": errloc, "q": q,
            "member_step_ms": step_ms, "member_busy_ms": busy_ms, "seconds": r.seconds,
            "setup_seconds": r.setup_seconds, "total_cells": r.total_cells,
            "total_edges": r.total_edges,
            "blocks": dict(zip(LOOPS, list(r.blocks))),
            "grids": dict(zip(LOOPS, list(r.grids))), "dead_res": bool(r.dead_res),
            "fused_bres": bool(r.fused_bres)}
