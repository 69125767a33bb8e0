"""ORACLE — test infrastructure only.

ctypes bindings of oracle/_build/liboracle.so (the CPU restatement of the
reference airfoil path, see op2_oracle.h) plus pure-Python restatements used
as independent checks: the RCM rule of include/op2amd_mesh.h and a brute-force
colouring checker.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline may import this module.
"""
import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REF_DIR = HERE / "_ref"


def build():
    subprocess.run(["make", "-s", "-C", os.fspath(HERE)], check=True)


def _load():
    if not LIB.exists():
        build()
    return C.CDLL(os.fspath(LIB))


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        _setup(_lib)
    return _lib


class Consts(C.Structure):
    _fields_ = [("gam", C.c_double), ("gm1", C.c_double), ("cfl", C.c_double),
                ("eps", C.c_double), ("mach", C.c_double), ("alpha", C.c_double),
                ("qinf", C.c_double * 4)]


class Plan(C.Structure):
    _fields_ = [("nblocks", C.c_int), ("ncolors", C.c_int), ("bsize", C.c_int),
                ("offset", C.POINTER(C.c_int)), ("nelems", C.POINTER(C.c_int)),
                ("nthrcol", C.POINTER(C.c_int)), ("thrcol", C.POINTER(C.c_int)),
                ("blk_col", C.POINTER(C.c_int)), ("ncolblk", C.POINTER(C.c_int)),
                ("blkmap", C.POINTER(C.c_int)), ("ninds", C.c_int),
                ("ind_sizes", C.POINTER(C.c_int)), ("ind_offs", C.POINTER(C.c_int)),
                ("ind_map", C.POINTER(C.c_int)), ("nindirect", C.POINTER(C.c_int)),
                ("loc_map", C.POINTER(C.c_short))]


def _setup(L):
    from op2amd.mesh import op2amd_mesh  # struct layout of the mesh views
    M = C.POINTER(op2amd_mesh)
    dp = C.POINTER(C.c_double)
    L.orc_airfoil_consts.argtypes = [C.c_int, C.POINTER(Consts)]
    L.orc_save_soln.argtypes = [C.c_int, dp, dp]
    L.orc_adt_calc.argtypes = [M, C.POINTER(Consts), dp, dp]
    L.orc_res_calc.argtypes = [M, C.POINTER(Consts), dp, dp, dp]
    L.orc_bres_calc.argtypes = [M, C.POINTER(Consts), dp, dp, dp]
    L.orc_update.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, C.POINTER(C.c_int)]
    L.orc_airfoil_run.restype = C.c_int
    L.orc_airfoil_run.argtypes = [M, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  dp, dp, dp, C.POINTER(C.c_int), dp, dp]
    L.orc_airfoil_run_qinf.restype = C.c_int
    L.orc_airfoil_run_qinf.argtypes = [M, C.c_int, dp, C.c_int, C.c_int, C.c_int,
                                       dp, dp, dp, C.POINTER(C.c_int), dp, dp]
    L.orc_plan_build.restype = C.c_int
    L.orc_plan_build.argtypes = [C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_int)),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                 C.POINTER(Plan)]
    L.orc_plan_free.argtypes = [C.POINTER(Plan)]


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def consts(alpha_mode=1):
    c = Consts()
    lib().orc_airfoil_consts(alpha_mode, C.byref(c))
    return c


def save_soln(q):
    qold = np.zeros_like(q)
    lib().orc_save_soln(q.shape[0], _p(q), _p(qold))
    return qold


def adt_calc(mesh, q, alpha_mode=1):
    adt = np.zeros(mesh.ncell)
    lib().orc_adt_calc(C.byref(mesh.c()), C.byref(consts(alpha_mode)),
                       _p(np.ascontiguousarray(q)), _p(adt))
    return adt


def res_calc(mesh, q, adt, res, alpha_mode=1):
    res = np.array(res, dtype=np.float64, copy=True)
    lib().orc_res_calc(C.byref(mesh.c()), C.byref(consts(alpha_mode)),
                       _p(np.ascontiguousarray(q)), _p(np.ascontiguousarray(adt)), _p(res))
    return res


def bres_calc(mesh, q, adt, res, alpha_mode=1):
    res = np.array(res, dtype=np.float64, copy=True)
    lib().orc_bres_calc(C.byref(mesh.c()), C.byref(consts(alpha_mode)),
                        _p(np.ascontiguousarray(q)), _p(np.ascontiguousarray(adt)), _p(res))
    return res


def update(qold, res, adt, rms=0.0, maxerr=0.0, errloc=0):
    q = np.zeros_like(qold)
    res = np.array(res, copy=True)
    r, m, e = C.c_double(rms), C.c_double(maxerr), C.c_int(errloc)
    lib().orc_update(qold.shape[0], _p(qold), _p(q), _p(res), _p(adt),
                     C.byref(r), C.byref(m), C.byref(e))
    return q, res, r.value, m.value, e.value


# "seq": the sequential executor; "openmp": the OpenMP executor (block colours
# in sequence, blocks in parallel); "seqpar": the sequential executor's results
# bit for bit, evaluated on `nthreads` threads (tests/test_oracle.py pins it
# to "seq").
EXECS = {"seq": 0, "openmp": 1, "seqpar": 2}


def airfoil_run(mesh, niter, alpha_mode=1, exec_="seq", nthreads=0, part_size=0):
    rms_k0 = np.zeros(niter)
    rms = np.zeros(niter)
    maxerr = np.zeros(niter)
    errloc = np.zeros(niter, np.int32)
    q = np.zeros((mesh.ncell, 4))
    secs = np.zeros(5)
    rc = lib().orc_airfoil_run(C.byref(mesh.c()), niter, alpha_mode,
                               EXECS[exec_], nthreads, part_size,
                               _p(rms_k0), _p(rms), _p(maxerr),
                               errloc.ctypes.data_as(C.POINTER(C.c_int)), _p(q),
                               _p(secs))
    if rc:
        raise RuntimeError("oracle airfoil run failed")
    return {"rms_k0": rms_k0, "rms": rms, "maxerr": maxerr, "errloc": errloc,
            "q": q, "loop_seconds": secs}


def airfoil_run_qinf(mesh, niter, qinf, exec_="seq", nthreads=0):
    """Airfoil run with an explicit free stream (ensemble members)."""
    rms_k0 = np.zeros(niter)
    rms = np.zeros(niter)
    maxerr = np.zeros(niter)
    errloc = np.zeros(niter, np.int32)
    q = np.zeros((mesh.ncell, 4))
    qi = np.ascontiguousarray(qinf, dtype=np.float64)
    rc = lib().orc_airfoil_run_qinf(C.byref(mesh.c()), niter, _p(qi),
                                    EXECS[exec_], nthreads, 0,
                                    _p(rms_k0), _p(rms), _p(maxerr),
                                    errloc.ctypes.data_as(C.POINTER(C.c_int)), _p(q),
                                    None)
    if rc:
        raise RuntimeError("oracle airfoil run failed")
    return {"rms_k0": rms_k0, "rms": rms, "maxerr": maxerr, "errloc": errloc, "q": q}


def plan_build(set_size, args, part_size=0):
    """op_plan_core(OP_STAGE_ALL) restatement.  `args`: list of dicts with keys
    map (2-D int array or None), idx, ind, acc, bytes, to_size."""
    n = len(args)
    keep = []
    maps = (C.POINTER(C.c_int) * n)()
    for i, a in enumerate(args):
        if a["map"] is None:
            maps[i] = C.POINTER(C.c_int)()
        else:
            arr = np.ascontiguousarray(a["map"], dtype=np.int32)
            keep.append(arr)
            maps[i] = arr.ctypes.data_as(C.POINTER(C.c_int))

    def iv(key, default=0):
        return (C.c_int * n)(*[a.get(key, default) for a in args])
    mdim = (C.c_int * n)(*[(a["map"].shape[1] if a["map"] is not None else 0) for a in args])
    P = Plan()
    lib().orc_plan_build(set_size, n, maps, mdim, iv("idx", -1), iv("ind", -1),
                         iv("acc"), iv("bytes"), iv("to_size"), part_size, C.byref(P))
    nb, nc, ni = P.nblocks, P.ncolors, P.ninds

    def arr(ptr, k):
        return np.ctypeslib.as_array(ptr, shape=(k,)).copy() if k else np.zeros(0, np.int32)
    out = {
        "nblocks": nb, "ncolors": nc, "bsize": P.bsize,
        "offset": arr(P.offset, nb), "nelems": arr(P.nelems, nb),
        "nthrcol": arr(P.nthrcol, nb), "thrcol": arr(P.thrcol, set_size),
        "blk_col": arr(P.blk_col, nb), "ncolblk": arr(P.ncolblk, nc),
        "blkmap": arr(P.blkmap, nb), "ninds": ni,
        "ind_sizes": arr(P.ind_sizes, nb * ni), "ind_offs": arr(P.ind_offs, nb * ni),
        "nindirect": arr(P.nindirect, ni),
    }
    uses = [sum(1 for a in args if a.get("ind", -1) == i) for i in range(ni)]
    base = np.cumsum([0] + [u * set_size for u in uses])
    allind = arr(P.ind_map, int(base[-1]))
    out["ind_maps"] = [allind[base[i]:base[i] + out["nindirect"][i]] for i in range(ni)]
    loc = np.ctypeslib.as_array(P.loc_map, shape=(n * set_size,)).copy() if set_size else np.zeros(0, np.int16)
    out["loc_maps"] = [loc[i * set_size:(i + 1) * set_size] if args[i].get("ind", -1) >= 0 else None
                       for i in range(n)]
    lib().orc_plan_free(C.byref(P))
    return out


def airfoil_plan_args(mesh, loop):
    """Argument descriptors of the airfoil loops as airfoil.cpp:235-261 declares them."""
    nc, nn, nb = mesh.ncell, mesh.nnode, mesh.nbedge
    if loop == "adt_calc":
        return mesh.ncell, [dict(map=mesh.cell, idx=k, ind=0, acc=0, bytes=16, to_size=nn) for k in range(4)] + [
            dict(map=None, idx=-1, ind=-1, acc=0, bytes=32, to_size=nc),
            dict(map=None, idx=-1, ind=-1, acc=1, bytes=8, to_size=nc)]
    if loop == "res_calc":
        return mesh.nedge, [
            dict(map=mesh.edge, idx=0, ind=0, acc=0, bytes=16, to_size=nn),
            dict(map=mesh.edge, idx=1, ind=0, acc=0, bytes=16, to_size=nn),
            dict(map=mesh.ecell, idx=0, ind=1, acc=0, bytes=32, to_size=nc),
            dict(map=mesh.ecell, idx=1, ind=1, acc=0, bytes=32, to_size=nc),
            dict(map=mesh.ecell, idx=0, ind=2, acc=0, bytes=8, to_size=nc),
            dict(map=mesh.ecell, idx=1, ind=2, acc=0, bytes=8, to_size=nc),
            dict(map=mesh.ecell, idx=0, ind=3, acc=3, bytes=32, to_size=nc),
            dict(map=mesh.ecell, idx=1, ind=3, acc=3, bytes=32, to_size=nc)]
    if loop == "bres_calc":
        becell = mesh.becell.reshape(-1, 1)
        return mesh.nbedge, [
            dict(map=mesh.bedge, idx=0, ind=0, acc=0, bytes=16, to_size=nn),
            dict(map=mesh.bedge, idx=1, ind=0, acc=0, bytes=16, to_size=nn),
            dict(map=becell, idx=0, ind=1, acc=0, bytes=32, to_size=nc),
            dict(map=becell, idx=0, ind=2, acc=0, bytes=8, to_size=nc),
            dict(map=becell, idx=0, ind=3, acc=3, bytes=32, to_size=nc),
            dict(map=None, idx=-1, ind=-1, acc=0, bytes=4, to_size=nb)]
    raise KeyError(loop)


# ---------------------------------------------------------------------------
# Pure-Python restatements (independent of both C++ implementations)
# ---------------------------------------------------------------------------
def colouring_conflicts(targets, colour, group=None):
    """Brute-force check: number of pairs of distinct elements with the same
    colour (and the same group, e.g. plan block, when given) that share an
    increment target.  `targets`: [n][k] target ids per element."""
    n = targets.shape[0]
    key = colour if group is None else colour.astype(np.int64) * (1 << 32) + group
    seen = {}
    conflicts = 0
    for e in range(n):
        for t in set(int(v) for v in targets[e]):
            k = (int(key[e]), t)
            if k in seen and seen[k] != e:
                conflicts += 1
            seen[k] = e
    return conflicts


def rcm_order(ncell, ecell):
    """Reverse Cuthill-McKee order of the cell graph with the rule stated in
    include/op2amd_mesh.h: George-Liu pseudo-peripheral start from the
    lowest-(degree, index) unvisited cell, neighbours queued by (degree, index),
    components one after another, the Cuthill-McKee order reversed.
    Returns new_of_old for cells."""
    adj = [[] for _ in range(ncell)]
    for a, b in ecell:
        adj[int(a)].append(int(b))
        adj[int(b)].append(int(a))
    for lst in adj:
        lst.sort()
    deg = [len(a) for a in adj]
    done = [False] * ncell

    def key(v):
        return (deg[v], v)

    def bfs(root):
        level = {root: 0}
        frontier = [root]
        last = [root]
        depth = 0
        while frontier:
            nxt = []
            for v in frontier:
                for w in adj[v]:
                    if not done[w] and w not in level:
                        level[w] = depth + 1
                        nxt.append(w)
            if not nxt:
                break
            last = nxt
            frontier = nxt
            depth += 1
        return depth, last

    cm = []
    while len(cm) < ncell:
        root = min((v for v in range(ncell) if not done[v]), key=key)
        ecc, last = bfs(root)
        for _ in range(16):
            cand = min(last, key=key)
            ecc2, last2 = bfs(cand)
            if ecc2 <= ecc:
                break
            root, ecc, last = cand, ecc2, last2
        head = len(cm)
        cm.append(root)
        done[root] = True
        while head < len(cm):
            v = cm[head]
            head += 1
            for w in sorted((w for w in adj[v] if not done[w]), key=key):
                done[w] = True
                cm.append(w)
    new_of_old = np.empty(ncell, np.int32)
    for k, v in enumerate(reversed(cm)):
        new_of_old[v] = k
    return new_of_old
