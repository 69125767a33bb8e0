"""Synthetic loops (include/op2amd_synth.h) and the MLP cost model
(include/op2amd_costmodel.h)."""
import ctypes as C
import math

import numpy as np

from ._lib import lib
from .api import PLAN_ATOMIC, PLAN_COLOR, PLAN_STAGED, PLAN_GATHER
from .mesh import ORDERINGS

DIRECT = 0
OP_READ, OP_INC, OP_RW = 0, 3, 2
ACCESS = {"read": OP_READ, "inc": OP_INC, "rw": OP_RW}


class Desc(C.Structure):
    _fields_ = [("I", C.c_int), ("J", C.c_int), ("ordering", C.c_int), ("arity", C.c_int),
                ("dim", C.c_int), ("access", C.c_int), ("is_double", C.c_int),
                ("k_fma", C.c_int), ("seed", C.c_uint64)]


class Shape(C.Structure):
    _fields_ = [("set_size", C.c_int), ("target_size", C.c_int),
                ("mean_bandwidth", C.c_double)]


lib.op2amd_synth_time.restype = C.c_int
lib.op2amd_synth_time.argtypes = [C.POINTER(Desc), C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(Shape)]
lib.op2amd_synth_run_once.restype = C.c_int
lib.op2amd_synth_run_once.argtypes = [C.POINTER(Desc), C.c_int, C.c_int,
                                      C.POINTER(C.c_double), C.c_int]
lib.op2amd_mlp_create.restype = C.c_void_p
lib.op2amd_mlp_create.argtypes = [C.c_int, C.c_int, C.c_uint64]
lib.op2amd_mlp_destroy.argtypes = [C.c_void_p]
lib.op2amd_mlp_num_params.restype = C.c_int
lib.op2amd_mlp_num_params.argtypes = [C.c_void_p]
for fn in ("op2amd_mlp_get_params", "op2amd_mlp_set_params"):
    getattr(lib, fn).restype = C.c_int
    getattr(lib, fn).argtypes = [C.c_void_p, C.POINTER(C.c_float)]
lib.op2amd_mlp_forward.restype = C.c_int
lib.op2amd_mlp_forward.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int,
                                   C.POINTER(C.c_float)]
lib.op2amd_mlp_loss_grad.restype = C.c_int
lib.op2amd_mlp_loss_grad.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
lib.op2amd_mlp_train.restype = C.c_int
lib.op2amd_mlp_train.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def desc_from(n_target, arity, dim, access, is_double, ordering, k, seed):
    """Descriptor whose loop set has about n_target elements: edges (about
    2 I J) for arity 2, cells (I J) otherwise; I = 2 J."""
    cells = n_target / 2 if arity == 2 else n_target
    J = max(4, int(round(math.sqrt(cells / 2))))
    return Desc(2 * J, J, ORDERINGS[ordering], arity, dim, ACCESS[access], int(is_double),
                k, seed)


def variants(access):
    """Candidate (plan, block) pairs of the exhaustive sweep for an access mode."""
    blocks = (64, 128, 256, 512, 1024)
    if access == "read":
        return [(DIRECT, b) for b in blocks]
    plans = [PLAN_COLOR, PLAN_GATHER] + ([PLAN_ATOMIC] if access == "inc" else [])
    out = [(p, b) for p in plans for b in blocks]
    out += [(PLAN_STAGED, b) for b in (128, 256, 512)]
    return out


def time_variants(desc, var, reps=3):
    n = len(var)
    plan = (C.c_int * n)(*[p for p, _ in var])
    block = (C.c_int * n)(*[b for _, b in var])
    ms = (C.c_double * n)()
    shape = Shape()
    if lib.op2amd_synth_time(C.byref(desc), n, plan, block, reps, ms, C.byref(shape)):
        raise RuntimeError("synthetic loop timing failed")
    return list(ms), {"set_size": shape.set_size, "target_size": shape.target_size,
                      "mean_bandwidth": shape.mean_bandwidth}


def run_once(desc, plan, block, size):
    out = np.zeros(size)
    rc = lib.op2amd_synth_run_once(C.byref(desc), plan, block,
                                   out.ctypes.data_as(C.POINTER(C.c_double)), size)
    if rc:
        raise RuntimeError(f"synthetic loop run failed ({rc})")
    return out


class MLP:
    """The library's HIP MLP (nin -> hidden -> hidden -> 1, ReLU, FP32)."""

    def __init__(self, nin, hidden=64, seed=0):
        self.h = lib.op2amd_mlp_create(nin, hidden, seed)
        if not self.h:
            raise RuntimeError("MLP creation failed (no GPU?)")
        self.nin, self.hidden = nin, hidden
        self.np = lib.op2amd_mlp_num_params(self.h)

    def params(self):
        p = np.zeros(self.np, np.float32)
        lib.op2amd_mlp_get_params(self.h, _fp(p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        lib.op2amd_mlp_set_params(self.h, _fp(p))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[0], np.float32)
        lib.op2amd_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        g = np.zeros(self.np, np.float32)
        lib.op2amd_mlp_loss_grad(self.h, _fp(x), _fp(t), x.shape[0], C.byref(loss), _fp(g))
        return loss.value, g

    def train(self, x, t, epochs, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        lib.op2amd_mlp_train(self.h, _fp(x), _fp(t), x.shape[0], epochs, lr, C.byref(loss))
        return loss.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.op2amd_mlp_destroy(self.h)
            self.h = None


ORDER_NAMES = ("random", "rcm", "hilbert")
PLAN_IDS = (DIRECT, PLAN_ATOMIC, PLAN_COLOR, PLAN_STAGED, PLAN_GATHER)


def features(rec, plan, block):
    """Feature vector of one (loop, variant): sizes, shape, arithmetic and the
    variant itself."""
    f = [math.log10(rec["set_size"]), math.log10(rec["target_size"]), rec["arity"] / 4.0,
         math.log2(rec["dim"]) / 3.0, float(rec["is_double"]), rec["k"] / 64.0,
         math.log10(1.0 + rec["mean_bandwidth"]) / 7.0,
         float(rec["access"] == "read"), float(rec["access"] == "inc"),
         float(rec["access"] == "rw")]
    f += [float(rec["ordering"] == o) for o in ORDER_NAMES]
    f += [float(plan == p) for p in PLAN_IDS]
    f += [math.log2(block) / 10.0]
    return f
