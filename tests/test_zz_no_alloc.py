"""op2h_update_noalloc, the update loop on caller-owned device arrays that
allocates nothing (its partials live in a device array of the library),
against the existing update stub op_par_loop_airfoil_5_update on the same
inputs: q, res, rms, maxerr and errloc bit for bit (the same kernels and
reduction tree), on two calls in a row, the second on fewer cells, in AoS and
SoA layouts.  The test itself allocates the device arrays it passes."""
import ctypes as C

import numpy as np
import pytest

import op2amd.api as api
from oracle import oracle

pytestmark = pytest.mark.gpu


class Result(C.Structure):
    _fields_ = [("rms", C.c_double), ("maxerr", C.c_double), ("errloc", C.c_int),
                ("pad", C.c_int)]


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


H2D, D2H = 1, 2


class Dev:
    """A device copy of a host array (hipMalloc'd by the test)."""

    def __init__(self, h, host):
        self.h, self.n = h, host.nbytes
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(self.n, 8)) == 0
        if self.n:
            assert h.hipMemcpy(self.p, host.ctypes.data, self.n, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        if self.n:
            assert self.h.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def stub_update(qold, q, res, adt, soa):
    """The existing entry point on a plain cell set declared through the API."""
    n = len(adt)
    api.init(*(["OP_AUTO_SOA"] if soa else []))
    try:
        cells = api.op_decl_set(n, b"cells")
        d_qold = api.decl_dat(cells, 4, "double", qold, "qold")
        d_q = api.decl_dat(cells, 4, "double", q, "q")
        d_res = api.decl_dat(cells, 4, "double", res, "res")
        d_adt = api.decl_dat(cells, 1, "double", adt, "adt")
        d = api.arg_dat
        rms, maxerr, errloc = C.c_double(0.0), C.c_double(-np.inf), C.c_int(-1)
        api.loops["op_par_loop_airfoil_5_update"](
            b"update", cells, d(d_qold, -1, None, 4, "double", api.OP_READ),
            d(d_q, -1, None, 4, "double", api.OP_WRITE), d(d_res, -1, None, 4, "double", api.OP_RW),
            d(d_adt, -1, None, 1, "double", api.OP_READ),
            api.op_arg_gbl_char(C.cast(C.pointer(rms), C.c_void_p), 1, b"double", 8, api.OP_INC),
            api.op_arg_gbl_char(C.cast(C.pointer(maxerr), C.c_void_p), 1, b"double", 8, api.OP_MAX),
            api.op_arg_idx(-1, api.OP_ID),
            api.op_arg_info_char(C.cast(C.pointer(errloc), C.c_void_p), 1, b"int", 4, 5))
        qn = api.fetch(d_q, (n, 4), np.float64)
        rn = api.fetch(d_res, (n, 4), np.float64)
    finally:
        api.op_exit()
    return qn, rn, rms.value, maxerr.value, errloc.value


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    c = oracle.consts(1)
    qold = np.tile(np.array(c.qinf[:]), (n, 1)) * (1 + 0.02 * rng.standard_normal((n, 4)))
    q = qold + 1e-3 * rng.standard_normal((n, 4))
    res = 1e-3 * rng.standard_normal((n, 4))
    adt = 0.5 + rng.random(n)
    return qold, q, res, adt


def to_soa(a, stride):
    out = np.zeros(4 * stride)
    for k in range(4):
        out[k * stride:k * stride + len(a)] = a[:, k]
    return out


def from_soa(a, n, stride):
    return np.stack([a[k * stride:k * stride + n] for k in range(4)], axis=1)


@pytest.mark.parametrize("soa", [False, True], ids=["aos", "soa"])
def test_update_noalloc_equals_update_stub_two_calls(soa):
    lib = api.lib
    lib.op2h_update_noalloc.restype = C.c_int
    lib.op2h_update_noalloc.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6
    h = hip()
    for n, seed in ((300_000, 1), (70_001, 2)):  # the second call on fewer cells
        qold, q, res, adt = inputs(n, seed)
        qn_s, rn_s, rms_s, max_s, loc_s = stub_update(qold, q, res, adt, soa)
        stride = (n + 63) // 64 * 64 if soa else 0
        lay = (lambda a: to_soa(a, stride)) if soa else (lambda a: a)
        bufs = [Dev(h, np.ascontiguousarray(lay(a))) for a in (qold, q, res)] + [Dev(h, adt)]
        out = Result(0.0, 0.0, 0, 0)
        dres = Dev(h, np.frombuffer(bytes(out), dtype=np.uint8).copy())
        try:
            print(f"op2h_update_noalloc: {n} cells, soa {soa}", flush=True)
            rc = lib.op2h_update_noalloc(n, stride, 256, bufs[0].p, bufs[1].p, bufs[2].p,
                                         bufs[3].p, dres.p, None)
            assert rc == 0
            assert h.hipDeviceSynchronize() == 0
            qn = bufs[1].get(lay(q))
            rn = bufs[2].get(lay(res))
            raw = dres.get(np.zeros(C.sizeof(Result), np.uint8))
            got = Result.from_buffer_copy(raw.tobytes())
        finally:
            for b in bufs + [dres]:
                b.free()
        if soa:
            qn, rn = from_soa(qn, n, stride), from_soa(rn, n, stride)
        assert np.array_equal(qn, qn_s) and np.array_equal(rn, rn_s)
        assert np.all(rn == 0.0)
        # the stub combines the loop's result with 0.0 / -inf / -1: the same values
        assert got.rms == rms_s and got.maxerr == max_s and got.errloc == loc_s
        assert 0 <= got.errloc < n


def test_update_noalloc_rejects_bad_arguments():
    lib = api.lib
    lib.op2h_update_noalloc.restype = C.c_int
    lib.op2h_update_noalloc.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6
    dummy = C.c_void_p(16)
    # checked on the host before any device work
    assert lib.op2h_update_noalloc(-1, 0, 256, dummy, dummy, dummy, dummy, dummy, None) == -1
    assert lib.op2h_update_noalloc(10, 0, 100, dummy, dummy, dummy, dummy, dummy, None) == -1
    assert lib.op2h_update_noalloc(10, 5, 256, dummy, dummy, dummy, dummy, dummy, None) == -1
    assert lib.op2h_update_noalloc(10, 0, 256, None, dummy, dummy, dummy, dummy, None) == -1
    assert lib.op2h_update_noalloc(10, 0, 256, dummy, dummy, dummy, dummy, None, None) == -1
