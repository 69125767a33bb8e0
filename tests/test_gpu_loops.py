"""Single op_par_loop calls through the generated stubs (C ABI, ctypes) vs the
oracle's single loops on the same seeded inputs."""
import ctypes as C

import numpy as np
import pytest

import op2amd
import op2amd.api as api
from conftest import rel_err
from oracle import oracle

pytestmark = pytest.mark.gpu


class App:
    """Declares the airfoil sets/maps/dats through the OP2 C API."""

    def __init__(self, mesh, q, res, soa=False):
        api.init(*(["OP_AUTO_SOA"] if soa else []))
        self.m = mesh
        self.nodes = api.op_decl_set(mesh.nnode, b"nodes")
        self.edges = api.op_decl_set(mesh.nedge, b"edges")
        self.bedges = api.op_decl_set(mesh.nbedge, b"bedges")
        self.cells = api.op_decl_set(mesh.ncell, b"cells")
        self.pedge = api.decl_map(self.edges, self.nodes, 2, mesh.edge, "pedge")
        self.pecell = api.decl_map(self.edges, self.cells, 2, mesh.ecell, "pecell")
        self.pbedge = api.decl_map(self.bedges, self.nodes, 2, mesh.bedge, "pbedge")
        self.pbecell = api.decl_map(self.bedges, self.cells, 1, mesh.becell, "pbecell")
        self.pcell = api.decl_map(self.cells, self.nodes, 4, mesh.cell, "pcell")
        self.bound = api.decl_dat(self.bedges, 1, "int", mesh.bound, "p_bound")
        self.x = api.decl_dat(self.nodes, 2, "double", mesh.x, "p_x")
        self.q = api.decl_dat(self.cells, 4, "double", q, "p_q")
        self.qold = api.decl_dat(self.cells, 4, "double", q, "p_qold")
        self.adt = api.decl_dat(self.cells, 1, "double", np.zeros(mesh.ncell), "p_adt")
        self.res = api.decl_dat(self.cells, 4, "double", res, "p_res")
        c = oracle.consts(1)
        for name, val in (("gam", [c.gam]), ("gm1", [c.gm1]), ("cfl", [c.cfl]),
                          ("eps", [c.eps]), ("mach", [c.mach]), ("alpha", [c.alpha]),
                          ("qinf", list(c.qinf))):
            a = np.array(val, dtype=np.float64)
            api.op_decl_const_char(len(val), b"double", 8, a.ctypes.data_as(C.c_void_p),
                                   name.encode())

    def d(self, dat, idx, mp, dim, typ, acc):
        return api.arg_dat(dat, idx, mp, dim, typ, acc)

    def adt_calc(self):
        R = api.OP_READ
        api.loops["op_par_loop_airfoil_2_adt_calc"](
            b"adt_calc", self.cells, *[self.d(self.x, k, self.pcell, 2, "double", R) for k in range(4)],
            self.d(self.q, -1, None, 4, "double", R),
            self.d(self.adt, -1, None, 1, "double", api.OP_WRITE))

    def res_calc(self):
        R, I = api.OP_READ, api.OP_INC
        api.loops["op_par_loop_airfoil_3_res_calc"](
            b"res_calc", self.edges,
            self.d(self.x, 0, self.pedge, 2, "double", R), self.d(self.x, 1, self.pedge, 2, "double", R),
            self.d(self.q, 0, self.pecell, 4, "double", R), self.d(self.q, 1, self.pecell, 4, "double", R),
            self.d(self.adt, 0, self.pecell, 1, "double", R), self.d(self.adt, 1, self.pecell, 1, "double", R),
            self.d(self.res, 0, self.pecell, 4, "double", I), self.d(self.res, 1, self.pecell, 4, "double", I))

    def bres_calc(self):
        R = api.OP_READ
        api.loops["op_par_loop_airfoil_4_bres_calc"](
            b"bres_calc", self.bedges,
            self.d(self.x, 0, self.pbedge, 2, "double", R), self.d(self.x, 1, self.pbedge, 2, "double", R),
            self.d(self.q, 0, self.pbecell, 4, "double", R), self.d(self.adt, 0, self.pbecell, 1, "double", R),
            self.d(self.res, 0, self.pbecell, 4, "double", api.OP_INC),
            self.d(self.bound, -1, None, 1, "int", R))

    def update(self, rms, maxerr, errloc):
        api.loops["op_par_loop_airfoil_5_update"](
            b"update", self.cells,
            self.d(self.qold, -1, None, 4, "double", api.OP_READ),
            self.d(self.q, -1, None, 4, "double", api.OP_WRITE),
            self.d(self.res, -1, None, 4, "double", api.OP_RW),
            self.d(self.adt, -1, None, 1, "double", api.OP_READ),
            api.op_arg_gbl_char(C.cast(C.pointer(rms), C.c_void_p), 1, b"double", 8, api.OP_INC),
            api.op_arg_gbl_char(C.cast(C.pointer(maxerr), C.c_void_p), 1, b"double", 8, api.OP_MAX),
            api.op_arg_idx(-1, api.OP_ID),
            api.op_arg_info_char(C.cast(C.pointer(errloc), C.c_void_p), 1, b"int", 4, 5))

    def fetch(self, dat, shape):
        return api.fetch(dat, shape, np.float64)

    def close(self):
        api.op_exit()


def inputs(mesh, seed):
    rng = np.random.default_rng(seed)
    c = oracle.consts(1)
    q = np.tile(np.array(c.qinf[:]), (mesh.ncell, 1)) * (1 + 0.02 * rng.standard_normal((mesh.ncell, 4)))
    res = 1e-3 * rng.standard_normal((mesh.ncell, 4))
    return q, res


@pytest.mark.parametrize("soa", [False, True])
@pytest.mark.parametrize("plan", ["atomic", "color", "staged", "gather", "owner"])
@pytest.mark.parametrize("ordering", ["random", "rcm"])
def test_indirect_loops_match_oracle(plan, soa, ordering):
    mesh = op2amd.mesh.generate(60, 30, 11, ordering)
    q, res0 = inputs(mesh, 1)
    app = App(mesh, q, res0, soa=soa)
    try:
        api.op2h_set_loop_plan(b"res_calc", api.PLAN_NAMES[plan], 0)
        api.op2h_set_loop_plan(b"bres_calc", api.PLAN_NAMES[plan], 0)
        app.adt_calc()
        adt = app.fetch(app.adt, (mesh.ncell,))
        app.res_calc()
        res1 = app.fetch(app.res, (mesh.ncell, 4))
        app.bres_calc()
        res2 = app.fetch(app.res, (mesh.ncell, 4))
    finally:
        app.close()
    adt_o = oracle.adt_calc(mesh, q)
    assert np.array_equal(adt, adt_o)
    res1_o = oracle.res_calc(mesh, q, adt_o, res0)
    res2_o = oracle.bres_calc(mesh, q, adt_o, res1_o)
    if plan in ("gather", "owner"):
        assert np.array_equal(res1, res1_o) and np.array_equal(res2, res2_o)
    else:
        assert rel_err(res1, res1_o) < 1e-14
        assert rel_err(res2, res2_o) < 1e-14


@pytest.mark.parametrize("soa", [False, True])
def test_direct_loops_and_reductions(soa):
    mesh = op2amd.mesh.generate(97, 41, 2, "random")
    q, res = inputs(mesh, 3)
    app = App(mesh, q, res, soa=soa)
    try:
        app.adt_calc()
        adt = app.fetch(app.adt, (mesh.ncell,))
        rms, maxerr, errloc = C.c_double(0.25), C.c_double(0.0), C.c_int(-7)
        app.update(rms, maxerr, errloc)
        qn = app.fetch(app.q, (mesh.ncell, 4))
        rn = app.fetch(app.res, (mesh.ncell, 4))
        # a second call with a maxerr that no element exceeds leaves errloc alone
        big, loc2, r2 = C.c_double(1e300), C.c_int(-3), C.c_double(0.0)
        app.update(r2, big, loc2)
    finally:
        app.close()
    qn_o, rn_o, rms_o, mx_o, loc_o = oracle.update(q, res, adt, 0.25, 0.0, -7)
    assert np.array_equal(qn, qn_o) and np.array_equal(rn, rn_o)
    assert abs(rms.value - rms_o) <= 1e-14 * rms_o
    assert maxerr.value == mx_o and errloc.value == loc_o
    assert big.value == 1e300 and loc2.value == -3 and r2.value == 0.0


def test_op_renumber_matches_generator_rcm():
    """op_renumber(pecell) on the randomly numbered mesh gives the numbering
    of the generator's RCM ordering: same RCM rule, same propagation rules."""
    rnd = op2amd.mesh.generate(40, 20, 5, "random")
    rcm = op2amd.mesh.generate(40, 20, 5, "rcm")
    q, res = inputs(rnd, 0)
    app = App(rnd, q, res)
    try:
        api.lib.op_renumber.argtypes = [api.op_map]
        api.lib.op_renumber(app.pecell)
        x = app.fetch(app.x, (rnd.nnode, 2))

        def table(m, rows, dim):
            return np.ctypeslib.as_array(m.contents.map, shape=(rows * dim,)).reshape(rows, dim).copy()
        ecell = table(app.pecell, rnd.nedge, 2)
        edge = table(app.pedge, rnd.nedge, 2)
        cell = table(app.pcell, rnd.ncell, 4)
        becell = table(app.pbecell, rnd.nbedge, 1).ravel()
    finally:
        app.close()
    assert np.array_equal(x, rcm.x)
    assert np.array_equal(ecell, rcm.ecell) and np.array_equal(edge, rcm.edge)
    assert np.array_equal(cell, rcm.cell) and np.array_equal(becell, rcm.becell)


@pytest.mark.parametrize("plan,block,ordering", [
    ("owner", 128, "rcm"), ("gather", 256, "rcm"), ("owner", 256, "rcm"),
    ("owner", 128, "hilbert"), ("owner", 128, "random")])
def test_full_size_loops_bit_exact(plan, block, ordering):
    """At the benchmark's size (1200x600, 720k cells): adt_calc, and res_calc
    then bres_calc with an owner-compute plan, equal the oracle's sequential
    loops bit for bit; the interior fluxes conserve (sum of res after res_calc
    from zero is zero up to rounding).  The owner plan's 8-byte records (16-bit
    offsets from a per-patch base) serve every patch of the RCM numbering,
    some of the Hilbert one's (the others span more than 65535 indices and
    keep 16-byte records) and none of the random one's."""
    mesh = op2amd.mesh.generate(1200, 600, 0, ordering)
    q, _ = inputs(mesh, 4)
    res0 = np.zeros((mesh.ncell, 4))
    app = App(mesh, q, res0)
    try:
        api.op2h_set_loop_plan(b"res_calc", api.PLAN_NAMES[plan], block)
        api.op2h_set_loop_plan(b"bres_calc", api.PLAN_NAMES["gather"], 128)
        app.adt_calc()
        adt = app.fetch(app.adt, (mesh.ncell,))
        app.res_calc()
        res1 = app.fetch(app.res, (mesh.ncell, 4))
        app.bres_calc()
        res2 = app.fetch(app.res, (mesh.ncell, 4))
    finally:
        app.close()
    adt_o = oracle.adt_calc(mesh, q)
    assert np.array_equal(adt, adt_o)
    res1_o = oracle.res_calc(mesh, q, adt_o, res0)
    assert np.array_equal(res1, res1_o)
    assert np.array_equal(res2, oracle.bres_calc(mesh, q, adt_o, res1_o))
    scale = np.abs(res1).sum(axis=0)
    assert np.all(np.abs(res1.sum(axis=0)) <= 1e-12 * scale)


def test_empty_sets():
    """Loops over empty sets leave dats and gbl arguments as given (a
    reduction over nothing is the identity, combined with the caller's
    value)."""
    api.init()
    try:
        cells = api.op_decl_set(0, b"cells")
        empty = np.zeros((0, 4))
        q = api.decl_dat(cells, 4, "double", empty, "q")
        qold = api.decl_dat(cells, 4, "double", empty, "qold")
        res = api.decl_dat(cells, 4, "double", empty, "res")
        adt = api.decl_dat(cells, 1, "double", np.zeros(0), "adt")
        d = api.arg_dat
        api.loops["op_par_loop_airfoil_1_save_soln"](
            b"save_soln", cells, d(q, -1, None, 4, "double", api.OP_READ),
            d(qold, -1, None, 4, "double", api.OP_WRITE))
        rms, maxerr, errloc = C.c_double(0.5), C.c_double(2.0), C.c_int(-3)
        api.loops["op_par_loop_airfoil_5_update"](
            b"update", cells, d(qold, -1, None, 4, "double", api.OP_READ),
            d(q, -1, None, 4, "double", api.OP_WRITE), d(res, -1, None, 4, "double", api.OP_RW),
            d(adt, -1, None, 1, "double", api.OP_READ),
            api.op_arg_gbl_char(C.cast(C.pointer(rms), C.c_void_p), 1, b"double", 8, api.OP_INC),
            api.op_arg_gbl_char(C.cast(C.pointer(maxerr), C.c_void_p), 1, b"double", 8, api.OP_MAX),
            api.op_arg_idx(-1, api.OP_ID),
            api.op_arg_info_char(C.cast(C.pointer(errloc), C.c_void_p), 1, b"int", 4, 5))
        api.op2h_device_sync()
        assert rms.value == 0.5 and maxerr.value == 2.0 and errloc.value == -3
    finally:
        api.op_exit()
