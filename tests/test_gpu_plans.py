"""Plans built by the runtime: the host builder reproduces the reference's
greedy colouring exactly; every colouring is conflict-free."""
import numpy as np
import pytest

import op2amd
import op2amd.api as api
from oracle import oracle
from test_gpu_loops import App, inputs

pytestmark = pytest.mark.gpu


def res_args(app):
    R, I = api.OP_READ, api.OP_INC
    d = app.d
    args = [d(app.x, 0, app.pedge, 2, "double", R), d(app.x, 1, app.pedge, 2, "double", R),
            d(app.q, 0, app.pecell, 4, "double", R), d(app.q, 1, app.pecell, 4, "double", R),
            d(app.adt, 0, app.pecell, 1, "double", R), d(app.adt, 1, app.pecell, 1, "double", R),
            d(app.res, 0, app.pecell, 4, "double", I), d(app.res, 1, app.pecell, 4, "double", I)]
    return (api.op_arg * 8)(*args)


def download_staged(p, n):
    thr = np.zeros(n, np.int32)
    blk = np.zeros(p.contents.nblocks, np.int32)
    off = np.zeros(p.contents.nblocks, np.int32)
    nel = np.zeros(p.contents.nblocks, np.int32)
    import ctypes as C
    api.op2h_staged_plan_download(p, *[a.ctypes.data_as(C.c_void_p) for a in (thr, blk, off, nel)])
    return thr, blk, off, nel


def block_conflicts(ecell, blkcol_of_edge, block_of_edge):
    """Two different blocks of one block colour must not touch one cell."""
    key = np.concatenate([blkcol_of_edge, blkcol_of_edge]).astype(np.int64) * (1 << 32) + \
        np.concatenate([ecell[:, 0], ecell[:, 1]])
    blk = np.concatenate([block_of_edge, block_of_edge])
    order = np.lexsort((blk, key))
    k, b = key[order], blk[order]
    same = (k[1:] == k[:-1]) & (b[1:] != b[:-1])
    return int(same.sum())


def thread_conflicts(ecell, thrcol, block_of_edge):
    key = (np.concatenate([block_of_edge, block_of_edge]).astype(np.int64) * 256 +
           np.concatenate([thrcol, thrcol])) * (1 << 32) + np.concatenate([ecell[:, 0], ecell[:, 1]])
    e = np.concatenate([np.arange(len(thrcol))] * 2)
    order = np.lexsort((e, key))
    k, ee = key[order], e[order]
    return int(((k[1:] == k[:-1]) & (ee[1:] != ee[:-1])).sum())


@pytest.mark.parametrize("part", [64, 256])
def test_host_staged_plan_equals_reference_greedy(part):
    mesh = op2amd.mesh.generate(60, 30, 0, "random")
    q, res = inputs(mesh, 0)
    app = App(mesh, q, res)
    try:
        api.op2h_set_plan_builder(0)
        p = api.op2h_plan_staged_for(b"res_calc", app.edges, part, 8, res_args(app))
        thr, blk, off, nel = download_staged(p, mesh.nedge)
        ncolors = p.contents.ncolors
    finally:
        app.close()
    n, args = oracle.airfoil_plan_args(mesh, "res_calc")
    ref = oracle.plan_build(n, args, part)
    assert np.array_equal(thr, ref["thrcol"])
    assert np.array_equal(blk, ref["blk_col"])
    assert np.array_equal(off, ref["offset"]) and np.array_equal(nel, ref["nelems"])
    assert ncolors == ref["ncolors"]


@pytest.mark.parametrize("I,J", [(60, 30), (1200, 600)])
@pytest.mark.parametrize("builder", [0, 1])
def test_staged_and_flat_colourings_conflict_free(I, J, builder):
    mesh = op2amd.mesh.generate(I, J, 0, "random")
    q, res = inputs(mesh, 0)
    app = App(mesh, q, res)
    try:
        api.op2h_set_plan_builder(builder)
        args = res_args(app)
        p = api.op2h_plan_staged_for(b"res_calc", app.edges, 256, 8, args)
        thr, blk, off, nel = download_staged(p, mesh.nedge)
        c = api.op2h_plan_color_for(b"res_calc", app.edges, 8, args)
        col = np.zeros(mesh.nedge, np.int32)
        import ctypes as C
        api.op2h_color_plan_download(c, col.ctypes.data_as(C.c_void_p))
        ncol_flat = c.contents.ncolors
    finally:
        app.close()
    block_of_edge = np.repeat(np.arange(len(nel)), nel)
    assert thread_conflicts(mesh.ecell, thr, block_of_edge) == 0
    assert block_conflicts(mesh.ecell, blk[block_of_edge], block_of_edge) == 0
    assert thread_conflicts(mesh.ecell, col, np.zeros(mesh.nedge, np.int64)) == 0
    assert ncol_flat <= 8


@pytest.mark.parametrize("builder", [0, 1])
def test_flat_colouring_of_four_increments_per_element(builder):
    """Cells incrementing their four nodes (a map of dim 4, four OP_INC args):
    no two cells of one colour share a node."""
    mesh = op2amd.mesh.generate(90, 45, 2, "rcm")
    q, res = inputs(mesh, 0)
    app = App(mesh, q, res)
    try:
        api.op2h_set_plan_builder(builder)
        w = api.decl_dat(app.nodes, 2, "double", np.zeros((mesh.nnode, 2)), "w")
        args = (api.op_arg * 4)(*[app.d(w, k, app.pcell, 2, "double", api.OP_INC)
                                  for k in range(4)])
        c = api.op2h_plan_color_for(b"cell_to_node", app.cells, 4, args)
        col = np.zeros(mesh.ncell, np.int32)
        import ctypes as C
        api.op2h_color_plan_download(c, col.ctypes.data_as(C.c_void_p))
        ncol = c.contents.ncolors
    finally:
        app.close()
    key = np.concatenate([col.astype(np.int64) * (1 << 32) + mesh.cell[:, k] for k in range(4)])
    cell = np.concatenate([np.arange(mesh.ncell)] * 4)
    order = np.lexsort((cell, key))
    k, e = key[order], cell[order]
    assert int(((k[1:] == k[:-1]) & (e[1:] != e[:-1])).sum()) == 0
    assert 4 <= ncol <= 16


@pytest.mark.parametrize("part", [64, 256])
def test_device_thread_colours_equal_reference_greedy(part):
    """The device builder colours threads with the reference's sequential
    greedy (one lane per block); its block colouring is its own but valid."""
    mesh = op2amd.mesh.generate(60, 30, 0, "rcm")
    q, res = inputs(mesh, 0)
    app = App(mesh, q, res)
    try:
        api.op2h_set_plan_builder(1)
        p = api.op2h_plan_staged_for(b"res_calc", app.edges, part, 8, res_args(app))
        assert p.contents.built_on_device == 1
        thr, blk, off, nel = download_staged(p, mesh.nedge)
        ncolors = p.contents.ncolors
    finally:
        app.close()
    n, args = oracle.airfoil_plan_args(mesh, "res_calc")
    ref = oracle.plan_build(n, args, part)
    assert np.array_equal(thr, ref["thrcol"])
    block_of_edge = np.repeat(np.arange(len(nel)), nel)
    assert block_conflicts(mesh.ecell, blk[block_of_edge], block_of_edge) == 0
    assert ncolors <= 2 * ref["ncolors"] + 2


@pytest.mark.parametrize("ordering,block", [("rcm", 256), ("random", 64), ("hilbert", 512)])
def test_owner_plan_invariants(ordering, block):
    """Every (edge, side) increment appears once, in its cell's own block, in
    ascending edge order; blocks own disjoint cells; boundary edges are the
    only ones evaluated twice."""
    mesh = op2amd.mesh.generate(80, 40, 3, ordering)
    q, res = inputs(mesh, 0)
    app = App(mesh, q, res)
    try:
        p = api.op2h_plan_owner_for(b"res_calc", app.edges, block, 8, res_args(app))
        P = p.contents
        nb, tot = P.nblocks, P.total
        owned_offs = np.zeros(nb + 1, np.int32)
        owned = np.zeros(mesh.ncell, np.int32)
        elem_offs = np.zeros(nb + 1, np.int32)
        elems = np.zeros(tot, np.int32)
        csr_offs = np.zeros(mesh.ncell + 1, np.int32)
        csr = np.zeros(2 * mesh.nedge, np.uint16)
        import ctypes as C
        api.op2h_owner_plan_download(p, *[a.ctypes.data_as(C.c_void_p) for a in
                                          (owned_offs, owned, elem_offs, elems, csr_offs, csr)])
    finally:
        app.close()
    ec = mesh.ecell
    assert owned_offs[-1] == mesh.ncell and np.array_equal(np.sort(owned), np.arange(mesh.ncell))
    block_of = np.empty(mesh.ncell, np.int64)
    for b in range(nb):
        cells = owned[owned_offs[b]:owned_offs[b + 1]]
        assert np.all(np.diff(cells) > 0) and 0 < len(cells) <= block
        block_of[cells] = b
    seen = np.zeros((mesh.nedge, 2), np.int32)
    for b in range(nb):
        el = elems[elem_offs[b]:elem_offs[b + 1]]
        assert np.all(np.diff(el) > 0)
        # an edge is an instance of b iff one of its cells is owned by b
        touch = (block_of[ec[:, 0]] == b) | (block_of[ec[:, 1]] == b)
        assert np.array_equal(el, np.nonzero(touch)[0])
        for i in range(owned_offs[b], owned_offs[b + 1]):
            ent = csr[csr_offs[i]:csr_offs[i + 1]].astype(np.int64)
            j, side = ent >> 1, ent & 1
            e = el[j]
            assert np.all(np.diff(e * 2 + side) > 0)
            assert np.all(ec[e, side] == owned[i])
            seen[e, side] += 1
    assert np.all(seen == 1)
    assert abs(P.redundancy - tot / mesh.nedge) < 1e-12
    assert P.n_owned == mesh.ncell
    assert P.max_incs == int(np.max(np.diff(csr_offs))) == 4  # quads: four edges
    assert P.redundancy < 1.35


def test_owner_plan_refuses_unsupported_loops():
    """Loops the owner partition cannot run: increments through two maps, and
    loops with effects that evaluating a boundary element twice would repeat
    (a gbl reduction, a direct OP_RW) or that read the incremented dat."""
    import ctypes as C
    mesh = op2amd.mesh.generate(20, 10, 0, "structured")
    q, res = inputs(mesh, 0)
    app = App(mesh, q, res)
    try:
        R, I, RW = api.OP_READ, api.OP_INC, api.OP_RW
        d = app.d
        inc = lambda: d(app.res, 0, app.pecell, 4, "double", I)  # noqa: E731
        args = (api.op_arg * 2)(inc(), d(app.x, 0, app.pedge, 2, "double", I))
        assert not api.op2h_plan_owner_for(b"two_maps", app.edges, 64, 2, args)
        total = np.zeros(1)
        gbl = api.op_arg_gbl_char(total.ctypes.data_as(C.c_void_p), 1, b"double", 8, I)
        args = (api.op_arg * 2)(inc(), gbl)
        assert not api.op2h_plan_owner_for(b"with_gbl_inc", app.edges, 64, 2, args)
        ew = np.zeros(mesh.nedge)
        edat = api.decl_dat(app.edges, 1, "double", ew, "p_ew")
        args = (api.op_arg * 2)(inc(), d(edat, -1, None, 1, "double", RW))
        assert not api.op2h_plan_owner_for(b"with_direct_rw", app.edges, 64, 2, args)
        args = (api.op_arg * 2)(inc(), d(app.res, 1, app.pecell, 4, "double", R))
        assert not api.op2h_plan_owner_for(b"reads_res", app.edges, 64, 2, args)
        # a gbl read and a direct write are fine
        args = (api.op_arg * 3)(inc(), api.op_arg_gbl_char(total.ctypes.data_as(C.c_void_p), 1,
                                                          b"double", 8, R),
                                d(edat, -1, None, 1, "double", api.OP_WRITE))
        assert api.op2h_plan_owner_for(b"read_gbl_write_direct", app.edges, 64, 3, args)
    finally:
        app.close()


@pytest.mark.parametrize("ordering", ["random", "rcm"])
def test_runtime_plan_checks_pass_on_built_plans(ordering, capfd):
    """With OP_diags > 0 the runtime checks every plan it builds (and exits
    on a failure); at OP_diags > 6 each passing check prints OK."""
    import ctypes as C
    mesh = op2amd.mesh.generate(120, 60, 0, ordering)
    q, res = inputs(mesh, 0)
    app = App(mesh, q, res)
    diags = C.c_int.in_dll(api.lib, "OP_diags")
    try:
        diags.value = 7
        for on_dev in (0, 1):
            api.op2h_set_plan_builder(on_dev)
            for part in (64, 256):
                name = f"res_calc_{on_dev}_{part}".encode()
                assert api.op2h_plan_staged_for(name, app.edges, part, 8, res_args(app))
            assert api.op2h_plan_color_for(f"res_calc_{on_dev}".encode(), app.edges, 8,
                                           res_args(app))
        for block in (64, 256):
            assert api.op2h_plan_owner_for(b"res_calc", app.edges, block, 8, res_args(app))
        assert api.op2h_plan_gather(app.pecell)
    finally:
        diags.value = 0
        api.op2h_set_plan_builder(1)
        app.close()
    out = capfd.readouterr().out
    assert "error" not in out
    assert out.count("OP_plan_check: thread colouring OK") == 4
    assert out.count("OP_plan_check: block colouring OK") == 4
    assert out.count("OP_plan_check: element colouring OK") == 2
    assert out.count("OP_plan_check: increment completeness OK") == 2
    assert out.count("OP_plan_check: gather entries OK") == 1
