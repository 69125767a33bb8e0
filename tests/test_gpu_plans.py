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
