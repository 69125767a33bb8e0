"""Reverse Cuthill-McKee built on the device equals the host build bit for
bit (same permutation), with the packed one-pass level sort and the two-pass
form, on small graphs of every shape and on the 720k-cell mesh; op_renumber
gives the same numbering with either build."""
import time

import numpy as np
import pytest

import op2amd
import op2amd.api as api
from test_rcm import CASES, csr, grid_edges, random_graph

pytestmark = pytest.mark.gpu


def same(off, adj):
    host = api.rcm_order(off, adj, 0)
    for mode in (1, 2):
        dev = api.rcm_order(off, adj, mode)
        assert np.array_equal(dev, host), mode
    return host


@pytest.mark.parametrize("name,n,make", CASES, ids=[c[0] for c in CASES])
def test_device_equals_host_small(name, n, make):
    print(f"rcm case {name} n={n}", flush=True)
    same(*csr(n, make()))


def test_device_equals_host_edge_cases():
    print("rcm edge cases", flush=True)
    assert api.rcm_order(np.zeros(1, np.int32), np.zeros(0, np.int32), 1).size == 0
    same(np.zeros(2, np.int32), np.zeros(0, np.int32))  # one vertex
    same(*csr(300, random_graph(300, 400, 11, loops=20)))  # self-loops
    # many small components and isolated vertices
    same(*csr(5000, random_graph(5000, 1500, 12)))
    # a long path: one vertex per level
    same(*csr(4000, [(i, i + 1) for i in range(3999)]))


@pytest.mark.parametrize("I,J,seed", [(300, 200, 1), (64, 900, 2)])
def test_device_equals_host_grids(I, J, seed):
    print(f"rcm grid {I}x{J}", flush=True)
    same(*csr(I * J, grid_edges(I, J, seed)))


def test_device_equals_host_full_mesh():
    """The 1200x600 mesh (720k cells), randomly numbered: same permutation."""
    mesh = op2amd.mesh.generate(1200, 600, 0, "random")
    e = np.sort(mesh.ecell.astype(np.int64), axis=1)
    n = mesh.ncell
    # CSR of the cell graph without Python loops
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    key = np.unique(src * n + dst)
    src, dst = key // n, key % n
    off = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(src, minlength=n), out=off[1:])
    adj = dst.astype(np.int32)
    t0 = time.perf_counter()
    host = api.rcm_order(off, adj, 0)
    t1 = time.perf_counter()
    dev = api.rcm_order(off, adj, 1)
    t2 = time.perf_counter()
    print(f"720k-cell RCM: host {t1 - t0:.3f} s, device {t2 - t1:.3f} s", flush=True)
    assert np.array_equal(dev, host)
    assert np.array_equal(np.sort(dev), np.arange(n))


@pytest.mark.parametrize("where", ["host", "gpu"])
def test_op_renumber_either_build(where, monkeypatch):
    """op_renumber with the base ordering built on the host or the device
    gives the generator's RCM numbering."""
    from test_gpu_loops import App, inputs
    monkeypatch.setenv("OP_HIP_RCM", where)
    rnd = op2amd.mesh.generate(60, 30, 6, "random")
    rcm = op2amd.mesh.generate(60, 30, 6, "rcm")
    q, res = inputs(rnd, 0)
    app = App(rnd, q, res)
    try:
        api.lib.op_renumber.argtypes = [api.op_map]
        api.lib.op_renumber(app.pecell)
        x = app.fetch(app.x, (rnd.nnode, 2))
        ecell = np.ctypeslib.as_array(app.pecell.contents.map,
                                      shape=(rnd.nedge * 2,)).reshape(rnd.nedge, 2).copy()
    finally:
        app.close()
    assert np.array_equal(x, rcm.x) and np.array_equal(ecell, rcm.ecell)
