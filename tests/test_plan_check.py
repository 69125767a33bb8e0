"""The runtime's plan checks (op_plan_check, op_rt_support.cpp:129-295, and
the colouring checks its stub leaves empty at :290-294) on the host, without
a GPU: plans made by the oracle's restatement of the reference's
op_plan_core pass; each planted fault (a thread- or block-colour conflict, a
block listed twice, a wrong local id, an unsorted staging list, a missing or
misordered owner increment, a foreign gather entry) is reported with the
reference's message.  The runtime runs these checks on every plan it builds
when OP_diags > 0 and exits with -1 on a failure."""
import ctypes as C

import numpy as np
import pytest

import op2amd
import op2amd.api as api
from oracle import oracle

P_INT = C.POINTER(C.c_int)


def iptr(a):
    return a.ctypes.data_as(P_INT)


class Staged:
    """Host image of the reference's plan for res_calc (oracle.plan_build),
    as op2h_check_staged_plan reads it: one map slot per indirect dat (ind),
    one loc slot per indirect arg."""

    def __init__(self, mesh, part):
        self.n, self.args = oracle.airfoil_plan_args(mesh, "res_calc")
        self.p = oracle.plan_build(self.n, self.args, part)
        p = self.p
        self.keep = []
        self.nb, self.ni = p["nblocks"], p["ninds"]
        self.colstart = np.concatenate([[0], np.cumsum(p["ncolblk"])]).astype(np.int32)
        self.blkmap = p["blkmap"].astype(np.int32)
        self.offset = p["offset"].astype(np.int32)
        self.nelems = p["nelems"].astype(np.int32)
        self.nthrcol = p["nthrcol"].astype(np.int32)
        self.thrcol = p["thrcol"].astype(np.uint8)
        io = p["ind_offs"].reshape(self.nb, self.ni)
        isz = p["ind_sizes"].reshape(self.nb, self.ni)
        self.ind_offs = [np.ascontiguousarray(io[:, i], np.int32) for i in range(self.ni)]
        self.ind_sizes = [np.ascontiguousarray(isz[:, i], np.int32) for i in range(self.ni)]
        self.ind_maps = [np.ascontiguousarray(m, np.int32) for m in p["ind_maps"]]
        self.tables = {}
        for a in self.args:
            if a["map"] is not None:
                self.tables[a["ind"]] = np.ascontiguousarray(a["map"], np.int32)
        self.locs = [(a["ind"], a["idx"], np.ascontiguousarray(p["loc_maps"][i], np.uint16))
                     for i, a in enumerate(self.args) if a["map"] is not None]

    def host(self):
        h = api.op2h_staged_plan_host()
        h.set_size, h.nblocks, h.ncolors = self.n, self.nb, len(self.colstart) - 1
        h.colstart, h.blkmap = iptr(self.colstart), iptr(self.blkmap)
        h.offset, h.nelems, h.nthrcol = iptr(self.offset), iptr(self.nelems), iptr(self.nthrcol)
        h.thrcol = self.thrcol.ctypes.data_as(C.POINTER(C.c_ubyte))
        h.nmaps = self.ni
        for i in range(self.ni):
            t = self.tables[i]
            h.map_table[i], h.map_dim[i] = iptr(t), t.shape[1]
            h.map_to_size[i] = [a for a in self.args if a.get("ind") == i][0]["to_size"]
            h.ind_offs[i], h.ind_sizes[i] = iptr(self.ind_offs[i]), iptr(self.ind_sizes[i])
            h.ind_map[i] = iptr(self.ind_maps[i])
        h.nlocs = len(self.locs)
        for l, (ind, idx, lm) in enumerate(self.locs):
            h.loc_map_slot[l], h.loc_idx[l] = ind, idx
            h.loc_map[l] = lm.ctypes.data_as(C.POINTER(C.c_ushort))
        return h

    def accesses(self):
        acc = (api.op2h_check_access * len(self.locs))()
        for l, (ind, idx, _) in enumerate(self.locs):
            t = self.tables[ind]
            a = [x for x in self.args if x.get("ind") == ind and x["idx"] == idx][0]
            acc[l].table, acc[l].dim, acc[l].idx = iptr(t), t.shape[1], idx
            acc[l].to_size, acc[l].dat, acc[l].writes = a["to_size"], ind, int(a["acc"] == 3)
        return acc

    def check(self):
        acc = self.accesses()
        return api.op2h_check_staged_plan(C.byref(self.host()), len(acc), acc)


@pytest.fixture(scope="module")
def mesh():
    return op2amd.mesh.generate(60, 30, 0, "random")


@pytest.mark.parametrize("part", [0, 64])
def test_reference_plans_pass(mesh, part, capfd):
    s = Staged(mesh, part)
    assert len(s.colstart) > 2  # several block colours: the block check bites
    assert s.check() == 0
    assert "OP_plan_check" not in capfd.readouterr().out


def res_targets(s, e):
    return set(int(x) for x in s.tables[3][e])


def test_thread_colour_conflict_reported(mesh, capfd):
    s = Staged(mesh, 64)
    b = int(s.blkmap[0])
    lo, hi = s.offset[b], s.offset[b] + s.nelems[b]
    pair = next((e1, e2) for e1 in range(lo, hi) for e2 in range(e1 + 1, hi)
                if res_targets(s, e1) & res_targets(s, e2))
    assert s.thrcol[pair[0]] != s.thrcol[pair[1]]
    s.thrcol[pair[1]] = s.thrcol[pair[0]]
    assert s.check() == 1
    assert "OP_plan_check: thread colouring error" in capfd.readouterr().out


def test_block_colour_conflict_reported(mesh, capfd):
    s = Staged(mesh, 64)
    s.colstart = np.array([0, s.nb], np.int32)  # every block in one colour
    assert s.check() == 1
    assert "OP_plan_check: block colouring error" in capfd.readouterr().out


def test_block_listed_twice_reported(mesh, capfd):
    s = Staged(mesh, 64)
    s.blkmap[1] = s.blkmap[0]
    assert s.check() >= 1
    assert "OP_plan_check: blkmap error" in capfd.readouterr().out


def test_wrong_local_id_reported(mesh, capfd):
    s = Staged(mesh, 64)
    lm = s.locs[6][2]  # res through pecell, idx 0
    b = int(s.blkmap[0])
    e = int(s.offset[b])
    lm[e] = (int(lm[e]) + 1) % int(s.ind_sizes[3][b])
    assert s.check() == 1
    assert "OP_plan_check: 1 maps error(s)" in capfd.readouterr().out


def test_unsorted_staging_list_reported(mesh, capfd):
    s = Staged(mesh, 64)
    o = int(s.ind_offs[1][3])
    s.ind_maps[1][o], s.ind_maps[1][o + 1] = s.ind_maps[1][o + 1], s.ind_maps[1][o]
    assert s.check() >= 1
    assert "OP_plan_check: ind_maps error" in capfd.readouterr().out


def test_element_total_reported(mesh, capfd):
    s = Staged(mesh, 64)
    s.nelems[-1] -= 1
    assert s.check() >= 1
    out = capfd.readouterr().out
    assert "OP_plan_check: nelems error" in out


def flat_colouring(mesh):
    """A flat colouring of res_calc: the reference's greedy over one block
    spanning the set (OP_COLOR2), grouped by colour."""
    n, args = oracle.airfoil_plan_args(mesh, "res_calc")
    p = oracle.plan_build(n, args, n)
    col = p["thrcol"].astype(np.int32)
    nc = int(col.max()) + 1
    offs = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=nc))]).astype(np.int32)
    reord = np.argsort(col, kind="stable").astype(np.int32)
    ecell = np.ascontiguousarray(mesh.ecell, np.int32)
    acc = (api.op2h_check_access * 2)()
    for k in range(2):
        acc[k].table, acc[k].dim, acc[k].idx = iptr(ecell), 2, k
        acc[k].to_size, acc[k].dat, acc[k].writes = mesh.ncell, 0, 1
    return n, nc, offs, reord, col, acc, ecell


def test_flat_colouring_checked(mesh, capfd):
    n, nc, offs, reord, col, acc, ecell = flat_colouring(mesh)
    assert api.op2h_check_color_plan(n, nc, iptr(offs), iptr(reord), 2, acc) == 0
    # move one element into the colour of an edge sharing its cell
    e1 = int(reord[0])
    e2 = next(e for e in range(n) if e != e1 and set(ecell[e]) & set(ecell[e1]))
    c2 = col[e2]
    bad = np.concatenate([np.delete(reord, 0)[:offs[c2] - 1], [e1], np.delete(reord, 0)[offs[c2] - 1:]])
    bad_offs = offs.copy()
    bad_offs[1:c2 + 1] -= 1
    bad = bad.astype(np.int32)
    assert api.op2h_check_color_plan(n, nc, iptr(bad_offs), iptr(bad), 2, acc) == 1
    assert "OP_plan_check: element colouring error" in capfd.readouterr().out


def owner_image(mesh, block):
    """Owner lists of res_calc's increments through pecell (both columns)
    from the runtime's host partition, built here in numpy: owned cells and
    their edges ascending, per owned cell its (instance * 2 + k) entries in
    edge order."""
    ecell = np.ascontiguousarray(mesh.ecell, np.int32)
    nb, block_of = api.owner_partition(mesh.ncell, ecell, [0, 1], block)
    owned_offs, owned, elem_offs, elems, csr_offs, csr = [0], [], [0], [], [0], []
    for b in range(nb):
        cells = np.flatnonzero(block_of == b)
        inst = np.flatnonzero((block_of[ecell[:, 0]] == b) | (block_of[ecell[:, 1]] == b))
        local = {int(e): i for i, e in enumerate(inst)}
        for c in cells:
            ents = sorted(local[int(e)] * 2 + k for k in (0, 1)
                          for e in np.flatnonzero(ecell[:, k] == c))
            csr.extend(ents)
            csr_offs.append(len(csr))
        owned.extend(cells)
        elems.extend(inst)
        owned_offs.append(len(owned))
        elem_offs.append(len(elems))
    arr = {k: np.array(v, np.int32) for k, v in dict(owned_offs=owned_offs, owned=owned,
                                                       elem_offs=elem_offs, elems=elems,
                                                       csr_offs=csr_offs).items()}
    arr["csr"] = np.array(csr, np.uint16)
    return nb, ecell, arr


def owner_check(mesh, nb, ecell, arr, csr4=None):
    h = api.op2h_owner_plan_host()
    h.set_size, h.to_size, h.nblocks, h.ninc = mesh.nedge, mesh.ncell, nb, 2
    h.map_table, h.map_dim = iptr(ecell), 2
    h.inc_idx[0], h.inc_idx[1] = 0, 1
    for k in ("owned_offs", "owned", "elem_offs", "elems", "csr_offs"):
        setattr(h, k, iptr(arr[k]))
    h.csr = arr["csr"].ctypes.data_as(C.POINTER(C.c_ushort))
    if csr4 is not None:
        h.csr4 = csr4.ctypes.data_as(C.POINTER(C.c_ushort))
    return api.op2h_check_owner_plan(C.byref(h))


def test_owner_plan_checked(capfd):
    mesh = op2amd.mesh.generate(24, 12, 0, "random")
    nb, ecell, arr = owner_image(mesh, 32)
    c4 = np.full((mesh.ncell, 4), 0xFFFF, np.uint16)
    for i in range(mesh.ncell):
        ents = arr["csr"][arr["csr_offs"][i]:arr["csr_offs"][i + 1]]
        c4[i, :len(ents)] = ents
    assert owner_check(mesh, nb, ecell, arr, c4) == 0
    assert "OP_plan_check" not in capfd.readouterr().out
    # two increments of one cell in the wrong order
    bad = {k: v.copy() for k, v in arr.items()}
    j = int(bad["csr_offs"][5])
    bad["csr"][j], bad["csr"][j + 1] = bad["csr"][j + 1], bad["csr"][j]
    assert owner_check(mesh, nb, ecell, bad) == 1
    assert "OP_plan_check: increment lists error" in capfd.readouterr().out
    # the fixed-width copy disagreeing with the list
    c4b = c4.copy()
    c4b[7, 0] ^= 2
    assert owner_check(mesh, nb, ecell, arr, c4b) == 1
    assert "OP_plan_check: increment lists error" in capfd.readouterr().out
    # a cell owned by two blocks
    bad = {k: v.copy() for k, v in arr.items()}
    b1 = int(bad["owned_offs"][1])
    bad["owned"][b1] = bad["owned"][0]
    assert owner_check(mesh, nb, ecell, bad) >= 1
    assert "OP_plan_check: owned targets error" in capfd.readouterr().out


def test_owner_plan_missing_increment_reported(capfd):
    mesh = op2amd.mesh.generate(24, 12, 0, "random")
    nb, ecell, arr = owner_image(mesh, 32)
    bad = {k: v.copy() for k, v in arr.items()}
    j = int(bad["csr_offs"][3])
    bad["csr"] = np.delete(bad["csr"], j)
    bad["csr_offs"][4:] -= 1
    assert owner_check(mesh, nb, ecell, bad) >= 1
    assert "OP_plan_check: increment completeness error" in capfd.readouterr().out


def test_gather_plan_checked(mesh, capfd):
    ecell = np.ascontiguousarray(mesh.ecell, np.int32)
    flat = ecell.ravel()
    order = np.lexsort((np.arange(flat.size), flat)).astype(np.int32)
    targets = np.unique(flat).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(np.bincount(flat, minlength=mesh.ncell)[targets])])
    offs = offs.astype(np.int32)
    args = (mesh.nedge, mesh.ncell, 2, iptr(flat), len(targets), iptr(offs))
    assert api.op2h_check_gather_plan(*args, iptr(order), iptr(targets), None) == 0
    bad = order.copy()
    bad[0], bad[offs[1]] = bad[offs[1]], bad[0]  # an entry under another cell
    assert api.op2h_check_gather_plan(*args, iptr(bad), iptr(targets), None) == 1
    assert "OP_plan_check: gather entries error" in capfd.readouterr().out
    # with an element key the entries must follow the key's order instead
    key = np.random.default_rng(3).permutation(mesh.nedge).astype(np.int32)
    keyed = np.concatenate([sorted(order[offs[i]:offs[i + 1]], key=lambda e: (key[e // 2], e % 2))
                            for i in range(len(targets))]).astype(np.int32)
    assert api.op2h_check_gather_plan(*args, iptr(keyed), iptr(targets), iptr(key)) == 0
    assert api.op2h_check_gather_plan(*args, iptr(order), iptr(targets), iptr(key)) == 1
    capfd.readouterr()
