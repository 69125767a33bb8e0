"""Reverse Cuthill-McKee (op2h_rcm_order) on the host: equal to the oracle's
queue formulation, and that queue formulation equal to the level-synchronous
one the device build uses (each BFS level sorted by (position of the first
parent, degree, index)), on graphs with ties, several components, isolated
vertices and self-loops."""
import numpy as np
import pytest

import op2amd
import op2amd.api as api
from oracle import oracle


def csr(n, edges):
    """Strictly ascending neighbour lists of an undirected edge list."""
    nbr = [set() for _ in range(n)]
    for a, b in edges:
        nbr[a].add(b)
        nbr[b].add(a)
    off = np.zeros(n + 1, np.int32)
    for v in range(n):
        off[v + 1] = off[v] + len(nbr[v])
    adj = np.array([w for v in range(n) for w in sorted(nbr[v])], np.int32)
    return off, adj


def random_graph(n, m, seed, loops=0):
    rng = np.random.default_rng(seed)
    e = set()
    while len(e) < m:
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b:
            e.add((min(a, b), max(a, b)))
    e = sorted(e)
    e += [(int(v), int(v)) for v in rng.choice(n, loops, replace=False)] if loops else []
    return e


def grid_edges(I, J, seed):
    """I x J grid graph with vertices numbered by a seeded permutation (many
    degree ties)."""
    p = np.random.default_rng(seed).permutation(I * J)
    e = []
    for i in range(I):
        for j in range(J):
            v = p[i * J + j]
            if i + 1 < I:
                e.append((int(v), int(p[(i + 1) * J + j])))
            if j + 1 < J:
                e.append((int(v), int(p[i * J + j + 1])))
    return e


def level_sync_rcm(off, adj):
    """The device build's rule, restated: roots and pseudo-peripheral search as
    the queue rule; each new BFS level is every unplaced neighbour of the
    current level, ordered by (position of its first-placed neighbour in the
    current level, degree, index)."""
    n = len(off) - 1
    deg = np.diff(off)
    pos = -np.ones(n, np.int64)

    def nbrs(v):
        return adj[off[v]:off[v + 1]]

    def bfs(root):
        level = {root: 0}
        frontier, last, depth = [root], [root], 0
        while True:
            nxt = []
            for v in frontier:
                for w in nbrs(v):
                    w = int(w)
                    if pos[w] < 0 and w not in level:
                        level[w] = depth + 1
                        nxt.append(w)
            if not nxt:
                return depth, last
            last, frontier, depth = nxt, nxt, depth + 1

    placed = 0
    order = sorted(range(n), key=lambda v: (deg[v], v))
    for root in order:
        if pos[root] >= 0:
            continue
        ecc, last = bfs(root)
        for _ in range(16):
            cand = min(last, key=lambda v: (deg[v], v))
            ecc2, last2 = bfs(cand)
            if ecc2 <= ecc:
                break
            root, ecc, last = cand, ecc2, last2
        pos[root] = placed
        placed += 1
        frontier = [root]
        while frontier:
            parent = {}
            for v in frontier:
                for w in nbrs(v):
                    w = int(w)
                    if pos[w] < 0:
                        parent[w] = min(parent.get(w, n), int(pos[v]))
            nxt = sorted(parent, key=lambda w: (parent[w], deg[w], w))
            for w in nxt:
                pos[w] = placed
                placed += 1
            frontier = nxt
    return (n - 1 - pos).astype(np.int32)


def oracle_rcm(n, edges):
    """The oracle's queue rule (it takes edges without self-loops)."""
    return np.asarray(oracle.rcm_order(n, np.array([e for e in edges if e[0] != e[1]],
                                                   np.int64).reshape(-1, 2)), np.int32)


def bandwidth(new_of_old, edges):
    return max((abs(int(new_of_old[a]) - int(new_of_old[b])) for a, b in edges), default=0)


CASES = [
    ("random sparse", 300, lambda: random_graph(300, 330, 1)),
    ("random dense", 120, lambda: random_graph(120, 900, 2)),
    ("forest + isolated", 400, lambda: random_graph(400, 150, 3)),
    ("grid", 24 * 17, lambda: grid_edges(24, 17, 4)),
    ("path", 50, lambda: [(i, i + 1) for i in range(49)]),
    ("star", 40, lambda: [(7, i) for i in range(40) if i != 7]),
    ("none", 9, lambda: []),
]


@pytest.mark.parametrize("name,n,make", CASES, ids=[c[0] for c in CASES])
def test_host_rcm_matches_oracle_and_level_rule(name, n, make):
    edges = make()
    off, adj = csr(n, edges)
    host = api.rcm_order(off, adj, 0)
    assert sorted(host.tolist()) == list(range(n))
    assert np.array_equal(host, oracle_rcm(n, edges))
    assert np.array_equal(host, level_sync_rcm(off, adj))


def test_self_loops_do_not_change_the_rule():
    """A self-loop adds one to a vertex's degree and nothing else; the
    level-synchronous rule agrees with the host on such graphs."""
    edges = random_graph(200, 260, 9, loops=12)
    off, adj = csr(200, edges)
    assert np.array_equal(api.rcm_order(off, adj, 0), level_sync_rcm(off, adj))


def test_mesh_cell_graph_and_bandwidth():
    """On the generator's randomly numbered mesh: the host build equals the
    oracle on the cell graph, and cuts the bandwidth by far."""
    mesh = op2amd.mesh.generate(40, 20, 5, "random")
    edges = [(int(a), int(b)) for a, b in mesh.ecell]
    off, adj = csr(mesh.ncell, edges)
    host = api.rcm_order(off, adj, 0)
    assert np.array_equal(host, oracle_rcm(mesh.ncell, edges))
    ident = np.arange(mesh.ncell)
    assert bandwidth(host, edges) * 10 < bandwidth(ident, edges)


@pytest.mark.parametrize("bad", ["unsorted", "duplicate", "range", "offsets", "mode"])
def test_bad_arguments(bad):
    off, adj = csr(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    mode = 0
    if bad == "unsorted":
        adj[1:3] = adj[2:0:-1]
    elif bad == "duplicate":
        adj[2] = adj[1]
    elif bad == "range":
        adj[0] = 5
    elif bad == "offsets":
        off[2] = off[1] - 1
    else:
        mode = 3
    with pytest.raises(ValueError):
        api.rcm_order(off, adj, mode)


def test_empty_graph():
    assert api.rcm_order(np.zeros(1, np.int32), np.zeros(0, np.int32), 0).size == 0
