"""Host-side owner partition (op2h_owner_partition): complete, bounded,
connected and compact whatever the numbering of the mesh."""
import numpy as np
import pytest

import op2amd
import op2amd.api as api


def connected(cells, ecell, block_of, b):
    """Breadth-first search inside block b reaches all of its cells."""
    inside = (block_of[ecell[:, 0]] == b) & (block_of[ecell[:, 1]] == b)
    e = ecell[inside]
    adj = {c: [] for c in cells}
    for a, c in e:
        adj[a].append(c)
        adj[c].append(a)
    seen, stack = {cells[0]}, [cells[0]]
    while stack:
        for n in adj[stack.pop()]:
            if n not in seen:
                seen.add(n)
                stack.append(n)
    return len(seen) == len(cells)


@pytest.mark.parametrize("ordering", ["structured", "random", "rcm", "hilbert"])
@pytest.mark.parametrize("block", [64, 256])
def test_partition_complete_bounded_compact(ordering, block):
    mesh = op2amd.mesh.generate(160, 80, 5, ordering)
    nb, bo = api.owner_partition(mesh.ncell, mesh.ecell, [0, 1], block)
    assert bo.min() == 0 and bo.max() == nb - 1
    sizes = np.bincount(bo, minlength=nb)
    # every block is full but the last
    assert nb == -(-mesh.ncell // block)
    assert np.all(sizes[:-1] == block) and 1 <= sizes[-1] <= block
    # compact patches: few edges cross blocks (contiguous index ranges of an
    # RCM or random numbering would cut most edges)
    cross = np.mean(bo[mesh.ecell[:, 0]] != bo[mesh.ecell[:, 1]])
    assert cross < (0.25 if block == 64 else 0.15), cross
    # most blocks are single connected patches (a block that runs out of
    # reachable cells continues from the next seed)
    picks = range(0, nb, max(1, nb // 16))
    ok = sum(connected(np.nonzero(bo == b)[0], mesh.ecell, bo, b) for b in picks)
    assert ok >= 0.6 * len(picks)


def test_partition_is_deterministic_and_numbering_independent_in_quality():
    m1 = op2amd.mesh.generate(120, 60, 1, "rcm")
    a = api.owner_partition(m1.ncell, m1.ecell, [0, 1], 128)
    b = api.owner_partition(m1.ncell, m1.ecell, [0, 1], 128)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])


def test_partition_unreferenced_targets_and_single_increment():
    # 6 targets, elements touch 0..3 only; one increment per element: blocks
    # are runs of consecutive referenced targets
    table = np.array([[0, 5], [1, 5], [2, 5], [3, 5]], np.int32)
    nb, bo = api.owner_partition(6, table, [0], 2)
    assert nb == 2 and list(bo) == [0, 0, 1, 1, -1, -1]
    nb, bo = api.owner_partition(6, table, [0], 3)
    assert nb == 2 and list(bo) == [0, 0, 0, 1, -1, -1]
    with pytest.raises(ValueError):
        api.owner_partition(6, table, [2], 2)
