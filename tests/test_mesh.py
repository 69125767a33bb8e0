"""The synthetic mesh generator: sizes, orientation conventions of
naca0012.m:217-303, orderings, and the RCM renumbering rule."""
import hashlib

import numpy as np
import pytest

import op2amd
from conftest import golden
from oracle import oracle


def centroids(m):
    return m.x[m.cell].mean(axis=1)


@pytest.mark.parametrize("I,J", [(2, 2), (7, 3), (24, 12), (120, 60)])
def test_sizes(I, J):
    m = op2amd.mesh.generate(I, J, 0, "random")
    assert (m.nnode, m.ncell) == ((I + 1) * (J + 1), I * J)
    assert m.nedge == (I - 1) * J + I * (J - 1)
    assert m.nbedge == 2 * I + 2 * J


@pytest.mark.parametrize("ordering", ["structured", "random", "rcm", "hilbert"])
def test_orientation_conventions(ordering):
    m = op2amd.mesh.generate(30, 15, 4, ordering)
    x, cen = m.x, centroids(m)
    # cells counter-clockwise
    xc = x[m.cell]
    area = 0.5 * np.sum(xc[:, :, 0] * np.roll(xc[:, :, 1], -1, axis=1)
                        - np.roll(xc[:, :, 0], -1, axis=1) * xc[:, :, 1], axis=1)
    assert np.all(area > 0)
    # interior normals (dy,-dx), (dx,dy)=x0-x1, point from ecell[0] to ecell[1]
    d = x[m.edge[:, 0]] - x[m.edge[:, 1]]
    nrm = np.stack([d[:, 1], -d[:, 0]], axis=1)
    assert np.all(np.sum(nrm * (cen[m.ecell[:, 1]] - cen[m.ecell[:, 0]]), axis=1) > 0)
    # boundary normals point out of the domain
    d = x[m.bedge[:, 0]] - x[m.bedge[:, 1]]
    nrm = np.stack([d[:, 1], -d[:, 0]], axis=1)
    mid = 0.5 * (x[m.bedge[:, 0]] + x[m.bedge[:, 1]])
    assert np.all(np.sum(nrm * (mid - cen[m.becell]), axis=1) > 0)
    # bound codes: walls y=0,1 -> 1; far field x=0,2 -> 2
    wall = np.isclose(mid[:, 1], 0) | np.isclose(mid[:, 1], 1)
    assert np.all(m.bound[wall] == 1) and np.all(m.bound[~wall] == 2)
    # every cell is closed: outward normals of its faces sum to zero
    acc = np.zeros((m.ncell, 2))
    d = x[m.edge[:, 0]] - x[m.edge[:, 1]]
    nrm = np.stack([d[:, 1], -d[:, 0]], axis=1)
    np.add.at(acc, m.ecell[:, 0], nrm)
    np.add.at(acc, m.ecell[:, 1], -nrm)
    d = x[m.bedge[:, 0]] - x[m.bedge[:, 1]]
    np.add.at(acc, m.becell, np.stack([d[:, 1], -d[:, 0]], axis=1))
    assert np.max(np.abs(acc)) < 1e-15


def test_orderings_are_relabelings():
    a = op2amd.mesh.generate(20, 10, 9, "structured")
    for ordering in ("random", "rcm", "hilbert"):
        b = op2amd.mesh.generate(20, 10, 9, ordering)
        ka = np.sort(np.round(centroids(a), 12).view(np.complex128).ravel())
        kb = np.sort(np.round(centroids(b), 12).view(np.complex128).ravel())
        assert np.array_equal(ka, kb)


def test_rcm_equals_independent_restatement():
    for I, J, seed in [(24, 12, 0), (40, 20, 7)]:
        m = op2amd.mesh.generate(I, J, seed, "random")
        expect = oracle.rcm_order(m.ncell, m.ecell)
        perms = m.renumber_rcm()
        assert np.array_equal(perms[0], expect)
        # edges: stable sort by new ecell[0]; nodes: first visit through cell->node
        assert np.all(np.diff(m.ecell[:, 0]) >= 0)
        first = []
        seen = set()
        for row in m.cell:
            for v in row:
                if v not in seen:
                    seen.add(v)
                    first.append(v)
        assert first[:len(first)] == sorted(first)


def test_rcm_reduces_bandwidth():
    rnd = op2amd.mesh.generate(120, 60, 0, "random")
    rcm = op2amd.mesh.generate(120, 60, 0, "rcm")
    assert rcm.bandwidth()[0] * 20 < rnd.bandwidth()[0]
    assert rcm.bandwidth()[0] <= 2 * 61


@pytest.mark.parametrize("name", ["airfoil_120x60_random_a1", "airfoil_120x60_rcm_a1",
                                  "airfoil_24x12_random_a1"])
def test_generator_pinned_by_fixture_checksum(name, tmp_path):
    g = golden(name)
    m = op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))
    m.write(tmp_path / "g.dat")
    assert hashlib.sha256((tmp_path / "g.dat").read_bytes()).hexdigest() == str(g["grid_sha256"])


def test_write_read_round_trip(tmp_path):
    m = op2amd.mesh.generate(13, 7, 2, "rcm")
    m.write(tmp_path / "g.dat")
    r = op2amd.mesh.read(tmp_path / "g.dat")
    for f in ("x", "cell", "edge", "ecell", "bedge", "becell", "bound"):
        assert np.array_equal(getattr(m, f), getattr(r, f)), f
