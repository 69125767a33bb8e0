"""Bytes moved per loop: OP_kernels[k].transfer, the figure op_timing_output
divides by the kernel time for its GB/s column (the reference's
op_timing_output, op_lib_core.cpp:1096-1165, prints the same column from the
plan's transfer), here the loop's algorithmic bytes: each map index read once
(4 bytes), each distinct dat element the loop touches once, twice for OP_RW
and OP_INC.

On an I x J airfoil mesh (Nc = IJ cells, Nn = (I+1)(J+1) nodes,
Ne = 2IJ - I - J interior edges, Nb = 2(I+J) boundary edges) that gives, per
call:
  save_soln  64 Nc                          q read, qold written
  adt_calc   56 Nc + 16 Nn                  pcell, every node, q, adt
  res_calc   16 Ne + 16 (Nn - 4) + 104 Nc   pedge, pecell; the four corner
                                            nodes lie on no interior edge
  bres_calc  32 Nb + 104 (Nb - 4)           pbedge, pbecell, bound, the Nb
                                            nodes of the boundary loop; the
                                            four corner cells hold two
                                            boundary edges each
  update     136 Nc                         qold, q, res (RW), adt
The CPU tests pin those counts on the meshes themselves; the GPU tests check
the runtime's figures against the closed forms exactly, and the GB/s column
of the translated program's op_timing_output (OP_HIP_KERNEL_TIMING) against
count x bytes / kernel time."""
import subprocess

import numpy as np
import pytest

import op2amd
from conftest import REPO

SIZES = [(24, 12), (37, 23), (120, 60)]
CALLS = {"save_soln": 1, "adt_calc": 2, "res_calc": 2, "bres_calc": 2, "update": 2}


def closed_form(I, J):
    nc, nn = I * J, (I + 1) * (J + 1)
    ne, nb = 2 * I * J - I - J, 2 * (I + J)
    return {"save_soln": 64 * nc, "adt_calc": 56 * nc + 16 * nn,
            "res_calc": 16 * ne + 16 * (nn - 4) + 104 * nc,
            "bres_calc": 32 * nb + 104 * (nb - 4), "update": 136 * nc}


@pytest.mark.parametrize("I,J", SIZES)
def test_mesh_counts_behind_the_closed_forms(I, J):
    m = op2amd.mesh.generate(I, J, 0, "random")
    assert (m.ncell, m.nnode, m.nedge, m.nbedge) == (I * J, (I + 1) * (J + 1),
                                                     2 * I * J - I - J, 2 * (I + J))
    assert np.unique(m.cell).size == m.nnode
    assert np.unique(m.edge).size == m.nnode - 4
    assert np.unique(m.ecell).size == m.ncell
    assert np.unique(m.bedge).size == m.nbedge
    assert np.unique(m.becell).size == m.nbedge - 4


@pytest.mark.gpu
@pytest.mark.parametrize("I,J", SIZES)
def test_runtime_transfer_equals_closed_form(I, J):
    m = op2amd.mesh.generate(I, J, 0, "random")
    t = op2amd.airfoil.run(m, 2, mode="tuned", kernel_timing=True)
    want = closed_form(I, J)
    for lp in CALLS:
        assert t["loop_bytes"][lp] == want[lp], (lp, t["loop_bytes"][lp], want[lp])


@pytest.mark.gpu
def test_timing_output_gbs_is_transfer_over_kernel_time(tmp_path):
    I, J, niter = 240, 120, 100
    exe = REPO / "tests" / "apps" / "_build" / "airfoil_tr" / "airfoil_tr"
    assert exe.exists(), f"{exe} missing: run __graft_entry__.build()"
    grid = tmp_path / "new_grid.dat"
    op2amd.mesh.generate(I, J, 0, "rcm").write(str(grid))
    print("running airfoil_tr", I, J, niter, "OP_HIP_KERNEL_TIMING", flush=True)
    r = subprocess.run([str(exe), str(grid), str(niter), "OP_HIP_KERNEL_TIMING"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = {}
    for line in r.stdout.splitlines():
        f = [v.strip() for v in line.split(";")]
        if len(f) == 6 and f[0].isdigit():
            rows[f[5]] = (int(f[0]), float(f[3]), float(f[4]))
    want = closed_form(I, J)
    assert set(rows) == set(CALLS)
    for lp, (count, kms, gbs) in rows.items():
        assert count == CALLS[lp] * niter
        assert kms > 0.0
        expect = count * want[lp] / (kms * 1e-3) / 1e9
        # kernel ms printed to 4 decimals, GB/s to 2; transfer summed in float
        assert abs(gbs - expect) <= 1e-3 * expect + 0.01, (lp, gbs, expect)
