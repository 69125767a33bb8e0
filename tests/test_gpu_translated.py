"""Programs translated by op2amd.translator and built by tests/apps/build.py
(run by __graft_entry__.build()), on the GPU:
- the functional programs (global reductions, dat increments, op_arg_idx,
  SoA strides, constants, temporary dats) with the default plans (owner partition for
  increments), with every dat in SoA layout (OP_AUTO_SOA), and with
  coloured or atomic increments;
- the translated airfoil program against the reference-pinned trajectory,
  eagerly on each plan and as graph launches of 9 captured steps (the rest of
  the 999 iterations after an eager one), rms/maxerr/errloc from the graph
  history; the final q read back from its op_print_dat_to_binfile dump.  The
  owner-partition plans (the default, and the graphed run) sum each cell's
  increments in the sequential order: their q and every iteration's maxerr
  and errloc are the reference's bits, also when the device works in its own
  RCM numbering (OP_HIP_RENUMBER) while the program keeps the mesh's.  Atomic and coloured increments sum
  in another order: q within 1e-10, and errloc wherever maxerr came out
  identical."""
import subprocess

import numpy as np
import pytest

import op2amd
from conftest import REPO, golden, pointwise_rel_err, rel_err

pytestmark = pytest.mark.gpu
BUILD = REPO / "tests" / "apps" / "_build"


def run(name, *args, timeout=120):
    exe = BUILD / name / name
    assert exe.exists(), f"{exe} missing: run __graft_entry__.build()"
    print("running", name, *args, flush=True)
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=timeout)
    return r


@pytest.mark.parametrize("variant", [[], ["OP_AUTO_SOA"], ["OP_HIP_PLAN=*:color"],
                                     ["OP_HIP_PLAN=*:atomic"], ["OP_HIP_RENUMBER"]],
                         ids=["default", "soa", "color", "atomic", "renumbered"])
@pytest.mark.parametrize("name", ["gbl", "reductions", "idx", "strides", "consts", "helpers",
                                  "tempdats"])
def test_functional_program(name, variant):
    r = run(name, *variant)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert f"{name}: all passed" in r.stdout


@pytest.mark.parametrize("variant", [[], ["OP_HIP_PLAN=*:atomic"], ["OP_HIP_PLAN=*:color"],
                                     ["OP_AUTO_SOA"], ["graph=10"], ["OP_HIP_RENUMBER=pecell"],
                                     ["graph=10", "OP_HIP_RENUMBER=pecell"]],
                         ids=["owner", "atomic", "color", "soa", "graphed", "renumbered",
                              "renumbered-graphed"])
def test_translated_airfoil_matches_reference_trajectory(variant, tmp_path):
    g = golden("airfoil_120x60_random_a1")
    mesh = op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))
    grid = tmp_path / "new_grid.dat"
    mesh.write(str(grid))
    dump = tmp_path / "q.bin"
    r = run("airfoil_tr", str(grid), str(int(g["niter"])), *variant, f"dump={dump}")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [l.split() for l in r.stdout.splitlines() if l and l[0].isdigit()]
    rms = np.array([float(x[1]) for x in rows])
    maxerr = np.array([float(x[2]) for x in rows])
    errloc = np.array([int(x[3]) for x in rows])
    assert len(rms) == int(g["niter"])
    assert pointwise_rel_err(rms, g["rms"]) <= 1e-10
    raw = dump.read_bytes()
    count, dim = np.frombuffer(raw[:8], dtype=np.int32)
    q = np.frombuffer(raw[8:], dtype=np.float64).reshape(count, dim)
    exact = not any(v.startswith(("OP_HIP_PLAN", "OP_AUTO_SOA")) for v in variant)
    if exact:
        assert np.array_equal(q, g["q"])
        assert np.array_equal(maxerr, g["maxerr"])  # printed with 17 digits: exact
        assert np.array_equal(errloc, g["errloc"])
    else:
        assert rel_err(q, g["q"]) <= 1e-10
        assert pointwise_rel_err(maxerr, g["maxerr"]) <= 1e-10
        same = maxerr == g["maxerr"]
        assert np.array_equal(errloc[same], g["errloc"][same])
    qline = [l.split() for l in r.stdout.splitlines() if l.startswith("q ")][0]
    assert abs(float(qline[1]) - q.sum()) <= 1e-12 * abs(q.sum())
