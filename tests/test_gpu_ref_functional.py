"""The reference's own functional tests (tests/functional/{gbl,dat_reductions,
idx,strides,const}/*.cpp), unmodified: translated by op2amd.translator and
linked against libop2hip.so by tests/apps/build.py where the reference tree
is present (only the executables are kept), then run on the GPU.  Each
program checks its own results against closed forms (TOL 1e-9), prints one
"<check> passed" line per check and exits nonzero at the first failure; it
calls op_init with diagnostics level 2, so every plan the runtime builds is
also checked on the host.  Run with the default plans, every dat in SoA
layout, coloured and atomic increments, and under transparent renumbering."""
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu
BUILD = REPO / "tests" / "apps" / "_build"

# checks each program runs (its "passed" lines)
CHECKS = {"ref_gbl": 8, "ref_reduc": 4, "ref_idx": 3, "ref_strides": 2, "ref_const": 2}


@pytest.mark.parametrize("variant", [[], ["OP_AUTO_SOA"], ["OP_HIP_PLAN=*:color"],
                                     ["OP_HIP_PLAN=*:atomic"], ["OP_HIP_RENUMBER"]],
                         ids=["default", "soa", "color", "atomic", "renumbered"])
@pytest.mark.parametrize("name", sorted(CHECKS))
def test_reference_functional_test(name, variant, tmp_path):
    import subprocess
    exe = BUILD / name / name
    assert exe.exists(), f"{exe} missing: run __graft_entry__.build() where the reference is present"
    print("running", name, *variant, flush=True)
    r = subprocess.run([str(exe), *variant], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    passed = [l for l in r.stdout.splitlines() if l.rstrip().split(" [rank")[0].endswith("passed")]
    assert len(passed) == CHECKS[name], r.stdout[-2000:]
    assert "ERROR" not in r.stdout
    assert "OP_plan_check" not in r.stdout
