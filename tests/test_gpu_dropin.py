"""The reference's own airfoil application as a drop-in: its unmodified
source (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp), translated by
op2amd.translator and linked against libop2hip.so by tests/apps/build.py
(where the reference tree is present; only the executable is kept), run on
the GPU in a directory holding new_grid.dat, as the application expects
(airfoil.cpp:89).

Checked against the oracle and the reference's own output:
- out_grid_seq.bin (op_print_dat_to_binfile, airfoil.cpp:303) is bit for bit
  the file the reference's sequential build writes for the same mesh (the
  golden fixture made by running that build unmodified): owner-partition
  res_calc sums each cell's increments in the sequential order;
- the printed history (airfoil.cpp:285-286: every 100 iterations rms,
  maxerr and errloc) against the oracle's sequential run: errloc exactly,
  maxerr to 1e-14 relative, rms to the printed 6 digits;
- the op_profile tree the application asks for (airfoil.cpp:215,299,313) at
  OP_PROFILE_LEVEL=3, written to OP_PROFILE_JSON_OUTPUT: one loop node per
  op_par_loop with the application's call counts, each holding its kernels'
  device time.

Two more of the reference's airfoil applications run the same way,
unmodified: the single-precision one (airfoil_plain/sp) and the one whose
work arrays are temporary dats (airfoil_tempdats/dp); their printed
histories are checked against the reference's sequential builds of the same
sources on the same mesh."""
import json
import os
import subprocess

import numpy as np
import pytest

import op2amd
from conftest import REPO, golden
from oracle import oracle

pytestmark = pytest.mark.gpu
EXE = REPO / "tests" / "apps" / "_build" / "airfoil_ref" / "airfoil_ref"


def run_app(tmp_path, *tokens, env_extra=None):
    assert EXE.exists(), f"{EXE} missing: run __graft_entry__.build() where the reference is present"
    g = golden("airfoil_24x12_random_app_a0")
    mesh = op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))
    mesh.write(str(tmp_path / "new_grid.dat"))
    env = dict(os.environ)
    env.update(env_extra or {})
    print("running the reference airfoil app", *tokens, flush=True)
    r = subprocess.run([str(EXE), *tokens], cwd=tmp_path, env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return g, mesh, r.stdout


def run_variant(tmp_path, exe, golden_name, *tokens):
    """A reference app variant on its golden's mesh; its printed history."""
    path = REPO / "tests" / "apps" / "_build" / exe / exe
    assert path.exists(), f"{path} missing: run __graft_entry__.build() where the reference is present"
    g = golden(golden_name)
    mesh = op2amd.mesh.generate(int(g["I"]), int(g["J"]), 0, str(g["ordering"]))
    mesh.write(str(tmp_path / "new_grid.dat"))
    print("running", exe, *tokens, flush=True)
    r = subprocess.run([str(path), *tokens], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    rows = [l.split() for l in r.stdout.splitlines()
            if len(l.split()) == 2 and l.split()[0].isdigit()]
    return g, mesh, r.stdout, rows


def read_bin(path):
    raw = path.read_bytes()
    count, dim = np.frombuffer(raw[:8], dtype=np.int32)
    return np.frombuffer(raw[8:], dtype=np.float64).reshape(count, dim)


@pytest.mark.parametrize("tokens", [[], ["OP_AUTO_SOA"], ["OP_HIP_RENUMBER=pecell"]],
                         ids=["aos", "soa", "renumbered"])
def test_unmodified_reference_app_bit_identical(tmp_path, tokens):
    g, mesh, out = run_app(tmp_path, *tokens)
    q = read_bin(tmp_path / "out_grid_seq.bin")
    assert q.shape == g["q"].shape
    assert np.array_equal(q, g["q"]), float(np.max(np.abs(q - g["q"])))
    assert (tmp_path / "out_grid_seq.dat").exists()
    ref = oracle.airfoil_run(mesh, 1000, 0, "seq")
    rows = [l.split() for l in out.splitlines() if l.strip() and l.split()[0].isdigit()]
    assert [int(r[0]) for r in rows] == list(range(100, 1001, 100))
    for r in rows:
        it = int(r[0]) - 1
        assert int(r[-1]) == ref["errloc"][it]
        maxerr = float(r[4].rstrip(","))
        assert abs(maxerr - ref["maxerr"][it]) <= 1e-14 * ref["maxerr"][it]
        assert abs(float(r[1].rstrip(",")) - ref["rms"][it]) <= 1e-5 * ref["rms"][it]


def test_unmodified_reference_app_profile_tree(tmp_path):
    js = tmp_path / "timing.json"
    _, _, out = run_app(tmp_path, env_extra={"OP_PROFILE_LEVEL": "3",
                                             "OP_PROFILE_JSON_OUTPUT": str(js)})
    assert f"Writing timing JSON to: {js}" in out
    tree = json.loads(js.read_text())
    assert tree["name"] == "Airfoil" and tree["clock"]["n"] == 1
    loops = {c["name"]: c for c in tree["children"] if c["type"] == 1}
    calls = {"save_soln": 1000, "adt_calc": 2000, "res_calc": 2000, "bres_calc": 2000,
             "update": 2000}
    for loop, n in calls.items():
        (name,) = [k for k in loops if k.split("/")[0] == loop]
        node = loops[name]
        assert name.split("/")[1] == "hip"
        assert node["clock"]["n"] == n
        comp = [c for c in node["children"] if c["name"] == "Computation"][0]
        kern = [c for c in comp["children"] if c["name"] == "Kernel"][0]
        assert comp["clock"]["n"] == n and kern["clock"]["n"] == n
        assert 0 < kern["clock"]["total"] <= node["clock"]["total"]
    assert loops[[k for k in loops if k.startswith("save_soln/")][0]]["name"].endswith("/direct")
    total = sum(c["clock"]["total"] for c in tree["children"])
    assert total <= tree["clock"]["total"]
    assert "%kern" in out


@pytest.mark.parametrize("tokens", [[], ["OP_AUTO_SOA"]], ids=["aos", "soa"])
def test_reference_tempdats_app(tmp_path, tokens):
    """apps/c/airfoil/airfoil_tempdats/dp/airfoil.cpp, unmodified: res, adt and
    qold are temporary dats declared zeroed and freed every iteration
    (:225-307), the update reduces rms only.  Its printed history (rms every
    100 iterations, 6 significant digits) against the reference's sequential
    build of the same app on the same mesh (golden), and against the
    oracle's rms; no free reported an error."""
    g, mesh, out, rows = run_variant(tmp_path, "airfoil_ref_tempdats",
                                     "airfoil_120x60_random_app_tempdats", *tokens)
    assert "cannot be removed" not in out
    assert [int(r[0]) for r in rows] == list(g["iters"])
    rms = np.array([float(r[1]) for r in rows])
    assert np.all(np.abs(rms - g["rms"]) <= 1e-5 * g["rms"]), (rms, g["rms"])
    ref = oracle.airfoil_run(mesh, 1000, 0, "seq")
    assert np.all(np.abs(rms - ref["rms"][g["iters"] - 1]) <= 1e-5 * rms)


def test_reference_single_precision_app(tmp_path):
    """apps/c/airfoil/airfoil_plain/sp/airfoil.cpp, unmodified: float dats,
    constants and rms reduction.  On the reference's free stream the flow is
    steady and the printed rms is float rounding noise, which after 1000
    iterations reproduces only if every step's float arithmetic does: the
    history against the reference's sequential build of the same app on the
    same mesh (golden), to 1e-4 relative (the rms sum's own order differs)."""
    g, _, out, rows = run_variant(tmp_path, "airfoil_ref_sp", "airfoil_120x60_random_app_sp")
    assert [int(r[0]) for r in rows] == list(g["iters"])
    rms = np.array([float(r[1]) for r in rows])
    assert np.all(np.abs(rms - g["rms"]) <= 1e-4 * g["rms"]), (rms, g["rms"])
