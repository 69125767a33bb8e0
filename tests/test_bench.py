"""bench.py's contract pieces that run without a GPU: it refuses more than
one GPU with a message, and the roofline's traffic comes from the committed
counter summaries under profiles/ (the line the summaries end with)."""
import importlib.util
import subprocess
import sys

from conftest import REPO


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", REPO / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=str(REPO))
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_roofline_traffic_from_profiles():
    b = load_bench()
    hand = b.profiled_traffic("res_calc/owner/128", "1200x600/rcm")
    gen = b.profiled_traffic("res_calc/generated-owner/128", "1200x600/rcm")
    large = b.profiled_traffic("res_calc/owner/256", "4800x2400/rcm")
    # the committed summaries hold HBM bytes per launch near the algorithmic
    # 109.44 MB (720k cells) and 1751.0 MB (11.5M cells)
    assert hand and 1.0 <= hand / 109.44e6 <= 1.1
    assert gen and 1.0 <= gen / 109.44e6 <= 1.1
    assert large and 1.0 <= large / 1751.0e6 <= 1.1
    assert b.profiled_traffic("res_calc/owner/128", "no-such-mesh") is None


def test_parity_check_flags_a_wrong_iteration():
    """The bench's parity check covers every iteration: one rms off by 1e-9
    relative anywhere in the trajectory, or q off by 1e-9, fails it."""
    import numpy as np
    b = load_bench()
    orc = {"rms": np.linspace(1e-3, 1e-4, 520), "q": np.linspace(1.0, 2.0, 40).reshape(10, 4)}
    good = {"rms": orc["rms"] * (1 + 1e-14), "q": orc["q"] + 1e-15}
    late = {"rms": orc["rms"].copy(), "q": orc["q"]}
    late["rms"][517] *= 1 + 1e-9
    qbad = {"rms": orc["rms"], "q": orc["q"].copy()}
    qbad["q"][3, 2] += 2e-9 * 2.0
    nan = {"rms": orc["rms"].copy(), "q": orc["q"].copy()}
    nan["rms"][100] = np.nan
    nan["q"][0, 0] = np.nan
    res = b.parity({"good": good, "late": late, "qbad": qbad, "nan": nan}, orc)
    assert max(res["good"].values()) <= b.PARITY_TOL
    assert res["late"]["rms_rel_err_max"] > b.PARITY_TOL
    assert res["qbad"]["q_rel_err"] > b.PARITY_TOL and res["qbad"]["rms_rel_err_max"] == 0.0
    assert not res["nan"]["rms_rel_err_max"] <= b.PARITY_TOL
    assert not res["nan"]["q_rel_err"] <= b.PARITY_TOL
