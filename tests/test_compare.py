"""op2amd.compare: the reports of the reference's airfoil result checkers
(apps/c/airfoil/compare_results/{compare,comparebin}.cpp, built unmodified
into oracle/_ref) line for line, plus exact mode and permutation files."""
import pathlib
import subprocess
import sys

import numpy as np
import pytest

from op2amd import compare as cmp

REPO = pathlib.Path(__file__).resolve().parent.parent
REF = REPO / "oracle" / "_ref"


def write_txt(path, v):
    """A text dump as op_print_dat_to_txtfile writes doubles
    (op_lib_core.cpp:1316-1336)."""
    with open(path, "w") as f:
        f.write(f"{v.shape[0]} {v.shape[1]}\n")
        for row in v:
            f.write("".join(" %+2.15E" % (0.0 if x == 0 else x) for x in row) + "\n")


def ours(f1, f2, text):
    lines = []
    try:
        cmp.compare_files(str(f1), str(f2), text=text, out=lines.append)
    except cmp.DumpError as e:
        lines.append(str(e))
    return lines


def reference(f1, f2, text):
    exe = REF / ("compare" if text else "comparebin")
    if not exe.exists():
        pytest.skip("reference checkers not built (make -C oracle ref)")
    out = subprocess.run([str(exe), str(f1), str(f2)], capture_output=True, text=True,
                         timeout=60).stdout
    return out.splitlines()


def base(seed=0, n=50, d=4):
    return np.random.default_rng(seed).normal(size=(n, d))


def perturbed(text):
    v = base()
    w = v.copy()
    if text:
        w[7, 3] += 2e-6    # flagged
        w[12, 0] += 5e-7   # within 1e-6
        w[40, 1] -= 3e-6   # flagged
    else:
        w[3, 1] += 1e-13   # flagged
        w[10, 0] -= 5e-14  # flagged
        w[20, 3] += 1e-15  # within 1e-14 (may round to no change)
    return v, w


CASES = {
    "identical": lambda text: (base(), base()),
    "perturbed": perturbed,
    "nan": lambda text: (base(), np.where(np.arange(200).reshape(50, 4) == 9, np.nan, base())),
    "shape": lambda text: (base(), base(n=49)),
}


@pytest.mark.parametrize("text", [False, True], ids=["bin", "txt"])
@pytest.mark.parametrize("case", list(CASES))
def test_report_matches_reference_checker(tmp_path, case, text):
    a, b = CASES[case](text)
    f1, f2 = tmp_path / "a", tmp_path / "b"
    for f, v in ((f1, a), (f2, b)):
        write_txt(f, v) if text else cmp.write_bin(f, v)
    got = ours(f1, f2, text)
    assert got == reference(f1, f2, text)
    if case == "perturbed":
        assert got[-1] == "Files Differ" and len(got) == 4


def test_exact_mode_and_nan(tmp_path):
    v = base()
    w = v.copy()
    w[9 // 4, 9 % 4] = np.nan
    assert cmp.compare_arrays(v, w) == []              # as the checkers: NaN passes
    assert cmp.compare_arrays(v, w, exact=True) == [(2, 1)]
    z = np.zeros((2, 1))
    assert cmp.compare_arrays(z, -z, exact=True) == [(0, 0), (1, 0)]
    assert cmp.compare_arrays(z, -z) == []


def test_permutation_files(tmp_path):
    v = base(n=30)
    p = np.random.default_rng(1).permutation(30)  # new_of_old
    w = np.empty_like(v)
    w[p] = v                                       # the renumbered run's rows
    cmp.write_bin(tmp_path / "a.bin", v)
    cmp.write_bin(tmp_path / "b.bin", w)
    cmp.write_perm(tmp_path / "b.perm", p)
    assert np.array_equal(cmp.read_perm(tmp_path / "b.perm"), p)
    lines = []
    assert cmp.compare_files(str(tmp_path / "a.bin"), str(tmp_path / "b.bin"),
                             exact=True, perm=str(tmp_path / "b.perm"), out=lines.append)
    assert not cmp.compare_files(str(tmp_path / "a.bin"), str(tmp_path / "b.bin"),
                                 out=lambda s: None)
    # a file that is not a permutation is refused
    bad = p.copy()
    bad[0] = bad[1]
    cmp.write_perm(tmp_path / "bad.perm", bad)
    with pytest.raises(cmp.DumpError):
        cmp.read_perm(tmp_path / "bad.perm")


def test_command_line_exit_codes(tmp_path):
    v = base()
    cmp.write_bin(tmp_path / "a.bin", v)
    cmp.write_bin(tmp_path / "b.bin", v + 1e-12)
    cmp.write_bin(tmp_path / "c.bin", v[:10])

    def run(*args):
        return subprocess.run([sys.executable, "-m", "op2amd.compare", *map(str, args)],
                              capture_output=True, text=True, timeout=60,
                              cwd=REPO / "op2-common_amd").returncode
    assert run(tmp_path / "a.bin", tmp_path / "a.bin") == 0
    assert run(tmp_path / "a.bin", tmp_path / "b.bin") == 1
    assert run(tmp_path / "a.bin", tmp_path / "b.bin", "--tol", "1e-11") == 0
    assert run(tmp_path / "a.bin", tmp_path / "c.bin") == 2
