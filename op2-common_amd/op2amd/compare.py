"""Comparison of OP2 result dumps, and permutation files for renumbered runs.

Dumps are the files op_print_dat_to_binfile and op_print_dat_to_txtfile write
(op2/src/core/op_lib_core.cpp:1283-1307: int32 count, int32 dim, then the
elements' raw bytes; :1309-1369: a "count dim" line, then the values).  The
tolerance modes follow the reference's airfoil result checkers:
apps/c/airfoil/compare_results/comparebin.cpp (binary, |a - b| > 1e-14 flags a
value) and compare.cpp (text, |a - b| > 1e-6); a mismatch is reported as
"line n+2 element d+1" of element n, component d, as they report it, and (as
there) a NaN on either side is not flagged.  `exact` compares bit patterns
instead, which also flags NaNs and the sign of zero.

A permutation file holds new_of_old of one set (the new number of every old
element) as text: the count on the first line, then one integer per line.
Comparing a dump of a renumbered run against one of the original numbering
takes such a file: element i of the first file is checked against element
perm[i] of the second.

Command line:
    python -m op2amd.compare FILE1 FILE2 [--text] [--exact] [--tol T] [--perm P]
prints the reference checkers' messages and exits 0 when the files agree,
1 when they differ, 2 on unreadable or mismatched files."""
import argparse
import sys

import numpy as np

BIN_TOL = 1e-14  # comparebin.cpp:24
TXT_TOL = 1e-6   # compare.cpp:24


class DumpError(Exception):
    pass


def read_bin(path, dtype=np.float64):
    """(count, dim, values[count, dim]) of a binary dump."""
    raw = open(path, "rb").read()
    if len(raw) < 8:
        raise DumpError(f"{path}: shorter than its header")
    count, dim = (int(v) for v in np.frombuffer(raw[:8], np.int32))
    itemsize = np.dtype(dtype).itemsize
    if count < 0 or dim < 1 or len(raw) - 8 != count * dim * itemsize:
        raise DumpError(f"{path}: {len(raw) - 8} data bytes for {count} x {dim} "
                        f"values of {itemsize} bytes")
    return count, dim, np.frombuffer(raw[8:], dtype).reshape(count, dim).copy()


def read_txt(path):
    """(count, dim, values[count, dim]) of a text dump."""
    with open(path) as f:
        head = f.readline().split()
        if len(head) != 2:
            raise DumpError(f"{path}: header must be 'count dim'")
        count, dim = int(head[0]), int(head[1])
        vals = np.array(f.read().split(), dtype=np.float64)
    if count < 0 or dim < 1 or vals.size != count * dim:
        raise DumpError(f"{path}: {vals.size} values for {count} x {dim}")
    return count, dim, vals.reshape(count, dim)


def write_bin(path, values):
    """A binary dump of a [count, dim] array in op_print_dat_to_binfile's form."""
    v = np.ascontiguousarray(values)
    v = v.reshape(v.shape[0], -1)
    with open(path, "wb") as f:
        f.write(np.array([v.shape[0], v.shape[1]], np.int32).tobytes())
        f.write(v.tobytes())


def write_perm(path, new_of_old):
    p = np.asarray(new_of_old, dtype=np.int64).ravel()
    with open(path, "w") as f:
        f.write(f"{p.size}\n")
        f.write("".join(f"{int(v)}\n" for v in p))


def read_perm(path):
    with open(path) as f:
        toks = f.read().split()
    if not toks:
        raise DumpError(f"{path}: empty permutation file")
    n = int(toks[0])
    p = np.array(toks[1:], dtype=np.int64)
    if p.size != n or not np.array_equal(np.sort(p), np.arange(n)):
        raise DumpError(f"{path}: not a permutation of {n} elements")
    return p


def compare_arrays(a, b, tol=BIN_TOL, exact=False, perm=None):
    """Mismatching (element, component) pairs of a[i] against b[perm[i]] (or
    b[i]), as a list of (n, d), in ascending order."""
    a = np.asarray(a)
    b = np.asarray(b)
    if perm is not None:
        perm = np.asarray(perm)
        if perm.size != b.shape[0]:
            raise DumpError(f"permutation of {perm.size} elements for {b.shape[0]} rows")
        b = b[perm]
    if a.shape != b.shape:
        raise DumpError(f"shapes differ: {a.shape} and {b.shape}")
    if exact:
        bad = a.view(np.uint8).reshape(a.shape + (-1,)) != b.view(np.uint8).reshape(b.shape + (-1,))
        bad = bad.any(axis=-1)
    else:
        with np.errstate(invalid="ignore"):
            bad = np.abs(a.astype(np.float64) - b.astype(np.float64)) > tol
    return [(int(n), int(d)) for n, d in np.argwhere(bad)]


def compare_files(f1, f2, text=False, exact=False, tol=None, perm=None, out=print):
    """The checkers' report for two dumps; returns True when they agree."""
    read = read_txt if text else read_bin
    c1, d1, v1 = read(f1)
    c2, d2, v2 = read(f2)
    out(f"File 1 lines {c1}, File2 lines {c2}")
    if c1 != c2 or d1 != d2:
        raise DumpError("File mismatch: number of lines or element dimensions not matching")
    bad = compare_arrays(v1, v2, tol=(TXT_TOL if text else BIN_TOL) if tol is None else tol,
                         exact=exact, perm=read_perm(perm) if isinstance(perm, str) else perm)
    for n, d in bad:
        out(f"File mismatch: at line {n + 2} element {d + 1}")
    out("Files Identical" if not bad else "Files Differ")
    return not bad


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m op2amd.compare",
                                 description="compare two OP2 dat dumps")
    ap.add_argument("file1")
    ap.add_argument("file2")
    ap.add_argument("--text", action="store_true", help="text dumps (default: binary)")
    ap.add_argument("--exact", action="store_true", help="bit-for-bit comparison")
    ap.add_argument("--tol", type=float, default=None, help="absolute tolerance")
    ap.add_argument("--perm", default=None,
                    help="permutation file: element i of FILE1 is element perm[i] of FILE2")
    a = ap.parse_args(argv)
    try:
        ok = compare_files(a.file1, a.file2, a.text, a.exact, a.tol, a.perm)
    except (DumpError, OSError, ValueError) as e:
        print(e)
        return 2
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
