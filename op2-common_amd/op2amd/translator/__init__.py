"""Source-to-source translator for OP2 C++ applications targeting this
backend (the role of translator-v2's `hip` target,
translator-v2/op2-translator/target.py:50-55, cpp/schemes.py:121-167).

    python -m op2amd.translator app.cpp -o outdir

writes, for each input `<stem>.cpp`:
  <stem>_op.cpp      the application with op_par_loop calls rewritten to stubs
  <stem>_op2gen.h    the stubs' prototypes
  <stem>_kernels.hip the stubs, their HIP kernels and the constants' hook
Build: hipcc --offload-arch=gfx950 -I include <stem>_op.cpp <stem>_kernels.hip
       -L op2-common_amd/lib -lop2hip
"""
import os

from . import hip, op, source  # noqa: F401

TARGETS = {"hip": hip}


def translate(path, outdir, target="hip"):
    """Translate one application file; returns the written paths."""
    if target not in TARGETS:
        raise ValueError(f"unknown target {target!r} (have {sorted(TARGETS)})")
    prog = op.parse_program(path)
    kern, hdr = TARGETS[target].generate(prog)
    os.makedirs(outdir, exist_ok=True)
    header = f"{prog.stem}_op2gen.h"
    paths = {"app": os.path.join(outdir, f"{prog.stem}_op.cpp"),
             "header": os.path.join(outdir, header),
             "kernels": os.path.join(outdir, f"{prog.stem}_kernels.hip")}
    with open(paths["app"], "w") as f:
        f.write(op.translate_program(prog, header))
    with open(paths["header"], "w") as f:
        f.write(hdr)
    with open(paths["kernels"], "w") as f:
        f.write(kern)
    return paths
