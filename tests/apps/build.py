"""Translate the OP2 programs under tests/apps with op2amd.translator and
compile them for gfx950 against the in-tree libop2hip.so; where the
reference tree is present, also its unmodified airfoil application.  Outputs
go to tests/apps/_build/<program>/ (the executable has the program's name).

    python tests/apps/build.py [program ...]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "op2-common_amd"))

PROGRAMS = {
    "gbl": "functional/gbl.cpp",
    "reductions": "functional/reductions.cpp",
    "idx": "functional/idx.cpp",
    "strides": "functional/strides.cpp",
    "consts": "functional/consts.cpp",
    "helpers": "functional/helpers.cpp",
    "tempdats": "functional/tempdats.cpp",
    "airfoil_tr": "airfoil/airfoil_tr.cpp",
}

# The reference's own airfoil application, translated unchanged, where the
# reference tree is present (this container).  Only its executable is kept:
# the generated sources and objects, which carry the application's text, are
# removed after the link.
REFERENCE = "/root/reference"
REF_PROGRAMS = {
    "airfoil_ref": "apps/c/airfoil/airfoil_plain/dp/airfoil.cpp",
    # single precision throughout, rms-only update
    "airfoil_ref_sp": "apps/c/airfoil/airfoil_plain/sp/airfoil.cpp",
    # res, adt and qold as temporary dats declared and freed every iteration
    "airfoil_ref_tempdats": "apps/c/airfoil/airfoil_tempdats/dp/airfoil.cpp",
    # the reference's own functional tests (self-checking; exit 1 on a failure)
    "ref_gbl": "tests/functional/gbl/gbl_tests.cpp",
    "ref_reduc": "tests/functional/dat_reductions/reduc_tests.cpp",
    "ref_idx": "tests/functional/idx/idx_tests.cpp",
    "ref_strides": "tests/functional/strides/stride_tests.cpp",
    "ref_const": "tests/functional/const/const_tests.cpp",
}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(name):
    from op2amd.translator import translate
    if name in REF_PROGRAMS:
        src = os.path.join(REFERENCE, REF_PROGRAMS[name])
    else:
        src = os.path.join(HERE, PROGRAMS[name])
    out = os.path.join(HERE, "_build", name)
    paths = translate(src, out)
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + os.path.dirname(src), "-I" + out]
    lib = os.path.join(REPO, "op2-common_amd", "lib")
    objs = []
    for key, lang in (("kernels", "hip"), ("app", "c++")):
        obj = paths[key] + ".o"
        cmd = [HIPCC, *FLAGS, *inc, "-x", lang, "-c", paths[key], "-o", obj, "-w"]
        subprocess.run(cmd, check=True)
        objs.append(obj)
    exe = os.path.join(out, name)
    subprocess.run([HIPCC, "--offload-arch=gfx950", *objs, "-L" + lib, "-lop2hip",
                    "-Wl,-rpath," + lib, "-Wl,-rpath,$ORIGIN/../../../../op2-common_amd/lib",
                    "-o", exe], check=True)
    if name in REF_PROGRAMS:
        for p in [*paths.values(), *objs]:
            os.remove(p)
    return exe


if __name__ == "__main__":
    names = sys.argv[1:] or [*PROGRAMS, *(REF_PROGRAMS if os.path.isdir(REFERENCE) else [])]
    for n in names:
        print(build(n), flush=True)
