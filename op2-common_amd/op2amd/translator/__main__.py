import argparse
import sys

from . import translate
from .op import OpError


def main():
    ap = argparse.ArgumentParser(prog="python -m op2amd.translator",
                                 description="Translate OP2 C++ applications for the MI355X backend")
    ap.add_argument("sources", nargs="+")
    ap.add_argument("-o", "--outdir", default=".")
    ap.add_argument("-t", "--target", default="hip")
    a = ap.parse_args()
    for src in a.sources:
        try:
            paths = translate(src, a.outdir, a.target)
        except (OpError, ValueError, SyntaxError) as e:
            print(f"{src}: {e}", file=sys.stderr)
            return 1
        print(" ".join(paths.values()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
