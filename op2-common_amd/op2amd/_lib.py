"""Loads libop2hip.so built in-tree by op2-common_amd/Makefile."""
import ctypes
import os
import pathlib

PKG_ROOT = pathlib.Path(__file__).resolve().parent.parent
REPO_ROOT = PKG_ROOT.parent
LIB_PATH = PKG_ROOT / "lib" / "libop2hip.so"


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C op2-common_amd` "
            "(or __graft_entry__.build()); this backend has no CPU fallback")
    return ctypes.CDLL(os.fspath(LIB_PATH), mode=ctypes.RTLD_GLOBAL)


lib = _load()
