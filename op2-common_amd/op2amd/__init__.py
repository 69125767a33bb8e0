"""Python bindings of the MI355X-native OP2 backend (libop2hip.so).

The library itself is C++/HIP; these ctypes bindings expose its C ABI
(include/op2_hip.h, include/op2_hip_rt.h, include/op2amd_mesh.h,
include/op2amd_airfoil.h) to tests and benchmarks.  There is no CPU fallback:
importing the bindings without the built library raises.
"""
from ._lib import lib, LIB_PATH, REPO_ROOT  # noqa: F401
from . import api, mesh, airfoil, ensemble, costmodel  # noqa: F401
