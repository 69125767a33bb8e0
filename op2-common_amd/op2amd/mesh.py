"""Synthetic airfoil-channel meshes (include/op2amd_mesh.h) as numpy arrays."""
import ctypes as C
import os

import numpy as np

from ._lib import lib

STRUCTURED, RANDOM, RCM, HILBERT = 0, 1, 2, 3
ORDERINGS = {"structured": STRUCTURED, "random": RANDOM, "rcm": RCM, "hilbert": HILBERT}


class op2amd_mesh(C.Structure):
    _fields_ = [("nnode", C.c_int), ("ncell", C.c_int), ("nedge", C.c_int),
                ("nbedge", C.c_int), ("x", C.POINTER(C.c_double)),
                ("cell", C.POINTER(C.c_int)), ("edge", C.POINTER(C.c_int)),
                ("ecell", C.POINTER(C.c_int)), ("bedge", C.POINTER(C.c_int)),
                ("becell", C.POINTER(C.c_int)), ("bound", C.POINTER(C.c_int))]


lib.op2amd_mesh_generate.restype = C.c_int
lib.op2amd_mesh_generate.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_int,
                                     C.POINTER(op2amd_mesh)]
lib.op2amd_mesh_free.argtypes = [C.POINTER(op2amd_mesh)]
lib.op2amd_mesh_write.restype = C.c_int
lib.op2amd_mesh_write.argtypes = [C.POINTER(op2amd_mesh), C.c_char_p]
lib.op2amd_mesh_read.restype = C.c_int
lib.op2amd_mesh_read.argtypes = [C.c_char_p, C.POINTER(op2amd_mesh)]
lib.op2amd_mesh_renumber_rcm.restype = C.c_int
lib.op2amd_mesh_renumber_rcm.argtypes = [C.POINTER(op2amd_mesh)] + [C.c_void_p] * 4
lib.op2amd_mesh_cell_bandwidth.argtypes = [C.POINTER(op2amd_mesh),
                                           C.POINTER(C.c_int),
                                           C.POINTER(C.c_double)]

_FIELDS = [("x", np.float64, 2, "nnode"), ("cell", np.int32, 4, "ncell"),
           ("edge", np.int32, 2, "nedge"), ("ecell", np.int32, 2, "nedge"),
           ("bedge", np.int32, 2, "nbedge"), ("becell", np.int32, 1, "nbedge"),
           ("bound", np.int32, 1, "nbedge")]


class Mesh:
    """Owns numpy copies of a mesh; `.c()` gives a C view for the library."""

    def __init__(self, nnode, ncell, nedge, nbedge, **arrays):
        self.nnode, self.ncell, self.nedge, self.nbedge = nnode, ncell, nedge, nbedge
        for name, dtype, dim, count in _FIELDS:
            a = np.ascontiguousarray(arrays[name], dtype=dtype)
            n = getattr(self, count)
            setattr(self, name, a.reshape((n, dim)) if dim > 1 else a.reshape(n))

    @classmethod
    def _from_c(cls, m):
        arrays = {}
        for name, dtype, dim, count in _FIELDS:
            n = getattr(m, count) * dim
            ptr = getattr(m, name)
            arrays[name] = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, dtype)
        mesh = cls(m.nnode, m.ncell, m.nedge, m.nbedge, **arrays)
        lib.op2amd_mesh_free(C.byref(m))
        return mesh

    def c(self):
        m = op2amd_mesh()
        m.nnode, m.ncell, m.nedge, m.nbedge = self.nnode, self.ncell, self.nedge, self.nbedge
        for name, dtype, _, _ in _FIELDS:
            ctype = C.c_double if dtype is np.float64 else C.c_int
            setattr(m, name, getattr(self, name).ctypes.data_as(C.POINTER(ctype)))
        self._keep = m
        return m

    def write(self, path):
        rc = lib.op2amd_mesh_write(C.byref(self.c()), os.fspath(path).encode())
        if rc:
            raise OSError(f"cannot write mesh to {path}")

    def bandwidth(self):
        mx, mean = C.c_int(), C.c_double()
        lib.op2amd_mesh_cell_bandwidth(C.byref(self.c()), C.byref(mx), C.byref(mean))
        return mx.value, mean.value

    def renumber_rcm(self):
        """RCM-renumber in place; returns new-of-old permutations
        (cells, edges, bedges, nodes)."""
        m = self.c()
        perms = [np.empty(n, np.int32) for n in (self.ncell, self.nedge, self.nbedge, self.nnode)]
        rc = lib.op2amd_mesh_renumber_rcm(C.byref(m), *[p.ctypes.data_as(C.c_void_p) for p in perms])
        if rc:
            raise RuntimeError("RCM renumbering failed")
        return perms


def generate(I, J, seed=0, ordering="random"):
    m = op2amd_mesh()
    rc = lib.op2amd_mesh_generate(I, J, seed, ORDERINGS[ordering], C.byref(m))
    if rc:
        raise ValueError(f"mesh generation failed ({rc}) for {I}x{J}")
    return Mesh._from_c(m)


def read(path):
    m = op2amd_mesh()
    rc = lib.op2amd_mesh_read(os.fspath(path).encode(), C.byref(m))
    if rc:
        raise OSError(f"cannot read mesh {path} ({rc})")
    return Mesh._from_c(m)
