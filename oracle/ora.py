"""ctypes wrapper over oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check of the repository's entry module and bench.py's CPU-baseline leg
import this module, and only as the checker (or the reported CPU baseline). The product never
loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U64P = ctypes.POINTER(ctypes.c_uint64)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_fe_run.restype = _P
        L.ora_fe_run.argtypes = [_P, _I64, _I64, _I64, _P, _P, ctypes.c_float, ctypes.c_int]
        L.ora_fe_sizes.argtypes = [_P, _P, _P]
        L.ora_fe_times.argtypes = [_P, _P]
        L.ora_fe_mesh.argtypes = [_P, _P, _P, _P]
        L.ora_fe_rows.argtypes = [_P] * 7
        L.ora_fe_offsets.argtypes = [_P] * 5
        L.ora_fe_cases.argtypes = [_P, _P, _P]
        L.ora_fe_free.argtypes = [_P]
        L.ora_mc_run.restype = _P
        L.ora_mc_run.argtypes = [_P, _I64, _I64, _I64, _P, _P, ctypes.c_float]
        L.ora_mc_sizes.argtypes = [_P, _P, _P, _P]
        L.ora_mc_mesh.argtypes = [_P, _P, _P, _P]
        L.ora_mc_edge_ids.argtypes = [_P, _P]
        L.ora_mc_free.argtypes = [_P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class FEResult:
    """Everything the oracle's Flying Edges computes for one (volume, isovalue)."""

    def __init__(self, vol, iso, spacing=(1, 1, 1), origin=(0, 0, 0), nthreads=1, want_debug=True):
        L = lib()
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        nz, ny, nx = vol.shape
        sp = (ctypes.c_float * 3)(*spacing)
        org = (ctypes.c_float * 3)(*origin)
        h = L.ora_fe_run(vol.ctypes.data, nx, ny, nz, sp, org, ctypes.c_float(iso), int(nthreads))
        try:
            v, t = ctypes.c_uint64(), ctypes.c_uint64()
            L.ora_fe_sizes(h, ctypes.byref(v), ctypes.byref(t))
            self.nverts, self.ntris = v.value, t.value
            times = (ctypes.c_double * 6)()
            L.ora_fe_times(h, times)
            self.times = list(times)
            self.points = np.zeros((self.nverts, 3), np.float32)
            self.normals = np.zeros((self.nverts, 3), np.float32)
            self.triangles = np.zeros((self.ntris, 3), np.uint64)
            L.ora_fe_mesh(h, _ptr(self.points), _ptr(self.normals), _ptr(self.triangles))
            self.rows = None
            if want_debug and nx >= 2 and ny >= 2 and nz >= 2:
                g = ny * nz
                c = (ny - 1) * (nz - 1)
                xl, xr, xc, yc, zc = (np.zeros(g, np.uint64) for _ in range(5))
                tc = np.zeros(c, np.uint64)
                L.ora_fe_rows(h, _ptr(xl), _ptr(xr), _ptr(xc), _ptr(yc), _ptr(zc), _ptr(tc))
                xs, ys, zs = (np.zeros(g, np.uint64) for _ in range(3))
                to = np.zeros(c, np.uint64)
                L.ora_fe_offsets(h, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(to))
                cube = np.zeros((nx - 1) * (ny - 1) * (nz - 1), np.uint8)
                edge = np.zeros((nx - 1) * ny * nz, np.uint8)
                L.ora_fe_cases(h, _ptr(cube), _ptr(edge))
                self.rows = dict(xl=xl, xr=xr, xcount=xc, ycount=yc, zcount=zc, tcount=tc,
                                 xstart=xs, ystart=ys, zstart=zs, toffset=to,
                                 cube_cases=cube, edge_cases=edge)
        finally:
            L.ora_fe_free(h)


class MCResult:
    def __init__(self, vol, iso, spacing=(1, 1, 1), origin=(0, 0, 0)):
        L = lib()
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        nz, ny, nx = vol.shape
        sp = (ctypes.c_float * 3)(*spacing)
        org = (ctypes.c_float * 3)(*origin)
        h = L.ora_mc_run(vol.ctypes.data, nx, ny, nz, sp, org, ctypes.c_float(iso))
        try:
            v, t, s = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
            L.ora_mc_sizes(h, ctypes.byref(v), ctypes.byref(t), ctypes.byref(s))
            self.nverts, self.ntris, self.seconds = v.value, t.value, s.value
            self.points = np.zeros((self.nverts, 3), np.float32)
            self.normals = np.zeros((self.nverts, 3), np.float32)
            self.triangles = np.zeros((self.ntris, 3), np.uint64)
            L.ora_mc_mesh(h, _ptr(self.points), _ptr(self.normals), _ptr(self.triangles))
            self.edge_ids = np.zeros(self.nverts, np.uint64)
            L.ora_mc_edge_ids(h, _ptr(self.edge_ids))
        finally:
            L.ora_mc_free(h)

    def edge_order(self):
        """(points, normals, triangles) renumbered as util::duplicateRemover renumbers them
        (marchingCubes/util/DuplicateRemover.h:34-70) when the duplicate-free OpenMP driver runs
        one section (marchingCubes/openmpDupFree/main.cpp:198-331): points sorted by global edge
        id, which is unique per point here, and each triangle's ids sent through the old -> new
        map; triangles keep their cell order."""
        order = np.argsort(self.edge_ids, kind="stable")
        old_to_new = np.empty(self.nverts, np.uint64)
        old_to_new[order] = np.arange(self.nverts, dtype=np.uint64)
        tris = old_to_new[self.triangles.astype(np.int64)] if self.ntris else self.triangles.copy()
        return self.points[order], self.normals[order], tris


def ref_binary(name: str) -> str | None:
    """Path of a reference binary compiled into oracle/_ref, or None when absent."""
    p = os.path.join(REF_DIR, name)
    return p if os.access(p, os.X_OK) else None
