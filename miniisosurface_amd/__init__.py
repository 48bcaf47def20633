"""MI355X-native isosurface extraction (Flying Edges and Marching Cubes) behind the miniapp's API.

Host-side mirror of the reference interfaces, over the C ABI of include/miniiso.h:

    FlyingEdgesAlgorithm(image, isoval)  .pass1() .pass2() .pass3() .pass4() .moveOutput()
        -- flyingEdges/serial/FlyingEdgesAlgorithm.h:27-49 (OpenMP variant identical)
    MarchingCubes(image, isoval[, nSectionsX, nSectionsY, nSectionsZ])
        -- marchingCubes/serial/main.cpp:42-44, marchingCubes/openmp/main.cpp:159-162

`image` is a vtkio.Image3D (host numpy volume) or a DeviceVolume (a device pointer, e.g. from a
torch tensor on the GPU). Outputs are vtkio.TriangleMesh with float32 points/normals and uint64
triangle indices, exactly the reference's types.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import MisoError, Volume, check, lib
from .vtkio import Image3D, TriangleMesh, load_dat_image, load_image, load_mesh, save_image, save_mesh

__all__ = ["FlyingEdgesAlgorithm", "MarchingCubes", "DeviceVolume", "Image3D", "TriangleMesh",
           "MisoError", "load_image", "load_dat_image", "save_image", "load_mesh", "save_mesh", "flying_edges",
           "FlyingEdgesSlabs", "flying_edges_slabs"]


@dataclass
class DeviceVolume:
    """A float32 volume already resident in device memory (x fastest)."""
    ptr: int
    dims: tuple                 # (nx, ny, nz)
    spacing: tuple = (1.0, 1.0, 1.0)
    origin: tuple = (0.0, 0.0, 0.0)
    owner: object = None        # keeps e.g. the torch tensor alive

    @classmethod
    def from_tensor(cls, t, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
        if t.dtype.__str__() != "torch.float32" or t.dim() != 3 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("expected a contiguous float32 (nz, ny, nx) tensor on the GPU")
        nz, ny, nx = t.shape
        return cls(t.data_ptr(), (nx, ny, nz), tuple(spacing), tuple(origin), t)


def _volume_struct(image):
    v = Volume()
    if isinstance(image, DeviceVolume):
        v.data = image.ptr
        v.nx, v.ny, v.nz = (int(d) for d in image.dims)
        v.on_device = 1
        keep = image
        sp, org = image.spacing, image.origin
    else:
        if not isinstance(image, Image3D):
            image = Image3D(np.asarray(image, dtype=np.float32))
        v.data = image.data.ctypes.data
        v.nx, v.ny, v.nz = image.dims
        v.on_device = 0
        keep = image
        sp, org = image.spacing, image.origin
    for a in range(3):
        v.spacing[a] = float(sp[a])
        v.origin[a] = float(org[a])
    return v, keep


def _stream_ptr(stream, device_input=False):
    """The HIP stream the library runs on. With none given, the library's own stream, which is
    ordered after work on the null stream (torch's default stream); for device-resident input
    while torch's current stream is another one, that stream, so that the kernels are ordered
    after the torch work that produced the volume."""
    if stream is None and device_input:
        import torch
        cur = torch.cuda.current_stream()
        if cur.cuda_stream != 0:
            stream = cur
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


class FlyingEdgesAlgorithm:
    """FlyingEdgesAlgorithm (flyingEdges/serial/FlyingEdgesAlgorithm.h:27-49) on the GPU.

    Passes must run in order; each waits for its kernels. moveOutput() returns the mesh and, as in
    the reference, the algorithm no longer holds it afterwards."""

    def __init__(self, image, isoval, stream=None):
        self._vol, self._keep = _volume_struct(image)
        self.isoval = float(np.float32(isoval))
        h = ctypes.c_void_p()
        check(lib().miso_fe_create(ctypes.byref(self._vol), ctypes.c_float(self.isoval),
                                   _stream_ptr(stream, bool(self._vol.on_device)), ctypes.byref(h)),
              "FlyingEdgesAlgorithm")
        self._h = h
        self._destroy = lib().miso_fe_destroy      # kept: module globals may be gone at exit
        self.dims = (self._vol.nx, self._vol.ny, self._vol.nz)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def _pass(self, n):
        check(lib().miso_fe_pass(self._h, n), f"pass{n}")

    def pass1(self):
        self._pass(1)

    def pass2(self):
        self._pass(2)

    def pass3(self):
        self._pass(3)

    def pass4(self):
        self._pass(4)

    def run(self):
        """All four passes with one host synchronisation."""
        check(lib().miso_fe_run(self._h), "run")

    def set_isoval(self, isoval):
        self.isoval = float(np.float32(isoval))
        check(lib().miso_fe_set_isoval(self._h, ctypes.c_float(self.isoval)), "set_isoval")

    def set_index_bits(self, bits):
        check(lib().miso_fe_set_index_bits(self._h, int(bits)), "set_index_bits")

    def counts(self):
        v, t = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().miso_fe_counts(self._h, ctypes.byref(v), ctypes.byref(t)), "counts")
        return v.value, t.value

    def set_timing(self, passes=(1, 2, 3, 4)):
        """Time only these passes with HIP events in later runs (the others report 0); each
        timed pass adds an event pair between the kernels."""
        mask = 0
        for p in passes:
            if p not in (1, 2, 3, 4):
                raise ValueError("passes are 1..4")
            mask |= 1 << (p - 1)
        check(lib().miso_fe_set_timing(self._h, ctypes.c_uint(mask)), "set_timing")

    def pass_times_ms(self):
        ms = (ctypes.c_float * 4)()
        check(lib().miso_fe_pass_times(self._h, ms), "pass_times")
        return list(ms)

    def moveOutput(self) -> TriangleMesh:
        v, t = self.counts()
        pts = np.empty((v, 3), np.float32)
        nrm = np.empty((v, 3), np.float32)
        tris = np.empty((t, 3), np.uint64)
        check(lib().miso_fe_copy_out(self._h, pts.ctypes.data, nrm.ctypes.data, tris.ctypes.data), "moveOutput")
        return TriangleMesh(pts, nrm, tris)

    def device_output(self):
        """(points_ptr, normals_ptr, triangles_ptr, index_bits, nverts, ntris) on the device."""
        p, n, t = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        bits = ctypes.c_int()
        check(lib().miso_fe_device_output(self._h, ctypes.byref(p), ctypes.byref(n), ctypes.byref(t),
                                          ctypes.byref(bits)), "device_output")
        v, tt = self.counts()
        return p.value, n.value, t.value, bits.value, v, tt

    def rows(self):
        """Per-row state after pass 3 (see miso_fe_rows)."""
        nx, ny, nz = self.dims
        R = ny * nz
        out = {k: np.zeros(R, np.uint32) for k in ("xl", "xr", "xcount", "ycount", "zcount", "tcount")}
        out["voffset"] = np.zeros(R, np.uint64)
        out["toffset"] = np.zeros(R, np.uint64)
        check(lib().miso_fe_rows(self._h, *(out[k].ctypes.data for k in
                                            ("xl", "xr", "xcount", "ycount", "zcount", "tcount", "voffset",
                                             "toffset"))), "rows")
        return out

    def cube_cases(self, all_cells=False):
        nx, ny, nz = self.dims
        n = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
        out = np.zeros(n, np.uint8)
        if n:
            check(lib().miso_fe_cube_cases(self._h, out.ctypes.data, int(bool(all_cells))), "cube_cases")
        return out


def flying_edges(image, isoval, stream=None) -> TriangleMesh:
    algo = FlyingEdgesAlgorithm(image, isoval, stream)
    algo.set_timing(())   # no pass times are returned: no event pairs between the passes
    algo.run()
    return algo.moveOutput()


class FlyingEdgesSlabs:
    """Flying Edges over z-slabs of at most `planes` grid planes (miso_fe_slabs_*): the section
    decomposition with ghost planes of marchingCubes/mpi/mpiutil.h:87-221 along z, one slab after
    another on one GPU. Iterating yields, per slab, (part, points, normals, triangles) with part =
    {k0, k1, vbase, nverts, tbase, ntris}; the triangles index the whole mesh's vertices, and the
    parts in order concatenate to flying_edges()'s mesh. A host volume is read slab by slab (it
    stays borrowed until the iteration ends); a DeviceVolume is read in place."""

    def __init__(self, image, isoval, planes, stream=None):
        self._h = None
        self._vol, self._keep = _volume_struct(image)
        h = ctypes.c_void_p()
        check(lib().miso_fe_slabs_create(ctypes.byref(self._vol), ctypes.c_float(float(np.float32(isoval))),
                                         int(planes), _stream_ptr(stream, bool(self._vol.on_device)),
                                         ctypes.byref(h)), "FlyingEdgesSlabs")
        self._h = h
        self._destroy = lib().miso_fe_slabs_destroy

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def next(self) -> bool:
        """Extract the next slab; False once every plane is done."""
        done = ctypes.c_int()
        check(lib().miso_fe_slabs_next(self._h, ctypes.byref(done)), "FlyingEdgesSlabs")
        return not done.value

    def restart(self, isoval):
        """Start over from the first slab at a new isovalue (the device buffers are kept)."""
        check(lib().miso_fe_slabs_restart(self._h, ctypes.c_float(float(np.float32(isoval)))), "restart")

    def part(self) -> dict:
        k0, k1 = ctypes.c_int64(), ctypes.c_int64()
        vb, nv, tb, nt = (ctypes.c_uint64() for _ in range(4))
        check(lib().miso_fe_slabs_part(self._h, ctypes.byref(k0), ctypes.byref(k1), ctypes.byref(vb), ctypes.byref(nv),
                                       ctypes.byref(tb), ctypes.byref(nt)), "part")
        return dict(k0=k0.value, k1=k1.value, vbase=vb.value, nverts=nv.value, tbase=tb.value, ntris=nt.value)

    def copy_part(self):
        p = self.part()
        pts = np.empty((p["nverts"], 3), np.float32)
        nrm = np.empty((p["nverts"], 3), np.float32)
        tris = np.empty((p["ntris"], 3), np.uint64)
        check(lib().miso_fe_slabs_copy_part(self._h, pts.ctypes.data, nrm.ctypes.data, tris.ctypes.data), "copy_part")
        return pts, nrm, tris

    def device_part(self):
        """Device pointers (points, normals, 64-bit triangles) of the current part."""
        p, n, t = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().miso_fe_slabs_device_part(self._h, ctypes.byref(p), ctypes.byref(n), ctypes.byref(t)),
              "device_part")
        return p.value, n.value, t.value

    def __iter__(self):
        while self.next():
            yield (self.part(), *self.copy_part())


def flying_edges_slabs(image, isoval, planes, stream=None) -> TriangleMesh:
    """flying_edges() computed slab by slab (FlyingEdgesSlabs); the same mesh."""
    parts = list(FlyingEdgesSlabs(image, isoval, planes, stream))
    if not parts:
        return TriangleMesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint64))
    return TriangleMesh(np.concatenate([q[1] for q in parts]), np.concatenate([q[2] for q in parts]),
                        np.concatenate([q[3] for q in parts]))


MC_ORDERS = {"first_visit": 0, "edge": 1}


class _MC:
    """order "first_visit": the serial reference's vertex numbering; "edge": vertices sorted by
    global edge id, as util::duplicateRemover leaves them (marchingCubes/util/DuplicateRemover.h:34-70)."""

    def __init__(self, image, isoval, stream=None, order="first_visit"):
        if order not in MC_ORDERS:
            raise ValueError(f"order must be one of {sorted(MC_ORDERS)}")
        self._h = None
        self._vol, self._keep = _volume_struct(image)
        h = ctypes.c_void_p()
        check(lib().miso_mc_create(ctypes.byref(self._vol), ctypes.c_float(float(np.float32(isoval))),
                                   _stream_ptr(stream, bool(self._vol.on_device)), ctypes.byref(h)),
              "MarchingCubes")
        self._h = h
        self._destroy = lib().miso_mc_destroy
        if MC_ORDERS[order]:
            self.set_order(order)

    def set_order(self, order):
        """Vertex order of the next run()s (miso_mc_set_order)."""
        if order not in MC_ORDERS:
            raise ValueError(f"order must be one of {sorted(MC_ORDERS)}")
        check(lib().miso_mc_set_order(self._h, MC_ORDERS[order]), "MarchingCubes")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def run(self):
        check(lib().miso_mc_run(self._h), "MarchingCubes")

    def counts(self):
        v, t = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().miso_mc_counts(self._h, ctypes.byref(v), ctypes.byref(t)), "counts")
        return v.value, t.value

    def time_ms(self):
        ms = ctypes.c_float()
        check(lib().miso_mc_time_ms(self._h, ctypes.byref(ms)), "time")
        return ms.value

    def mesh(self) -> TriangleMesh:
        v, t = self.counts()
        pts = np.empty((v, 3), np.float32)
        nrm = np.empty((v, 3), np.float32)
        tris = np.empty((t, 3), np.uint64)
        check(lib().miso_mc_copy_out(self._h, pts.ctypes.data, nrm.ctypes.data, tris.ctypes.data), "copy_out")
        return TriangleMesh(pts, nrm, tris)


def MarchingCubes(image, isoval, nSectionsX=None, nSectionsY=None, nSectionsZ=None, stream=None,
                  order="first_visit") -> TriangleMesh:
    """MarchingCubes(image, isoval) (marchingCubes/serial/main.cpp:42-192). The OpenMP overload's
    section counts (marchingCubes/openmp/main.cpp:159-162) are accepted and ignored: the result is
    the serial numbering, which the OpenMP variant's output equals up to vertex order and seam
    duplicates. order="edge" gives the duplicate-free OpenMP variant's vertex order instead
    (marchingCubes/openmpDupFree/main.cpp:324-331)."""
    mc = _MC(image, isoval, stream, order)
    mc.run()
    return mc.mesh()
