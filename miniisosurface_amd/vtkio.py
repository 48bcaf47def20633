"""VTK legacy file I/O with the reference miniapp's exact on-disk formats.

Input volumes (reader: flyingEdges/util/LoadImage.h:27-198, ConvertBuffer.h:24-72,
TypeInfo.h:23-116):
    # vtk DataFile Version ...      (line ignored)
    <title>                         (line ignored)
    BINARY
    DATASET STRUCTURED_POINTS
    DIMENSIONS nx ny nz  | SPACING sx sy sz | ORIGIN ox oy oz   (three lines, any order)
    POINT_DATA n
    SCALARS <name> <type>           char | short | unsigned_short | int | float | double
    LOOKUP_TABLE <name>
    <n big-endian values, converted to float32>

Output meshes (writer: flyingEdges/util/SaveTriangleMesh.h:33-118):
    POINTS V float, V*3 big-endian float32, newline,
    POLYGONS T 4T, T*(3, a, b, c) as big-endian 8-byte unsigned integers, newline,
    POINT_DATA V / NORMALS Normals float, V*3 big-endian float32, newline.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

# TypeInfo names understood by the reference reader (flyingEdges/util/TypeInfo.h:23-116)
_VTK_TYPES = {
    "char": ">i1",
    "short": ">i2",
    "unsigned_short": ">u2",
    "int": ">i4",
    "float": ">f4",
    "double": ">f8",
}


class BadFormat(RuntimeError):
    """Mirrors util::bad_format (flyingEdges/util/Errors.h:38-48)."""


class FileNotFound(RuntimeError):
    """Mirrors util::file_not_found (flyingEdges/util/Errors.h:26-36)."""


@dataclass
class Image3D:
    """Host-side volume: data[k][j][i] (x fastest), spacing and origin per axis.

    Mirrors util::Image3D (flyingEdges/util/Image3D.h:27-86)."""
    data: np.ndarray                      # float32, shape (nz, ny, nx), C-contiguous
    spacing: tuple = (1.0, 1.0, 1.0)
    origin: tuple = (0.0, 0.0, 0.0)

    def __post_init__(self):
        self.data = np.ascontiguousarray(self.data, dtype=np.float32)
        if self.data.ndim != 3:
            raise ValueError("volume must be 3-D (nz, ny, nx)")
        self.spacing = tuple(float(np.float32(s)) for s in self.spacing)
        self.origin = tuple(float(np.float32(o)) for o in self.origin)

    @property
    def dims(self):
        nz, ny, nx = self.data.shape
        return nx, ny, nz

    def xdimension(self):
        return self.dims[0]

    def ydimension(self):
        return self.dims[1]

    def zdimension(self):
        return self.dims[2]


@dataclass
class TriangleMesh:
    """Output mesh: float32 points/normals (V,3), triangles (T,3) of unsigned indices.

    Mirrors util::TriangleMesh (flyingEdges/util/TriangleMesh.h:24-90)."""
    points: np.ndarray
    normals: np.ndarray
    triangles: np.ndarray
    extra: dict = field(default_factory=dict)

    def numberOfVertices(self):
        return int(self.points.shape[0])

    def numberOfTriangles(self):
        return int(self.triangles.shape[0])


def _fmt(x: float) -> str:
    # 9 significant digits round-trip float32 exactly through the reference's `>>` parse
    return "%.9g" % np.float32(x)


def save_image(path: str, img: Image3D, vtk_type: str = "float") -> None:
    """Writes a volume the reference loader accepts."""
    nx, ny, nz = img.dims
    if vtk_type not in _VTK_TYPES:
        raise BadFormat("Unsupported data type")
    header = (
        "# vtk DataFile Version 3.0\n"
        "volume\n"
        "BINARY\n"
        "DATASET STRUCTURED_POINTS\n"
        f"DIMENSIONS {nx} {ny} {nz}\n"
        f"SPACING {_fmt(img.spacing[0])} {_fmt(img.spacing[1])} {_fmt(img.spacing[2])}\n"
        f"ORIGIN {_fmt(img.origin[0])} {_fmt(img.origin[1])} {_fmt(img.origin[2])}\n"
        f"POINT_DATA {nx * ny * nz}\n"
        f"SCALARS image_data {vtk_type}\n"
        "LOOKUP_TABLE default\n"
    )
    with open(path, "wb") as f:
        f.write(header.encode())
        f.write(img.data.astype(_VTK_TYPES[vtk_type]).tobytes())


def load_dat_image(path: str) -> Image3D:
    """loadDatImage (marchingCubes/util/LoadImage.h:177-206): three little-endian uint16
    dimensions, then nx*ny*nz little-endian uint16 samples, x fastest; unit spacing, zero origin.
    Samples missing at the end of a short file read as 0, as the reference's failed reads leave
    them."""
    if not os.path.exists(path):
        raise FileNotFound(path)
    raw = np.fromfile(path, dtype="<u2")
    if raw.size < 3:
        raise BadFormat(f"{path}: no dimensions")
    nx, ny, nz = (int(d) for d in raw[:3])
    n = nx * ny * nz
    data = np.zeros(n, np.float32)
    got = raw[3:3 + n]
    data[:got.size] = got
    return Image3D(data.reshape(nz, ny, nx))


def load_image(path: str) -> Image3D:
    """util::loadImage (flyingEdges/util/LoadImage.h:168-198)."""
    if not os.path.exists(path):
        raise FileNotFound(path)
    with open(path, "rb") as f:
        def line():
            return f.readline().decode("latin-1").rstrip("\n")
        line()
        line()
        if line() != "BINARY":
            raise BadFormat("Only 'BINARY' format supported")
        toks = line().split()
        if len(toks) < 2 or toks[0] != "DATASET" or toks[1] != "STRUCTURED_POINTS":
            raise BadFormat("Expecting STRUCTURED_POINTS dataset")
        dims = spacing = origin = None
        for _ in range(3):
            toks = line().split()
            tag = toks[0] if toks else ""
            try:
                if tag == "DIMENSIONS":
                    dims = [int(t) for t in toks[1:4]]
                elif tag == "SPACING":
                    spacing = [float(np.float32(float(t))) for t in toks[1:4]]
                elif tag == "ORIGIN":
                    origin = [float(np.float32(float(t))) for t in toks[1:4]]
                else:
                    raise BadFormat("Expecting DIMENSIONS, SPACING and ORIGIN")
            except ValueError as e:
                raise BadFormat(f"Expecting {tag} [3]") from e
        if dims is None or spacing is None or origin is None:
            raise BadFormat("Expecting DIMENSIONS, SPACING and ORIGIN")
        toks = line().split()
        if len(toks) < 2 or toks[0] != "POINT_DATA":
            raise BadFormat("Expecting POINT_DATA <npoints>")
        npoints = int(toks[1])
        toks = line().split()
        if len(toks) < 1 or toks[0] != "SCALARS":
            raise BadFormat("Expecting SCALARS <name> <type>")
        typ = toks[2] if len(toks) > 2 else ""
        if typ not in _VTK_TYPES:
            raise BadFormat("Unsupported data type")
        toks = line().split()
        if len(toks) < 1 or toks[0] != "LOOKUP_TABLE":
            raise BadFormat("Expecting LOOKUP_TABLE name")
        dt = np.dtype(_VTK_TYPES[typ])
        raw = f.read(npoints * dt.itemsize)
    vals = np.frombuffer(raw, dtype=dt, count=len(raw) // dt.itemsize).astype(np.float32)
    nx, ny, nz = dims
    data = np.zeros(nx * ny * nz, np.float32)
    n = min(vals.size, data.size)
    data[:n] = vals[:n]
    return Image3D(data.reshape(nz, ny, nx), tuple(spacing), tuple(origin))


def save_mesh(path: str, mesh: TriangleMesh) -> None:
    """util::saveTriangleMesh (flyingEdges/util/SaveTriangleMesh.h:33-118)."""
    pts = np.ascontiguousarray(mesh.points, dtype=np.float32).reshape(-1, 3)
    nrm = np.ascontiguousarray(mesh.normals, dtype=np.float32).reshape(-1, 3)
    tris = np.asarray(mesh.triangles).reshape(-1, 3).astype(np.uint64)
    v, t = pts.shape[0], tris.shape[0]
    poly = np.empty((t, 4), dtype=">u8")
    poly[:, 0] = 3
    poly[:, 1:] = tris
    with open(path, "wb") as f:
        f.write(b"# vtk DataFile Version 3.0\nIsosurface Mesh\nBINARY\nDATASET POLYDATA\n")
        f.write(f"POINTS {v} float\n".encode())
        f.write(pts.astype(">f4").tobytes())
        f.write(b"\n")
        f.write(f"POLYGONS {t} {4 * t}\n".encode())
        f.write(poly.tobytes())
        f.write(b"\n")
        f.write(f"POINT_DATA {v}\nNORMALS Normals float\n".encode())
        f.write(nrm.astype(">f4").tobytes())
        f.write(b"\n")


def load_mesh(path: str) -> TriangleMesh:
    """Reads a mesh in the format save_mesh (and the reference writer) produces."""
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0

    def line():
        nonlocal pos
        e = buf.index(b"\n", pos)
        s = buf[pos:e].decode("latin-1")
        pos = e + 1
        return s

    for _ in range(4):
        line()
    toks = line().split()
    v = int(toks[1])
    pts = np.frombuffer(buf, dtype=">f4", count=3 * v, offset=pos).astype(np.float32).reshape(v, 3)
    pos += 12 * v
    line()
    toks = line().split()
    t = int(toks[1])
    poly = np.frombuffer(buf, dtype=">u8", count=4 * t, offset=pos).astype(np.uint64).reshape(t, 4)
    pos += 32 * t
    line()
    line()
    line()
    nrm = np.frombuffer(buf, dtype=">f4", count=3 * v, offset=pos).astype(np.float32).reshape(v, 3)
    return TriangleMesh(pts, nrm, poly[:, 1:].copy())
