"""Command-line tools with the miniapp's flags, report and file formats, built in-tree:

    lib/miniiso_flying_edges   -i volume.vtk -o mesh.vtk -v iso [-y report.yaml]
                               (flyingEdges/serial/main.cpp:27-205)
    lib/miniiso_marching_cubes -i volume.vtk | -input_dat volume.dat -o mesh.vtk -v iso [-y ...]
                               (marchingCubes/serial/main.cpp:194-330)
    lib/miniiso_same_contents  a.vtk b.vtk     (utilities/tests/SameContentsCheck.cpp:190-349)

The drivers reach the GPU through include/miniiso.hpp (the C++ shim over the C ABI) and link
lib/libminiiso.so with an $ORIGIN run path, so the lib/ folder can be moved as a whole.
"""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG)
APPS = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")
CXX = os.environ.get("CXX", "g++")

TOOLS = {
    "miniiso_flying_edges": ("flying_edges.cpp", True),
    "miniiso_marching_cubes": ("marching_cubes.cpp", True),
    "miniiso_same_contents": ("same_contents.cpp", False),
}


def tool_path(name: str) -> str:
    return os.path.join(LIBDIR, name)


def build(verbose: bool = False) -> None:
    """Compiles the tools with the host C++ compiler against the built libminiiso.so."""
    os.makedirs(LIBDIR, exist_ok=True)
    for name, (src, needs_lib) in TOOLS.items():
        out = tool_path(name)
        cmd = [CXX, "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", f"-I{APPS}",
               os.path.join(APPS, src), "-o", out + ".tmp"]
        if needs_lib:
            cmd += [f"-L{LIBDIR}", "-lminiiso", "-Wl,-rpath,$ORIGIN", "-Wl,--enable-new-dtags"]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
