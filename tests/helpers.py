"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits_equal(a, b) -> bool:
    """Bit-exact equality of float32 arrays or integer arrays. A NaN must sit where the other
    array has a NaN; its sign and payload are not compared (IEEE 754 leaves them to the
    hardware: x86 SSE produces 0xFFC00000, the GPU the canonical 0x7FC00000)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == np.float32 and b.dtype == np.float32:
        na, nb = np.isnan(a), np.isnan(b)
        if not np.array_equal(na, nb):
            return False
        return np.array_equal(a.view(np.uint32)[~na], b.view(np.uint32)[~nb])
    return np.array_equal(a.astype(np.uint64), b.astype(np.uint64))


def max_ulp(a, b) -> int:
    """Largest distance in units of the last place between two float32 arrays (NaN == NaN)."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    if a.size == 0:
        return 0
    both_nan = np.isnan(a) & np.isnan(b)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    d = np.abs(ia - ib)
    d[both_nan] = 0
    return int(d.max())


def golden_files():
    return sorted(f for f in glob.glob(os.path.join(GOLDEN, "*.npz")) if not f.endswith("cases256.npz"))


def load_fixture(path):
    """Returns (volume, iso, spacing, origin, dict of arrays); regenerates inputs not stored."""
    d = dict(np.load(path, allow_pickle=False))
    if "volume" in d:
        vol = d["volume"]
    else:
        from miniisosurface_amd import synth
        name = os.path.basename(path)
        if name.startswith("c1_gaussians_128"):
            vol = synth.gaussians(128, 4, 0)
        else:
            raise KeyError(name)
    return vol, float(d["iso"]), tuple(float(x) for x in d["spacing"]), tuple(float(x) for x in d["origin"]), d
