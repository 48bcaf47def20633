"""Order-independent mesh equivalence, with the miniapp checker's rule
(utilities/tests/SameContentsCheck.cpp:190-335):

  - equal triangle counts;
  - every point of each mesh occurs among the other's points, and likewise every normal
    (float ==: -0 equals 0, NaN matches nothing);
  - every triangle of the first, as its three corner positions in order, occurs among the
    second's triangles.

Vertex counts and numbering may differ. Returns (equivalent, reason).
"""
from __future__ import annotations

import numpy as np


def _keys(a):
    """Rows of float32 values as hashable byte keys under == equality (-0 -> +0)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    a = a + np.float32(0.0)           # -0.0 + 0.0 == +0.0
    return a.view(np.uint32)


def _rows_in(a, b):
    """Every row of a occurs among the rows of b."""
    if a.shape[0] == 0:
        return True
    if b.shape[0] == 0:
        return False
    w = a.shape[1]
    dt = np.dtype((np.void, 4 * w))
    av = np.ascontiguousarray(a).view(dt).ravel()
    bv = np.unique(np.ascontiguousarray(b).view(dt).ravel())
    return bool(np.isin(av, bv).all())


def same_contents(mesh_a, mesh_b):
    ta, tb = np.asarray(mesh_a.triangles), np.asarray(mesh_b.triangles)
    if ta.shape[0] != tb.shape[0]:
        return False, "not the same number of triangles"
    pa, pb = np.asarray(mesh_a.points, np.float32), np.asarray(mesh_b.points, np.float32)
    na, nb = np.asarray(mesh_a.normals, np.float32), np.asarray(mesh_b.normals, np.float32)
    if np.isnan(pa).any() or np.isnan(pb).any():
        return False, "point not found"
    ka, kb = _keys(pa), _keys(pb)
    if not (_rows_in(ka, kb) and _rows_in(kb, ka)):
        return False, "point not found"
    if np.isnan(na).any() or np.isnan(nb).any():
        return False, "normal not found"
    ma, mb = _keys(na), _keys(nb)
    if not (_rows_in(ma, mb) and _rows_in(mb, ma)):
        return False, "normal not found"
    ca = pa[ta.astype(np.int64)].reshape(-1, 9) if ta.size else np.zeros((0, 9), np.float32)
    cb = pb[tb.astype(np.int64)].reshape(-1, 9) if tb.size else np.zeros((0, 9), np.float32)
    if not _rows_in(_keys(ca), _keys(cb)):
        return False, "triangle not found"
    return True, ""
