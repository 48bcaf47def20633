"""GPU, beyond 2^32 voxels: a 2048^3 gyroid (8.6 G voxels, 32 GiB) extracted whole and in z-slabs.

The whole-volume run indexes voxels, rows and vertex ids far beyond 32 bits; the slab run covers
the same volume with small, slab-local offsets. Their parts must be identical (compared on the
device), and the whole run's per-row counts and triangle total must equal what torch computes
from the sign field directly (sign changes along each axis per row; the case table's triangle
count summed over all cells)."""
import ctypes

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


class _DeviceArray:
    """__cuda_array_interface__ over a raw device pointer (no ownership)."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 3}


def _view(torch, ptr, rows, typestr):
    return torch.as_tensor(_DeviceArray(ptr, 3 * rows, typestr), device="cuda").view(rows, 3)


def test_fe_2048_whole_equals_slabs_and_counts(oracle_mod):
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd import synth

    n, iso = 2048, 0.1
    vol = synth.gyroid(n, seed=0, periods=16, noise=0.0, device="cuda")
    dv = m.DeviceVolume.from_tensor(vol)
    whole = m.FlyingEdgesAlgorithm(dv, iso)
    whole.run()
    wp, wn, wt, bits, V, T = whole.device_output()
    print(f"[2048] V={V} T={T} index_bits={bits}", flush=True)
    assert V > 2 ** 28 and bits == 32
    itype = "<u4" if bits == 32 else "<u8"
    Wp, Wn = _view(torch, wp, V, "<f4"), _view(torch, wn, V, "<f4")
    Wt = _view(torch, wt, T, "<i4" if bits == 32 else "<i8")

    # the same mesh slab by slab (300-plane slabs: 7 of them, the last shorter)
    slabs = m.FlyingEdgesSlabs(dv, iso, 300)
    nslabs = 0
    while slabs.next():
        part = slabs.part()
        pp, pn, pt = slabs.device_part()
        vb, nv, tb, nt = part["vbase"], part["nverts"], part["tbase"], part["ntris"]
        assert torch.equal(_view(torch, pp, nv, "<f4").view(torch.int32), Wp[vb:vb + nv].view(torch.int32))
        assert torch.equal(_view(torch, pn, nv, "<f4").view(torch.int32), Wn[vb:vb + nv].view(torch.int32))
        assert torch.equal(_view(torch, pt, nt, "<i8"), Wt[tb:tb + nt].to(torch.int64))
        nslabs += 1
        print(f"[2048] slab {nslabs} planes [{part['k0']}, {part['k1']}) ok", flush=True)
    assert nslabs == 7 and vb + nv == V and tb + nt == T
    del slabs

    # per-row vertex counts and the triangle total from the sign field
    rows = whole.rows()
    L = oracle_mod.lib()
    L.ora_tables_init()
    ntri = torch.tensor(np.frombuffer((ctypes.c_ubyte * 256).in_dll(L, "ora_num_tris"), dtype=np.uint8).astype(np.int64),
                        device="cuda")
    xs = np.zeros((n, n), np.int64)
    ys = np.zeros((n, n), np.int64)
    zs = np.zeros((n, n), np.int64)
    total = 0
    for k0 in range(0, n, 128):
        k1 = min(k0 + 128, n)
        s = vol[k0:min(k1 + 1, n)] >= iso                       # planes k0 .. k1 (one more for z)
        xs[k0:k1] = (s[:k1 - k0, :, 1:] != s[:k1 - k0, :, :-1]).sum(dim=2).cpu().numpy()
        ys[k0:k1, :-1] = (s[:k1 - k0, 1:, :] != s[:k1 - k0, :-1, :]).sum(dim=2).cpu().numpy()
        if s.shape[0] > k1 - k0:
            zs[k0:k1] = (s[1:] != s[:-1]).sum(dim=2).cpu().numpy()
        else:
            zs[k0:k1 - 1] = (s[1:] != s[:-1]).sum(dim=2).cpu().numpy()
        for c0 in range(0, s.shape[0] - 1, 16):
            a = s[c0:c0 + 17].to(torch.int64)
            lo, hi = a[:-1], a[1:]
            case = (lo[:, :-1, :-1] | lo[:, :-1, 1:] << 1 | lo[:, 1:, 1:] << 2 | lo[:, 1:, :-1] << 3
                    | hi[:, :-1, :-1] << 4 | hi[:, :-1, 1:] << 5 | hi[:, 1:, 1:] << 6 | hi[:, 1:, :-1] << 7)
            total += int(ntri[case].sum())
            del a, lo, hi, case
        del s
    assert np.array_equal(rows["xcount"].astype(np.int64), xs.reshape(-1))
    assert np.array_equal(rows["ycount"].astype(np.int64), ys.reshape(-1))
    assert np.array_equal(rows["zcount"].astype(np.int64), zs.reshape(-1))
    assert V == int(xs.sum() + ys.sum() + zs.sum())
    assert T == total
