"""Seeded synthetic volumes for the benchmark and test workloads.

All fields are float32, unit spacing, zero origin, laid out (nz, ny, nx) with x fastest, exactly
like the miniapp's Image3D (flyingEdges/util/Image3D.h:27-86). Values are computed in float64 and
rounded once to float32. The generators are defined by what they compute, so two runs with the
same seed on the same host produce the same bits:

  gaussians(n, count, seed)   sum of isotropic Gaussians, amplitude 1; for each blob the centre
                              (x, y, z) ~ U[0, dim-1] is drawn, then sigma ~ U[8, 20]
                              (numpy default_rng(seed)). 128^3, 4 blobs, seed 0, iso 0.5 gives
                              10,790 vertices and 21,242 triangles.
  gyroid(n, seed)             sin(kx)cos(ky) + sin(ky)cos(kz) + sin(kz)cos(kx), k = 2*pi*8/n on
                              voxel indices, plus U(-0.05, 0.05) noise (default_rng(seed), one
                              draw per voxel in memory order).
  value_noise(n, seed)        3 octaves; octave o has a (16*2^o + 1)^3 lattice of U(0, 1) values
                              (default_rng(seed), octaves in order), amplitude 2^-o, trilinear
                              interpolation at voxel i -> lattice coordinate i * cells / n;
                              the sum is divided by 1.75.
  batch_volumes(count, seed)  dims log-uniform in [32, 256] per axis, 1-6 Gaussians each with
                              sigma scaled to the volume, isovalue U[0.2, 0.8].
  serving_pool(count, seed)   the serving workload's resident volumes: cubic dims drawn from
                              {64, 128, 256, 384, 512}, each a Gaussian-blob field (1-6 blobs,
                              sigma scaled to the volume) or a gyroid, with equal odds.
"""
from __future__ import annotations

import math

import numpy as np


def _dims3(n):
    if isinstance(n, int):
        return (n, n, n)
    return tuple(int(v) for v in n)


def gaussians(n=128, count=4, seed=0, sigma_range=(8.0, 20.0), device=None):
    """device="cuda" evaluates the same float64 formula with torch on the GPU and returns a
    float32 tensor there (the GPU's exp may differ from the host's in the last float64 bit, so
    a rare voxel can round to a neighbouring float32)."""
    nx, ny, nz = _dims3(n)
    rng = np.random.default_rng(seed)
    blobs = []
    for _ in range(count):
        c = rng.uniform(0, [nx - 1, ny - 1, nz - 1], size=3)
        s = rng.uniform(*sigma_range)
        blobs.append((c, s))
    if device is not None:
        import torch
        x = torch.arange(nx, dtype=torch.float64, device=device)
        y = torch.arange(ny, dtype=torch.float64, device=device)
        z = torch.arange(nz, dtype=torch.float64, device=device)
        f = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
        for c, s in blobs:
            d2 = ((x - c[0]) ** 2)[None, None, :] + ((y - c[1]) ** 2)[None, :, None] \
                + ((z - c[2]) ** 2)[:, None, None]
            f += torch.exp(-d2 / (2 * s ** 2))
        return f.to(torch.float32)
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    z = np.arange(nz, dtype=np.float64)
    f = np.zeros((nz, ny, nx), np.float64)
    for c, s in blobs:
        d2 = ((x - c[0]) ** 2)[None, None, :] + ((y - c[1]) ** 2)[None, :, None] \
            + ((z - c[2]) ** 2)[:, None, None]
        f += np.exp(-d2 / (2 * s ** 2))
    return f.astype(np.float32)


def gyroid(n=512, seed=0, periods=8, noise=0.05, device=None):
    """device="cuda": the trigonometric part in float64 on the GPU, the same host noise draws."""
    nx, ny, nz = _dims3(n)
    k = 2.0 * math.pi * periods / nx
    x = np.arange(nx, dtype=np.float64) * k
    y = np.arange(ny, dtype=np.float64) * k
    z = np.arange(nz, dtype=np.float64) * k
    sx, cx = np.sin(x), np.cos(x)
    sy, cy = np.sin(y), np.cos(y)
    sz, cz = np.sin(z), np.cos(z)
    rng = np.random.default_rng(seed)
    if device is not None:
        import torch
        T = lambda a: torch.from_numpy(a).to(device)
        tsx, tcy, tsy, tcx = T(sx), T(cy), T(sy), T(cx)
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device=device)
        for kk in range(nz):
            f = tsx[None, :] * tcy[:, None] + tsy[:, None] * float(cz[kk]) + float(sz[kk]) * tcx[None, :]
            if noise:
                f = f + T(rng.uniform(-noise, noise, size=(ny, nx)))
            out[kk] = f.to(torch.float32)
        return out
    out = np.empty((nz, ny, nx), np.float32)
    for kk in range(nz):  # plane by plane keeps float64 temporaries small
        f = sx[None, :] * cy[:, None] + sy[:, None] * cz[kk] + sz[kk] * cx[None, :]
        if noise:
            f = f + rng.uniform(-noise, noise, size=(ny, nx))
        out[kk] = f
    return out


def _lerp_matrix(n, cells):
    """(n, cells+1) matrix of trilinear weights along one axis: voxel i -> i*cells/n."""
    u = np.arange(n, dtype=np.float64) * (cells / n)
    i0 = np.minimum(np.floor(u).astype(np.int64), cells - 1)
    f = u - i0
    w = np.zeros((n, cells + 1), np.float64)
    w[np.arange(n), i0] = 1.0 - f
    w[np.arange(n), i0 + 1] += f
    return w


def value_noise(n=1024, seed=0, octaves=3, base_cells=16, device=None):
    """Returns a float32 volume. With device set (e.g. 'cuda'), builds it with torch on that
    device and returns a torch tensor (used for volumes too large to build on the host quickly)."""
    nx, ny, nz = _dims3(n)
    rng = np.random.default_rng(seed)
    lattices = []
    for o in range(octaves):
        cells = base_cells * (2 ** o)
        lattices.append((cells, 2.0 ** -o, rng.uniform(0.0, 1.0, size=(cells + 1,) * 3)))
    if device is None:
        f = np.zeros((nz, ny, nx), np.float64)
        for cells, amp, lat in lattices:
            wx, wy, wz = _lerp_matrix(nx, cells), _lerp_matrix(ny, cells), _lerp_matrix(nz, cells)
            t = np.einsum("cba,ia->cbi", lat, wx)
            t = np.einsum("cbi,jb->cji", t, wy)
            f += amp * np.einsum("cji,kc->kji", t, wz)
        return (f / 1.75).astype(np.float32)
    import torch
    dev = torch.device(device)
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    tw = []
    for cells, amp, lat in lattices:
        wx = torch.from_numpy(_lerp_matrix(nx, cells)).to(dev)
        wy = torch.from_numpy(_lerp_matrix(ny, cells)).to(dev)
        wz = torch.from_numpy(_lerp_matrix(nz, cells)).to(dev)
        t = torch.from_numpy(lat).to(dev)
        t = torch.einsum("cba,ia->cbi", t, wx)
        t = torch.einsum("cbi,jb->cji", t, wy)         # (cells+1, ny, nx)
        tw.append((amp, wz, t))
    slab = 64
    for k0 in range(0, nz, slab):
        k1 = min(nz, k0 + slab)
        acc = torch.zeros((k1 - k0, ny, nx), dtype=torch.float64, device=dev)
        for amp, wz, t in tw:
            acc += amp * torch.einsum("cji,kc->kji", t, wz[k0:k1].contiguous())
        out[k0:k1] = (acc / 1.75).to(torch.float32)
    return out


def batch_volumes(count=256, seed=0, lo=32, hi=256, device=None):
    """Mixed-size batch: list of (volume, isovalue)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        dims = np.exp(rng.uniform(math.log(lo), math.log(hi), size=3)).astype(np.int64)
        dims = np.clip(dims, lo, hi)
        nb = int(rng.integers(1, 7))
        sub_seed = int(rng.integers(0, 2 ** 31 - 1))
        iso = float(rng.uniform(0.2, 0.8))
        scale = float(min(dims)) / 128.0
        vol = gaussians(tuple(int(d) for d in dims), count=nb, seed=sub_seed,
                        sigma_range=(8.0 * scale, 20.0 * scale), device=device)
        out.append((vol, iso))
    return out


SERVING_DIMS = (64, 128, 256, 384, 512)


def serving_pool(count=16, seed=0, device=None):
    """Resident volumes of the serving workload: list of (volume, (min, max), n). A volume is an
    n^3 Gaussian-blob field (1-6 blobs, sigma U[8, 20] * n / 128) or an n^3 gyroid with noise,
    each with probability 1/2; n is drawn from SERVING_DIMS (numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(SERVING_DIMS))
        if rng.uniform() < 0.5:
            v = gaussians(n, count=int(rng.integers(1, 7)), seed=int(rng.integers(1 << 30)),
                          sigma_range=(8.0 * n / 128, 20.0 * n / 128), device=device)
        else:
            v = gyroid(n, seed=int(rng.integers(1 << 30)), device=device)
        out.append((v, (float(v.min()), float(v.max())), n))
    return out
