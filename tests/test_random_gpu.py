"""GPU: seeded random shapes, fields and isovalues, each extracted by Flying Edges, Marching Cubes
(first-visit order) and Marching Cubes (edge order), and compared bit for bit with the oracle;
the unit-spacing ones also as one batch. The shapes reach the kernels' corners: rows of up to
4100 voxels (up to 65 mask words, so both the lane-row and the general kernels), fewer grid rows
per plane than a wavefront holds rows (carries into later planes), thin volumes, isovalues on a
level of a quantized field (ties), nearly empty fields (wavefronts with nothing to count or
emit), smooth blobs, and non-unit spacing with an origin."""
import numpy as np
import pytest

from helpers import bits_equal

pytestmark = pytest.mark.gpu

NCASES = 160


def _field(rng, kind, shape):
    nz, ny, nx = shape
    if kind == 0:
        return rng.uniform(0, 1, shape).astype(np.float32)
    if kind == 1:   # quantized: the isovalue 0.5 sits on a level
        return (rng.integers(0, 5, shape) * 0.25).astype(np.float32)
    if kind == 2:   # nearly empty: a few voxels above
        v = np.zeros(shape, np.float32)
        v.reshape(-1)[rng.integers(0, v.size, max(1, v.size // 500))] = 1.0
        return v
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    v = np.zeros(shape, np.float64)
    for _ in range(int(rng.integers(1, 5))):
        c = rng.uniform(0, 1, 3) * np.array([nx, ny, nz])
        s = rng.uniform(2, 12)
        v += np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))
    return v.astype(np.float32)


def _cases(seed=7):
    rng = np.random.default_rng(seed)
    for case in range(NCASES):
        r = rng.uniform()
        if r < 0.15:
            dims = (int(rng.integers(900, 4100)), int(rng.integers(2, 12)), int(rng.integers(2, 8)))
        elif r < 0.4:
            dims = (int(rng.integers(2, 80)), int(rng.integers(2, 10)), int(rng.integers(2, 70)))
        else:
            dims = (int(rng.integers(2, 300)), int(rng.integers(2, 90)), int(rng.integers(2, 40)))
        nx, ny, nz = dims
        kind = int(rng.integers(0, 4))
        vol = _field(rng, kind, (nz, ny, nx))
        iso = 0.5 if kind == 1 else float(np.float32(rng.uniform(0.05, 0.95)))
        if rng.uniform() < 0.25:
            sp = tuple(float(np.float32(a)) for a in rng.uniform(0.3, 2.0, 3))
            org = tuple(float(np.float32(a)) for a in rng.uniform(-5, 5, 3))
        else:
            sp, org = (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
        yield case, dims, kind, vol, iso, sp, org


def _same(mesh, pts, nrm, tris):
    return bits_equal(mesh.points, pts) and bits_equal(mesh.normals, nrm) and np.array_equal(mesh.triangles, tris)


def test_random_shapes_against_oracle(oracle_mod):
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd.batch import BatchExtraction
    batch, nonempty = [], 0
    for case, dims, kind, vol, iso, sp, org in _cases():
        tag = f"case {case}: dims {dims} field {kind} iso {iso} spacing {sp} origin {org}"
        print(f"[random] {tag}", flush=True)
        fe = oracle_mod.FEResult(vol, iso, sp, org, nthreads=1, want_debug=False)
        a = m.FlyingEdgesAlgorithm(m.Image3D(vol, sp, org), iso)
        a.pass1()
        a.pass2()
        a.pass3()
        a.pass4()
        assert _same(a.moveOutput(), fe.points, fe.normals, fe.triangles), "Flying Edges " + tag
        mc = oracle_mod.MCResult(vol, iso, sp, org)
        assert _same(m.MarchingCubes(m.Image3D(vol, sp, org), iso), mc.points, mc.normals, mc.triangles), \
            "Marching Cubes " + tag
        assert _same(m.MarchingCubes(m.Image3D(vol, sp, org), iso, order="edge"), *mc.edge_order()), \
            "Marching Cubes edge order " + tag
        nonempty += fe.ntris > 0
        if sp == (1.0, 1.0, 1.0) and dims[0] <= 4096:
            batch.append((torch.from_numpy(vol).cuda(), iso, fe))
    assert nonempty > NCASES // 2
    # the unit-spacing volumes as one batch: every job equals its oracle mesh
    print(f"[random] batch of {len(batch)} jobs", flush=True)
    b = BatchExtraction([(m.DeviceVolume.from_tensor(t), iso) for t, iso, _ in batch])
    b.run()
    for q, (_, _, fe) in enumerate(batch):
        assert _same(b.mesh(q), fe.points, fe.normals, fe.triangles), f"batch job {q}"
    b.close()
