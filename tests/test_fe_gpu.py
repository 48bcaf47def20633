"""GPU parity of Flying Edges (HIP kernels through the C ABI) against the oracle and the fixtures.

Bar: per-row counts, trims, offsets, cube cases, triangle connectivity and order bit-exact;
points and normals bit-exact (0 ULP: the kernels evaluate the reference's float formulas in the
same order with no fused multiply-add and IEEE division)."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, bits_equal, golden_files, load_fixture, max_ulp

pytestmark = pytest.mark.gpu


def fe_gpu(vol, iso, sp=(1, 1, 1), org=(0, 0, 0), index_bits=None, passes=True):
    import miniisosurface_amd as m
    img = m.Image3D(vol, sp, org)
    algo = m.FlyingEdgesAlgorithm(img, iso)
    if index_bits:
        algo.set_index_bits(index_bits)
    if passes:
        algo.pass1(); algo.pass2(); algo.pass3(); algo.pass4()
    else:
        algo.run()
    return algo


def assert_mesh_equal(mesh, pts, nrm, tris, tag=""):
    assert mesh.points.shape == pts.shape, tag
    assert mesh.triangles.shape == tris.shape, tag
    assert bits_equal(mesh.triangles, tris), tag
    if not bits_equal(mesh.points, pts):
        raise AssertionError(f"{tag}: points differ, max ulp {max_ulp(mesh.points, pts)}")
    if not bits_equal(mesh.normals, nrm):
        raise AssertionError(f"{tag}: normals differ, max ulp {max_ulp(mesh.normals, nrm)}")


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: os.path.basename(p)[:-4])
def test_fe_matches_reference_fixture(path):
    vol, iso, sp, org, d = load_fixture(path)
    print(f"[fe fixture] {os.path.basename(path)} dims={vol.shape[::-1]}", flush=True)
    algo = fe_gpu(vol, iso, sp, org)
    mesh = algo.moveOutput()
    assert_mesh_equal(mesh, d["fe_points"], d["fe_normals"], d["fe_triangles"], os.path.basename(path))


def test_fe_all_256_cases():
    d = np.load(os.path.join(GOLDEN, "cases256.npz"), allow_pickle=False)
    vo = np.concatenate([[0], np.cumsum(d["fe_vcount"])])
    to = np.concatenate([[0], np.cumsum(d["fe_tcount"])])
    for c in range(256):
        mesh = fe_gpu(d["volumes"][c], 0.5).moveOutput()
        assert_mesh_equal(mesh, d["fe_points"][vo[c]:vo[c + 1]], d["fe_normals"][vo[c]:vo[c + 1]],
                          d["fe_triangles"][to[c]:to[c + 1]], f"case {c}")


def _cases():
    rng = np.random.default_rng(77)
    from miniisosurface_amd import synth
    yield "rand_67x33x21", rng.uniform(0, 1, (21, 33, 67)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "rand_130x7x5_sp", rng.uniform(0, 1, (5, 7, 130)).astype(np.float32), 0.42, (0.3, 1.7, 2.2), (-4.0, 1.0, 3.5)
    yield "rand_1030x4x3", rng.uniform(0, 1, (3, 4, 1030)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "rand_2100x3x2", rng.uniform(0, 1, (2, 3, 2100)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    # rows of 64 mask words (the widest the (row, word) lane kernels take) and just wider (the
    # per-row kernels)
    yield "rand_4096x3x3", rng.uniform(0, 1, (3, 3, 4096)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "rand_4097x3x2", rng.uniform(0, 1, (2, 3, 4097)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "rand_5000x4x3_sp", rng.uniform(0, 1, (3, 4, 5000)).astype(np.float32), 0.4, (0.5, 2.0, 1.5), (1.0, -2.0, 0.25)
    # rows of 9 to 16 mask words (4 rows per wavefront, 16 lanes per row, some lanes past the row
    # end), and blocks that run past the last row
    yield "rand_600x5x4", rng.uniform(0, 1, (4, 5, 600)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "rand_1024x7x6_sp", rng.uniform(0, 1, (6, 7, 1024)).astype(np.float32), 0.45, (0.8, 1.25, 0.6), (2.0, -1.0, 0.5)
    yield "gauss_777x40x12", synth.gaussians((777, 40, 12), 5, 21, sigma_range=(4.0, 12.0)), 0.4, (1, 1, 1), (0, 0, 0)
    yield "rand_513x33x3", rng.uniform(0, 1, (3, 33, 513)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    # pass 4 row blocks in the XCD tile order: 256-row blocks, 8 per plane, 48 of 56 blocks in
    # whole tile rows; 16-row blocks, tiles of 2 blocks, every block in a tile row
    yield "gauss_64x2048x7", synth.gaussians((64, 2048, 7), 4, 31, sigma_range=(3.0, 9.0)), 0.35, (1, 1, 1), (0, 0, 0)
    yield "rand_1024x256x9_sp", rng.uniform(0, 1, (9, 256, 1024)).astype(np.float32), 0.5, (0.5, 1.5, 1.0), (0.0, 1.0, -2.0)
    # empty but for isolated voxels at word boundaries (i = 63, 64, 127, 128, 129): pass 2 skips
    # the triangle count of wavefronts whose words are all below the isovalue
    iso_vox = np.zeros((7, 40, 130), np.float32)
    for (k, j, i) in [(1, 3, 63), (2, 17, 64), (3, 30, 127), (4, 31, 128), (5, 38, 129), (0, 0, 64), (6, 39, 0)]:
        iso_vox[k, j, i] = 1.0
    yield "isolated_130x40x7", iso_vox, 0.5, (1, 1, 1), (0, 0, 0)
    yield "checker_64", (np.indices((64, 64, 64)).sum(0) % 2).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    # every edge cut on rows of 4096 voxels: one wavefront per row writes about 12,000 vertices
    # (dozens of pass 4 staging windows) and 4,095 active cells per row (several cell rounds);
    # and rows of 1000 voxels, where 4 rows share a wavefront
    yield "checker_4096x4x3", (np.indices((3, 4, 4096)).sum(0) % 2).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "checker_1000x9x5_sp", (np.indices((5, 9, 1000)).sum(0) % 2).astype(np.float32), 0.5, (0.7, 1.3, 2.0), \
        (1.0, -1.0, 0.5)
    yield "gauss_96x80x72", synth.gaussians((96, 80, 72), 6, 3, sigma_range=(5.0, 14.0)), 0.3, (1, 1, 1), (0, 0, 0)
    yield "gyroid_128", synth.gyroid(128, 1), 0.0, (1, 1, 1), (0, 0, 0)
    yield "noise_128", synth.value_noise(128, 2), 0.5, (0.5, 0.5, 0.5), (1.0, 2.0, 3.0)
    v = rng.uniform(0, 1, (17, 19, 23)).astype(np.float32)
    v[rng.uniform(0, 1, v.shape) < 0.05] = np.nan
    yield "nan_23x19x17", v, 0.5, (1, 1, 1), (0, 0, 0)
    yield "all_above", np.ones((9, 10, 11), np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "flat_1x8x8", rng.uniform(0, 1, (8, 8, 1)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)
    yield "flat_8x8x1", rng.uniform(0, 1, (1, 8, 8)).astype(np.float32), 0.5, (1, 1, 1), (0, 0, 0)


CASES = list(_cases())


@pytest.mark.parametrize("name,vol,iso,sp,org", CASES, ids=[c[0] for c in CASES])
def test_fe_matches_oracle_with_row_state(oracle_mod, name, vol, iso, sp, org):
    print(f"[fe oracle] {name} dims={vol.shape[::-1]}", flush=True)
    ref = oracle_mod.FEResult(vol, iso, sp, org, nthreads=4)
    algo = fe_gpu(vol, iso, sp, org)
    assert algo.counts() == (ref.nverts, ref.ntris)
    mesh = algo.moveOutput()
    assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, name)
    if ref.rows is None:
        return
    rows = algo.rows()
    R = ref.rows
    for key in ("xl", "xr"):
        assert np.array_equal(rows[key].astype(np.uint64), R[key]), key
    assert np.array_equal(rows["xcount"].astype(np.uint64), R["xcount"])
    assert np.array_equal(rows["ycount"].astype(np.uint64), R["ycount"])
    assert np.array_equal(rows["zcount"].astype(np.uint64), R["zcount"])
    assert np.array_equal(rows["voffset"], R["xstart"])
    nx, ny, nz = vol.shape[::-1]
    cell_rows = np.array([j + ny * k for k in range(nz - 1) for j in range(ny - 1)], np.int64)
    assert np.array_equal(rows["tcount"][cell_rows].astype(np.uint64), R["tcount"])
    assert np.array_equal(rows["toffset"][cell_rows], R["toffset"])
    assert np.array_equal(algo.cube_cases(), R["cube_cases"])


def test_fe_index_bits_64_and_run_path(oracle_mod):
    name, vol, iso, sp, org = CASES[0]
    ref = oracle_mod.FEResult(vol, iso, sp, org)
    for bits, passes in ((64, True), (32, False), (64, False)):
        mesh = fe_gpu(vol, iso, sp, org, index_bits=bits, passes=passes).moveOutput()
        assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, f"bits={bits}")


def test_fe_handle_reuse_across_isovalues(oracle_mod):
    import miniisosurface_amd as m
    from miniisosurface_amd import synth
    vol = synth.gaussians((70, 60, 50), 4, 9, sigma_range=(6.0, 12.0))
    algo = m.FlyingEdgesAlgorithm(m.Image3D(vol), 0.2)
    for iso in (0.2, 0.7, 0.05, 0.5, 2.0):
        algo.set_isoval(iso)
        algo.run()
        mesh = algo.moveOutput()
        ref = oracle_mod.FEResult(vol, iso)
        assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, f"iso={iso}")


@pytest.mark.parametrize("case", ["rand_67x33x21", "rand_5000x4x3_sp"])
def test_fe_rerun_grows_and_switches_index_width(oracle_mod, case):
    """Repeated run() on one handle writes pass 4 into the previous buffers, checked on the
    device; a larger mesh or another index width must fall back to growing them. Both kernel
    paths (rows of at most 4096 voxels, and wider)."""
    import miniisosurface_amd as m
    name, vol, iso, sp, org = [c for c in CASES if c[0] == case][0]
    algo = m.FlyingEdgesAlgorithm(m.Image3D(vol, sp, org), iso)
    steps = ((iso, None), (iso, None), (0.95, None), (iso, None), (iso, 64), (iso, 64), (iso, 32), (0.05, None))
    forced64 = False
    for s_iso, bits in steps:
        if bits:
            algo.set_index_bits(bits)
            forced64 = bits == 64
        algo.set_isoval(s_iso)
        algo.run()
        ref = oracle_mod.FEResult(vol, s_iso, sp, org, want_debug=False)
        mesh = algo.moveOutput()
        assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, f"{name} iso={s_iso} bits={bits}")
        assert algo.device_output()[3] == (64 if forced64 else 32)


def test_fe_device_resident_volume(oracle_mod):
    import torch
    import miniisosurface_amd as m
    from miniisosurface_amd import synth
    vol = synth.gyroid(96, 4)
    t = torch.from_numpy(vol).cuda()
    algo = m.FlyingEdgesAlgorithm(m.DeviceVolume.from_tensor(t), 0.1)
    algo.run()
    mesh = algo.moveOutput()
    ref = oracle_mod.FEResult(vol, 0.1, nthreads=4)
    assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, "device volume")


def test_fe_out_of_order_pass_raises():
    import miniisosurface_amd as m
    vol = np.random.default_rng(0).uniform(0, 1, (8, 8, 8)).astype(np.float32)
    algo = m.FlyingEdgesAlgorithm(m.Image3D(vol), 0.5)
    with pytest.raises(m.MisoError):
        algo.pass3()
    with pytest.raises(m.MisoError):
        algo.moveOutput()


@pytest.mark.slow
def test_fe_c2_gyroid_512(oracle_mod):
    """512^3 gyroid + noise at iso 0: V = 10,804,576 and T = 21,539,597 as the reference gives,
    and the whole mesh bit-exact against the OpenMP oracle."""
    from miniisosurface_amd import synth
    vol = synth.gyroid(512, 0)
    algo = fe_gpu(vol, 0.0, passes=False)
    assert algo.counts() == (10804576, 21539597)
    mesh = algo.moveOutput()
    ref = oracle_mod.FEResult(vol, 0.0, nthreads=min(16, os.cpu_count() or 1), want_debug=False)
    assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, "C2")
