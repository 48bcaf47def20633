"""GPU parity of Marching Cubes against the serial-MC oracle and the reference fixtures:
vertex numbering (first-visit order), points, normals and triangles bit-exact."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, bits_equal, golden_files, load_fixture, max_ulp
from test_fe_gpu import CASES, assert_mesh_equal

pytestmark = pytest.mark.gpu


def mc_gpu(vol, iso, sp=(1, 1, 1), org=(0, 0, 0)):
    import miniisosurface_amd as m
    return m.MarchingCubes(m.Image3D(vol, sp, org), iso)


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: os.path.basename(p)[:-4])
def test_mc_matches_reference_fixture(path):
    vol, iso, sp, org, d = load_fixture(path)
    print(f"[mc fixture] {os.path.basename(path)}", flush=True)
    mesh = mc_gpu(vol, iso, sp, org)
    assert_mesh_equal(mesh, d["mc_points"], d["mc_normals"], d["mc_triangles"], os.path.basename(path))


def test_mc_all_256_cases():
    d = np.load(os.path.join(GOLDEN, "cases256.npz"), allow_pickle=False)
    vo = np.concatenate([[0], np.cumsum(d["mc_vcount"])])
    to = np.concatenate([[0], np.cumsum(d["mc_tcount"])])
    for c in range(256):
        mesh = mc_gpu(d["volumes"][c], 0.5)
        assert_mesh_equal(mesh, d["mc_points"][vo[c]:vo[c + 1]], d["mc_normals"][vo[c]:vo[c + 1]],
                          d["mc_triangles"][to[c]:to[c + 1]], f"case {c}")


@pytest.mark.parametrize("name,vol,iso,sp,org", CASES, ids=[c[0] for c in CASES])
def test_mc_matches_oracle(oracle_mod, name, vol, iso, sp, org):
    print(f"[mc oracle] {name}", flush=True)
    ref = oracle_mod.MCResult(vol, iso, sp, org)
    mesh = mc_gpu(vol, iso, sp, org)
    assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, name)


def test_mc_openmp_signature_accepts_sections(oracle_mod):
    import miniisosurface_amd as m
    vol = np.random.default_rng(3).uniform(0, 1, (12, 13, 14)).astype(np.float32)
    mesh = m.MarchingCubes(m.Image3D(vol), 0.5, 2, 2, 4)
    ref = oracle_mod.MCResult(vol, 0.5)
    assert_mesh_equal(mesh, ref.points, ref.normals, ref.triangles, "sections")


def mc_edge_gpu(vol, iso, sp=(1, 1, 1), org=(0, 0, 0)):
    import miniisosurface_amd as m
    return m.MarchingCubes(m.Image3D(vol, sp, org), iso, order="edge")


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: os.path.basename(p)[:-4])
def test_mc_edge_order_matches_dupfree_fixture(path):
    """order="edge" against the duplicate-free OpenMP driver's own output (one section)."""
    vol, iso, sp, org, d = load_fixture(path)
    print(f"[mc edge fixture] {os.path.basename(path)}", flush=True)
    mesh = mc_edge_gpu(vol, iso, sp, org)
    assert_mesh_equal(mesh, d["mcdf_points"], d["mcdf_normals"], d["mcdf_triangles"], os.path.basename(path))


def test_mc_edge_order_all_256_cases():
    d = np.load(os.path.join(GOLDEN, "cases256.npz"), allow_pickle=False)
    vo = np.concatenate([[0], np.cumsum(d["mcdf_vcount"])])
    to = np.concatenate([[0], np.cumsum(d["mcdf_tcount"])])
    for c in range(256):
        mesh = mc_edge_gpu(d["volumes"][c], 0.5)
        assert_mesh_equal(mesh, d["mcdf_points"][vo[c]:vo[c + 1]], d["mcdf_normals"][vo[c]:vo[c + 1]],
                          d["mcdf_triangles"][to[c]:to[c + 1]], f"case {c}")


@pytest.mark.parametrize("name,vol,iso,sp,org", CASES, ids=[c[0] for c in CASES])
def test_mc_edge_order_matches_oracle(oracle_mod, name, vol, iso, sp, org):
    """Every shape class, both kernel paths (rows of at most 4096 voxels and wider ones)."""
    print(f"[mc edge oracle] {name}", flush=True)
    p, n, t = oracle_mod.MCResult(vol, iso, sp, org).edge_order()
    mesh = mc_edge_gpu(vol, iso, sp, org)
    assert_mesh_equal(mesh, p, n, t, name)


def test_mc_order_switch_on_one_handle(oracle_mod):
    """One handle run in edge order, then first-visit order, then edge order again."""
    import miniisosurface_amd as m
    vol = np.random.default_rng(31).uniform(0, 1, (14, 15, 16)).astype(np.float32)
    ref = oracle_mod.MCResult(vol, 0.5)
    p, n, t = ref.edge_order()
    mc = m._MC(m.Image3D(vol), 0.5, order="edge")
    mc.run()
    assert_mesh_equal(mc.mesh(), p, n, t, "edge")
    mc.set_order("first_visit")
    mc.run()
    assert_mesh_equal(mc.mesh(), ref.points, ref.normals, ref.triangles, "first visit")
    mc.set_order("edge")
    mc.run()
    assert_mesh_equal(mc.mesh(), p, n, t, "edge again")
    with pytest.raises(ValueError):
        m._MC(m.Image3D(vol), 0.5, order="sorted")
    assert m.lib().miso_mc_set_order(mc._h, 2) != 0            # the C ABI rejects other values


@pytest.mark.slow
def test_mc_c2_gyroid_512_both_orders(oracle_mod):
    """The 512^3 gyroid with noise (seed 0) at isovalue 0: Marching Cubes in the serial
    reference's first-visit order and in the duplicate-free variant's edge order, each mesh bit
    for bit against the oracle (V = 10,804,576 and T = 21,539,597, as the reference gives)."""
    import miniisosurface_amd as m
    from miniisosurface_amd import synth
    vol = synth.gyroid(512, 0)
    ref = oracle_mod.MCResult(vol, 0.0)
    assert (ref.nverts, ref.ntris) == (10804576, 21539597)
    mc = m._MC(m.Image3D(vol), 0.0)
    mc.run()
    print("[mc 512] first visit", flush=True)
    assert_mesh_equal(mc.mesh(), ref.points, ref.normals, ref.triangles, "C2 first visit")
    p, n, t = ref.edge_order()
    del ref
    mc.set_order("edge")
    mc.run()
    print("[mc 512] edge order", flush=True)
    assert_mesh_equal(mc.mesh(), p, n, t, "C2 edge order")
