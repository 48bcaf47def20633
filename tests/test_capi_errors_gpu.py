"""GPU: the C ABI's error behaviour -- invalid arguments and calls out of order return a negative
status and a message (no exception crosses the boundary, nothing crashes), and a handle stays
usable after a refused call."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vol(data, dims, on_device=0):
    from miniisosurface_amd._lib import Volume
    v = Volume()
    v.data = data
    v.nx, v.ny, v.nz = dims
    for a in range(3):
        v.spacing[a] = 1.0
        v.origin[a] = 0.0
    v.on_device = on_device
    return v


def test_create_rejects_bad_volumes():
    from miniisosurface_amd._lib import lib
    L = lib()
    h = ctypes.c_void_p()
    assert L.miso_fe_create(None, ctypes.c_float(0.5), None, ctypes.byref(h)) == -3
    assert b"NULL" in L.miso_last_error()
    assert L.miso_fe_create(ctypes.byref(_vol(None, (4, 4, 4))), ctypes.c_float(0.5), None, ctypes.byref(h)) == -3
    assert L.miso_fe_create(ctypes.byref(_vol(None, (-1, 4, 4))), ctypes.c_float(0.5), None, ctypes.byref(h)) == -3
    assert L.miso_mc_create(None, ctypes.c_float(0.5), None, ctypes.byref(h)) == -3
    assert L.miso_fe_create(ctypes.byref(_vol(None, (4, 4, 4))), ctypes.c_float(0.5), None, None) == -3


def test_calls_out_of_order_and_bad_values():
    from miniisosurface_amd._lib import lib
    L = lib()
    arr = np.random.default_rng(0).uniform(0, 1, (6, 7, 8)).astype(np.float32)
    h = ctypes.c_void_p()
    assert L.miso_fe_create(ctypes.byref(_vol(arr.ctypes.data, (8, 7, 6))), ctypes.c_float(0.5), None,
                            ctypes.byref(h)) == 0
    v, t = ctypes.c_uint64(), ctypes.c_uint64()
    assert L.miso_fe_pass(h, 2) == -4                       # pass 2 before pass 1
    assert L.miso_fe_pass(h, 0) == -3 and L.miso_fe_pass(h, 5) == -3
    assert L.miso_fe_set_index_bits(h, 16) == -3
    out = np.empty((1, 3), np.float32)
    assert L.miso_fe_copy_out(h, out.ctypes.data, out.ctypes.data, None) == -4   # before pass 4
    assert L.miso_fe_pass(h, 1) == 0 and L.miso_fe_pass(h, 2) == 0
    assert L.miso_fe_pass(h, 4) == -4                       # pass 4 before pass 3
    assert L.miso_fe_pass(h, 3) == 0 and L.miso_fe_pass(h, 4) == 0
    assert L.miso_fe_counts(h, ctypes.byref(v), ctypes.byref(t)) == 0 and v.value > 0
    assert L.miso_fe_run(h) == 0                            # still usable
    assert L.miso_fe_set_timing(h, 0x10) == -3 and L.miso_fe_set_timing(None, 1) == -3
    L.miso_fe_destroy(h)
    L.miso_fe_destroy(None)                                 # no-op


def test_pass_timing_selection():
    """miso_fe_set_timing: only the selected passes report a device time, the others 0, and the
    mesh does not depend on it."""
    import miniisosurface_amd as m
    from miniisosurface_amd import synth
    from helpers import bits_equal
    vol = synth.gaussians(48, 3, 1, sigma_range=(4.0, 10.0))
    meshes = []
    for passes in [(1, 2, 3, 4), (4,), (), (1, 3)]:
        a = m.FlyingEdgesAlgorithm(m.Image3D(vol), 0.5)
        a.set_timing(passes)
        a.run()
        ms = a.pass_times_ms()
        for p in range(1, 5):
            assert (ms[p - 1] > 0.0) == (p in passes), (passes, ms)
        meshes.append(a.moveOutput())
    for mesh in meshes[1:]:
        assert bits_equal(mesh.points, meshes[0].points) and bits_equal(mesh.triangles, meshes[0].triangles)
    a = m.FlyingEdgesAlgorithm(m.Image3D(vol), 0.5)
    with pytest.raises(ValueError):
        a.set_timing((0,))


def test_batch_argument_errors():
    import torch
    from miniisosurface_amd._lib import Volume, lib
    L = lib()
    h = ctypes.c_void_p()
    isos = (ctypes.c_float * 1)(0.5)
    assert L.miso_batch_create(None, isos, 1, None, ctypes.byref(h)) == -3
    t = torch.zeros((4, 5, 6), dtype=torch.float32, device="cuda")
    arr = (Volume * 1)(_vol(t.data_ptr(), (6, 5, 4), on_device=1))
    assert L.miso_batch_create(arr, isos, 0, None, ctypes.byref(h)) == -3
    assert L.miso_batch_create(arr, isos, 1, None, ctypes.byref(h)) == 0
    v = ctypes.c_uint64()
    assert L.miso_batch_job_counts(h, 0, ctypes.byref(v), None) == -4    # before run
    assert L.miso_batch_set_tuning(h, 0, 4) == -3
    assert L.miso_batch_run(h) == 0
    assert L.miso_batch_job_counts(h, 1, ctypes.byref(v), None) == -3    # job out of range
    assert L.miso_batch_emit_async(h, 2, None, None, None, 0, 0) == -3
    L.miso_batch_destroy(h)
    e = ctypes.c_void_p()
    assert L.miso_batch_create_empty(None, ctypes.byref(e)) == 0
    assert L.miso_batch_count_async(e) == -4                 # no jobs yet
    L.miso_batch_destroy(e)
