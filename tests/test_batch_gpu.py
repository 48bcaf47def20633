"""GPU: batched Flying Edges (one persistent launch per pass over many jobs, HIP-graph replay).

Bar: every job's points, normals and triangles bit-identical to the single-volume extraction
of the same volume and isovalue, whose parity with the reference the single-volume tests pin
(and, for a few jobs, directly to the oracle); the same on every graph replay and after the
launch shape changes."""
import numpy as np
import pytest

from helpers import bits_equal

pytestmark = pytest.mark.gpu


def _jobs(seed=3):
    import torch
    from miniisosurface_amd import DeviceVolume, synth
    rng = np.random.default_rng(seed)
    vols = []
    for dims in [(32, 32, 32), (45, 77, 33), (256, 40, 37), (129, 65, 2), (2, 3, 5), (200, 31, 64), (64, 64, 64),
                 (97, 130, 41), (33, 250, 38)]:
        vols.append(synth.gaussians(dims, count=int(rng.integers(1, 7)), seed=int(rng.integers(1 << 30)),
                                    sigma_range=(4.0, 12.0)))
    vols.append(synth.gyroid(48, 5))
    tens = [torch.from_numpy(v).cuda() for v in vols]
    jobs = []
    for q, t in enumerate(tens):
        iso = 0.0 if q == len(tens) - 1 else float(rng.uniform(0.2, 0.8))
        jobs.append((DeviceVolume.from_tensor(t), iso, vols[q], (1, 1, 1), (0, 0, 0)))
    # one volume at several isovalues, and one with spacing and origin
    for iso in (-0.6, 0.25, 0.9):
        jobs.append((DeviceVolume.from_tensor(tens[-1]), iso, vols[-1], (1, 1, 1), (0, 0, 0)))
    sp, org = (0.7, 1.3, 0.45), (-3.1, 2.2, 10.5)
    jobs.append((DeviceVolume.from_tensor(tens[1], sp, org), 0.4, vols[1], sp, org))
    return jobs, tens


def _single(vol, iso, sp, org):
    import miniisosurface_amd as m
    a = m.FlyingEdgesAlgorithm(m.Image3D(vol, sp, org), iso)
    a.run()
    return a.moveOutput()


def _check(b, refs):
    for q, ref in enumerate(refs):
        got = b.mesh(q)
        assert got.triangles.shape == ref.triangles.shape, q
        assert bits_equal(got.triangles, ref.triangles), q
        assert bits_equal(got.points, ref.points), q
        assert bits_equal(got.normals, ref.normals), q


def test_batch_matches_single_volume():
    from miniisosurface_amd.batch import BatchExtraction
    jobs, tens = _jobs()
    print(f"[batch] {len(jobs)} jobs", flush=True)
    refs = [_single(v, iso, sp, org) for _, iso, v, sp, org in jobs]
    b = BatchExtraction([(dv, iso) for dv, iso, _, _, _ in jobs])
    for rep in range(3):          # first run (two graphs), then the single-graph replays
        b.run()
        _check(b, refs)
    tv, tt = b.totals()
    assert tv == sum(r.points.shape[0] for r in refs)
    assert tt == sum(r.triangles.shape[0] for r in refs)
    for shape in [(1, 1), (8, 16), (2, 4)]:
        b.set_tuning(*shape)
        for order in ("xcd", "dispatch"):
            b.set_queue_order(order)
            assert b.queue_order() == order
            b.run()
            _check(b, refs)
            b.run()          # the single graph replayed with this order
            _check(b, refs)
    assert b.time_ms() > 0.0
    b.close()


def test_batch_fused_launch_matches():
    """The fused launch (one cooperative kernel for passes 1-4) gives every job the same mesh as
    the graph launch, on each replay, at each launch shape and queue order."""
    from miniisosurface_amd.batch import BatchExtraction
    jobs, tens = _jobs(seed=7)
    print(f"[batch fused] {len(jobs)} jobs", flush=True)
    refs = [_single(v, iso, sp, org) for _, iso, v, sp, org in jobs]
    b = BatchExtraction([(dv, iso) for dv, iso, _, _, _ in jobs])
    b.set_launch("fused")
    assert b.launch() == "fused"
    b.run()                       # first run sizes the output through the graphs
    _check(b, refs)
    for shape in [(4, 4), (1, 1), (2, 16), (8, 1)]:
        b.set_tuning(*shape)
        for order in ("dispatch", "xcd"):
            b.set_queue_order(order)
            print(f"  shape {shape} order {order}", flush=True)
            for rep in range(2):
                b.run()
                _check(b, refs)
    b.set_launch("graphs")
    b.run()
    _check(b, refs)
    b.close()


def test_batch_fused_content_change_regrows_output():
    """A fused replay whose totals outgrow the output writes nothing in its pass 4; the run then
    grows the buffers and runs pass 4 again from the graph."""
    import torch
    from miniisosurface_amd import DeviceVolume
    from miniisosurface_amd.batch import BatchExtraction
    t = torch.zeros((40, 50, 60), dtype=torch.float32, device="cuda")
    t[10:20, 10:20, 10:20] = 1.0
    b = BatchExtraction([(DeviceVolume.from_tensor(t), 0.5)])
    b.set_launch("fused")
    b.run()
    b.run()
    v1, t1 = b.totals()
    t.copy_(torch.from_numpy(np.random.default_rng(1).uniform(0, 1, (40, 50, 60)).astype(np.float32)))
    b.run()
    v2, t2 = b.totals()
    assert v2 > v1 and t2 > t1
    _check(b, [_single(t.cpu().numpy(), 0.5, (1, 1, 1), (0, 0, 0))])
    b.run()
    _check(b, [_single(t.cpu().numpy(), 0.5, (1, 1, 1), (0, 0, 0))])
    b.close()


def test_batch_shared_volume_groups():
    """Consecutive jobs on one volume are classified together, at most 16 to a group (the rows
    loaded once for all of them): 20 isovalues of a volume with 300-voxel rows (16-byte loads and
    a partial step), a job on another volume, 3 more isovalues of the first volume, and 5 of a
    volume with 64-voxel rows (the __ballot path). Each job equals its single-volume run."""
    import torch
    from miniisosurface_amd import DeviceVolume, synth
    from miniisosurface_amd.batch import BatchExtraction
    a = synth.gaussians((300, 21, 9), count=4, seed=5, sigma_range=(3.0, 9.0))
    g = synth.gyroid(40, 2)
    c = synth.gaussians((64, 30, 17), count=3, seed=8, sigma_range=(3.0, 8.0))
    ta, tg, tc = (torch.from_numpy(x).cuda() for x in (a, g, c))
    jobs = [(ta, a, float(iso)) for iso in np.linspace(0.1, 0.9, 20)]
    jobs.append((tg, g, 0.1))
    jobs += [(ta, a, iso) for iso in (0.33, 0.5, 0.71)]
    jobs += [(tc, c, float(iso)) for iso in np.linspace(0.2, 0.8, 5)]
    print(f"[batch groups] {len(jobs)} jobs", flush=True)
    refs = [_single(v, iso, (1, 1, 1), (0, 0, 0)) for _, v, iso in jobs]
    assert sum(r.triangles.shape[0] for r in refs) > 0
    b = BatchExtraction([(DeviceVolume.from_tensor(t), iso) for t, _, iso in jobs])
    for rep in range(2):
        b.run()
        _check(b, refs)
    b.close()


def test_batch_against_oracle(oracle_mod):
    from miniisosurface_amd.batch import BatchExtraction
    jobs, tens = _jobs(seed=11)
    pick = [0, 2, len(jobs) - 1]
    b = BatchExtraction([(jobs[q][0], jobs[q][1]) for q in pick]).run()
    for n, q in enumerate(pick):
        _, iso, vol, sp, org = jobs[q]
        r = oracle_mod.FEResult(vol, iso, sp, org, want_debug=False)
        got = b.mesh(n)
        assert bits_equal(got.points, r.points), q
        assert bits_equal(got.normals, r.normals), q
        assert bits_equal(got.triangles, r.triangles), q


def test_batch_autotune_keeps_results():
    from miniisosurface_amd.batch import BatchExtraction
    jobs, tens = _jobs(seed=5)
    refs = [_single(v, iso, sp, org) for _, iso, v, sp, org in jobs]
    b = BatchExtraction([(dv, iso) for dv, iso, _, _, _ in jobs])
    best = b.autotune(reps=1)
    assert best > 0.0
    assert b.tuning() in [(x, c) for x in (2, 4, 8) for c in (1, 4, 16)]
    assert b.queue_order() in ("dispatch", "xcd")
    assert b.launch() in ("graphs", "fused")
    b.run()
    _check(b, refs)


def test_batch_volume_content_change_regrows_output():
    """Replays from the single graph must notice when the totals outgrow the output."""
    import torch
    from miniisosurface_amd import DeviceVolume
    from miniisosurface_amd.batch import BatchExtraction
    t = torch.zeros((40, 50, 60), dtype=torch.float32, device="cuda")
    t[10:20, 10:20, 10:20] = 1.0
    b = BatchExtraction([(DeviceVolume.from_tensor(t), 0.5)])
    b.run()
    v1, t1 = b.totals()
    b.run()
    t.copy_(torch.from_numpy(np.random.default_rng(0).uniform(0, 1, (40, 50, 60)).astype(np.float32)))
    b.run()
    v2, t2 = b.totals()
    assert v2 > v1 and t2 > t1
    ref = _single(t.cpu().numpy(), 0.5, (1, 1, 1), (0, 0, 0))
    _check(b, [ref])


def test_batch_rejects_host_and_flat_volumes():
    import miniisosurface_amd as m
    import torch
    from miniisosurface_amd import DeviceVolume
    from miniisosurface_amd.batch import BatchExtraction
    with pytest.raises(ValueError):
        BatchExtraction([(m.Image3D(np.zeros((4, 4, 4), np.float32)), 0.5)])
    t = torch.zeros((1, 4, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(m.MisoError):
        BatchExtraction([(DeviceVolume.from_tensor(t), 0.5)])


def test_split_phase_emit_twice_after_one_count(oracle_mod):
    """Split phase: one count, then pass 4 of the first k jobs, then pass 4 of every job into
    other buffers (a step that admits more once the budget frees). Each emit starts its work
    queue afresh, so both outputs hold whole meshes equal to the oracle's."""
    import ctypes

    import torch

    from miniisosurface_amd import _volume_struct
    from miniisosurface_amd._lib import Volume, check, lib
    jobs, tens = _jobs(seed=21)
    jobs = [j for j in jobs if j[3] == (1, 1, 1)][:8]
    L = lib()
    h = ctypes.c_void_p()
    check(L.miso_batch_create_empty(None, ctypes.byref(h)), "create_empty")
    arr = (Volume * len(jobs))()
    for q, j in enumerate(jobs):
        arr[q] = _volume_struct(j[0])[0]
    isos = (ctypes.c_float * len(jobs))(*[j[1] for j in jobs])
    check(L.miso_batch_set_jobs(h, arr, isos, len(jobs)), "set_jobs")
    check(L.miso_batch_count_async(h), "count")
    check(L.miso_batch_count_wait(h), "count_wait")
    refs = [oracle_mod.FEResult(v, iso, sp, org, want_debug=False) for _, iso, v, sp, org in jobs]
    v, t = ctypes.c_uint64(), ctypes.c_uint64()
    for q, r in enumerate(refs):
        check(L.miso_batch_job_counts(h, q, ctypes.byref(v), ctypes.byref(t)), "counts")
        assert (v.value, t.value) == (r.nverts, r.ntris), q
    nv_all = sum(r.nverts for r in refs)
    nt_all = sum(r.ntris for r in refs)

    def emit(k):
        pts = torch.full((3 * max(nv_all, 1),), float("nan"), device="cuda")
        nrm = torch.full_like(pts, float("nan"))
        tri = torch.full((3 * max(nt_all, 1),), -1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        check(L.miso_batch_emit_async(h, k, pts.data_ptr(), nrm.data_ptr(), tri.data_ptr(), max(nv_all, 1),
                                      max(nt_all, 1)), "emit")
        check(L.miso_batch_wait(h), "wait")
        return pts, nrm, tri

    fv, ft = ctypes.c_uint64(), ctypes.c_uint64()
    for k, order in ((3, 0), (len(jobs), 0), (len(jobs), 1), (5, 1)):
        check(L.miso_batch_set_queue_order(h, order), "queue order")
        pts, nrm, tri = emit(k)
        for q in range(k):
            check(L.miso_batch_job_offsets(h, q, ctypes.byref(fv), ctypes.byref(ft)), "offsets")
            v0, t0, r = fv.value, ft.value, refs[q]
            p = pts[3 * v0:3 * (v0 + r.nverts)].view(-1, 3).cpu().numpy()
            n = nrm[3 * v0:3 * (v0 + r.nverts)].view(-1, 3).cpu().numpy()
            tt = tri[3 * t0:3 * (t0 + r.ntris)].view(-1, 3).cpu().numpy().view(np.uint32).astype(np.uint64)
            assert bits_equal(p, r.points) and bits_equal(n, r.normals) and np.array_equal(tt, r.triangles), (k, q)
    L.miso_batch_destroy(h)


def _threads():
    import os
    return min(16, os.cpu_count() or 1)


@pytest.mark.slow
def test_batch_mixed_256_volumes_against_oracle(oracle_mod):
    """The batched workload: 256 Gaussian-blob volumes with dims log-uniform in [32, 256] per
    axis and their own isovalues U[0.2, 0.8] (seed 0), as one batch with the autotuned launch
    shape; every job's mesh bit-exact against the oracle on the same volume and isovalue."""
    import torch

    from miniisosurface_amd import DeviceVolume, synth
    from miniisosurface_amd.batch import BatchExtraction
    vols = synth.batch_volumes(256, seed=0, device="cuda")
    b = BatchExtraction([(DeviceVolume.from_tensor(t), iso) for t, iso in vols])
    b.autotune(reps=1)
    b.run()
    torch.cuda.synchronize()
    print(f"[batch 256] tuning {b.tuning()}, totals {b.totals()}", flush=True)
    for q, (t, iso) in enumerate(vols):
        r = oracle_mod.FEResult(t.cpu().numpy(), iso, nthreads=_threads(), want_debug=False)
        got = b.mesh(q)
        assert got.points.shape[0] == r.nverts and got.triangles.shape[0] == r.ntris, q
        assert bits_equal(got.points, r.points), q
        assert bits_equal(got.normals, r.normals), q
        assert np.array_equal(got.triangles, r.triangles), q
    b.close()


@pytest.mark.slow
def test_batch_mixed_256_fused_matches_graphs():
    """The 256-volume workload (seed 0) with the fused cooperative launch, in both queue orders:
    every job's mesh bit-identical to the graph launch's (which the oracle test above pins)."""
    import torch

    from miniisosurface_amd import DeviceVolume, synth
    from miniisosurface_amd.batch import BatchExtraction
    vols = synth.batch_volumes(256, seed=0, device="cuda")
    b = BatchExtraction([(DeviceVolume.from_tensor(t), iso) for t, iso in vols])
    b.set_launch("graphs")
    b.run()
    refs = [b.mesh(q) for q in range(len(vols))]
    b.set_launch("fused")
    for order in ("dispatch", "xcd"):
        b.set_queue_order(order)
        b.run()
        torch.cuda.synchronize()
        print(f"[batch 256 fused] order {order}, totals {b.totals()}", flush=True)
        _check(b, refs)
    b.close()


@pytest.mark.slow
def test_batch_gyroid_512_sixteen_isovalues_against_oracle(oracle_mod):
    """The 512^3 gyroid (with noise, seed 0) at 16 isovalues evenly spaced in [-1, 1] as one
    batch over the one resident volume; every job's mesh bit-exact against the oracle."""
    import torch

    from miniisosurface_amd import DeviceVolume, synth
    from miniisosurface_amd.batch import BatchExtraction
    g = synth.gyroid(512, seed=0, device="cuda")
    dv = DeviceVolume.from_tensor(g)
    isos = [float(x) for x in np.linspace(-1.0, 1.0, 16)]
    b = BatchExtraction([(dv, i) for i in isos])
    b.autotune(reps=1)
    b.run()
    torch.cuda.synchronize()
    host = g.cpu().numpy()
    print(f"[batch gyroid16] tuning {b.tuning()}, totals {b.totals()}", flush=True)
    for q, iso in enumerate(isos):
        r = oracle_mod.FEResult(host, iso, nthreads=_threads(), want_debug=False)
        got = b.mesh(q)
        print(f"  iso {iso:+.4f}: V={r.nverts} T={r.ntris}", flush=True)
        assert got.points.shape[0] == r.nverts and got.triangles.shape[0] == r.ntris, q
        assert bits_equal(got.points, r.points), q
        assert bits_equal(got.normals, r.normals), q
        assert np.array_equal(got.triangles, r.triangles), q
    b.close()
