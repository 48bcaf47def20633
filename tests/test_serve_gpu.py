"""GPU: the serving loop (continuous batching over resident volumes, admission by the output
size that the pass-3 counts predict).

Bar: every request's mesh is bit-identical to the same extraction run alone
(FlyingEdgesAlgorithm on that volume and isovalue), whichever step and mix it was served in;
the live output never exceeds the budget; a request that can never fit is rejected and the
others are still served."""
import numpy as np
import pytest

from helpers import bits_equal

pytestmark = pytest.mark.gpu


def _pool(seed=0):
    import torch
    from miniisosurface_amd import DeviceVolume, synth
    vols = [synth.gaussians((48, 40, 36), count=3, seed=seed, sigma_range=(4, 10)),
            synth.gyroid(32, seed + 1),
            synth.gaussians((64, 64, 64), count=5, seed=seed + 2, sigma_range=(5, 12)),
            synth.gyroid((70, 33, 41), seed + 3)]
    tens = [torch.from_numpy(v).cuda() for v in vols]
    ranges = [(float(v.min()), float(v.max())) for v in vols]
    return [DeviceVolume.from_tensor(t) for t in tens], vols, ranges


def _alone(vol, iso):
    import miniisosurface_amd as m
    a = m.FlyingEdgesAlgorithm(m.Image3D(vol), iso)
    a.run()
    return a.moveOutput()


def _same(r, ref):
    return (r.mesh.triangles.shape == ref.triangles.shape and bits_equal(r.mesh.triangles, ref.triangles)
            and bits_equal(r.mesh.points, ref.points) and bits_equal(r.mesh.normals, ref.normals))


def test_serving_results_match_alone():
    from miniisosurface_amd.serve import IsosurfaceServer, latency_stats, poisson_requests
    pool, vols, ranges = _pool()
    reqs = poisson_requests(60, rate=3000.0, pool_ranges=ranges, seed=4)
    srv = IsosurfaceServer(pool, budget_bytes=200e9, max_jobs=16, keep_meshes=True)
    done = srv.run(reqs)
    print(f"[serve] {len(done)} requests in {srv.steps} steps, sizes {srv.step_sizes}", flush=True)
    assert len(done) == 60 and not any(r.rejected for r in done)
    assert srv.steps < 60          # requests were batched together
    for r in done:
        assert _same(r, _alone(vols[r.volume], r.iso)), r.rid
    st = latency_stats(done)
    assert st["p50_ms"] is not None and st["p99_ms"] >= st["p50_ms"]
    srv.close()


def test_serving_admission_under_a_small_budget():
    """A budget of about two requests' output forces deferral; results are unchanged."""
    from miniisosurface_amd.serve import BYTES_PER_TRIANGLE, BYTES_PER_VERTEX, IsosurfaceServer, poisson_requests
    pool, vols, ranges = _pool(seed=7)
    reqs = poisson_requests(24, rate=1e5, pool_ranges=ranges, seed=9)
    refs = {r.rid: _alone(vols[r.volume], r.iso) for r in reqs}
    foot = {rid: m.points.shape[0] * BYTES_PER_VERTEX + m.triangles.shape[0] * BYTES_PER_TRIANGLE
            for rid, m in refs.items()}
    budget = 2 * max(foot.values())
    srv = IsosurfaceServer(pool, budget_bytes=budget, max_jobs=24, keep_meshes=True)
    done = srv.run(reqs)
    assert len(done) == 24 and not any(r.rejected for r in done)
    assert srv.peak_live_bytes <= budget
    assert srv.steps >= 2
    for r in done:
        assert _same(r, refs[r.rid]), r.rid
    srv.close()


def test_serving_rejects_what_never_fits():
    from miniisosurface_amd.serve import IsosurfaceServer, Request
    pool, vols, ranges = _pool(seed=2)
    big = _alone(vols[2], 0.5)
    assert big.triangles.shape[0] > 0
    small_iso = ranges[0][1] - 1e-3          # a sliver near the maximum of volume 0
    small = _alone(vols[0], small_iso)
    budget = small.points.shape[0] * 24 + small.triangles.shape[0] * 12 + 64
    reqs = [Request(volume=2, iso=0.5, arrival=0.0, rid=0), Request(volume=0, iso=small_iso, arrival=0.0, rid=1)]
    srv = IsosurfaceServer(pool, budget_bytes=budget, max_jobs=4, keep_meshes=True)
    done = {r.rid: r for r in srv.run(reqs)}
    assert done[0].rejected and not done[1].rejected
    assert _same(done[1], small)
    srv.close()
