"""Benchmark: Flying Edges on one MI355X at 1024^3 float32 (3-octave value noise, iso 0.5).

Contract: `python bench.py [--gpus 1] [--steps K] [--warmup W]` prints ONE JSON line.
A step is one complete extraction (passes 1-4: classify, count, scan, emit, and a check of the
totals against the output buffers) of the volume, already resident in HBM, by miso_fe_run on a
handle that has its output buffers from an earlier step and no pass timing. value = cells /
(wall time per step), in Gcells/s. The roofline object reports the extraction's compulsory HBM
bytes, B_alg = 4N + 24V + 12T (the volume read once, points and normals, 32-bit triangle
indices), over its device time per step, measured with a HIP event pair on the stream the
library runs on (a torch stream handed to the handle) around the timed steps, and the HBM bytes
the four kernels' rocprofv3 counters show. `roofline.kernels` gives each kernel, from steps run
before the timed ones with every pass timed, its own compulsory bytes (pass 1: the volume and
the sign mask; pass 2: mask and row counts; pass 3: row state; pass 4: points, normals,
triangles, the mask and the row state it reads), its time and rate, and its counter traffic.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _reference_fe_seconds(exe_name, vol_np, iso, threads, runs):
    """Median wall time of the reference Flying Edges driver (compiled from its sources into
    oracle/_ref) on vol_np, by its own timer (ctor .. moveOutput, file I/O excluded); falls back
    to the oracle's restatement of the same loops. Returns (seconds, kind)."""
    from oracle import ora
    from miniisosurface_amd import vtkio
    exe = ora.ref_binary(exe_name)
    times = []
    if exe:
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
            vtkio.save_image(os.path.join(tmp, "v.vtk"), vtkio.Image3D(vol_np))
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
            for _ in range(runs):
                out = subprocess.run([exe, "-i", "v.vtk", "-o", "m.vtk", "-v", repr(float(iso)),
                                      "-y", os.path.join(tmp, "y", "r.yaml")], cwd=tmp, env=env,
                                     check=True, capture_output=True, text=True).stdout
                for line in out.splitlines():
                    if "Total Program WALL Time (seconds)" in line:
                        times.append(float(line.split(":")[-1]))
        kind = "reference"
    else:
        for _ in range(runs):
            r = ora.FEResult(vol_np, iso, nthreads=threads, want_debug=False)
            times.append(r.times[5])
        kind = "port"
    return sorted(times)[len(times) // 2], kind


def host_cores():
    """(threads used, physical cores those threads may run on, logical CPUs of the host): the
    threads are OMP_NUM_THREADS (the host CPU share given to one GPU), else the CPUs this
    process may run on; physical cores are counted from `lscpu -p=CPU,CORE,SOCKET` over that
    CPU set."""
    allowed = sorted(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", len(allowed)))
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        core_of = {}
        for line in out.splitlines():
            if line and not line.startswith("#"):
                c, core, sock = line.split(",")[:3]
                core_of[int(c)] = (core, sock)
        phys = len({core_of[c] for c in allowed if c in core_of})
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    return threads, phys, os.cpu_count()


def _reference_mc_seconds(exe_name, vol_np, iso, threads, runs):
    """Median wall time of a reference Marching Cubes driver (oracle/_ref) by its own timer
    (MarchingCubes call only, file I/O excluded), or None when it is not built."""
    from oracle import ora
    from miniisosurface_amd import vtkio
    exe = ora.ref_binary(exe_name)
    if not exe:
        return None
    times = []
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
        vtkio.save_image(os.path.join(tmp, "v.vtk"), vtkio.Image3D(vol_np))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        for _ in range(runs):
            out = subprocess.run([exe, "-i", "v.vtk", "-o", "m.vtk", "-v", repr(float(iso)),
                                  "-y", os.path.join(tmp, "y", "r.yaml")], cwd=tmp, env=env,
                                 check=True, capture_output=True, text=True).stdout
            for line in out.splitlines():
                if "Total Program WALL Time (seconds)" in line:
                    times.append(float(line.split(":")[-1]))
    return sorted(times)[len(times) // 2]


def cpu_baseline(vol_np, iso, threads, runs=5):
    """The reference's OpenMP Flying Edges on a bounded slab of the same volume, all threads."""
    nz, ny, nx = vol_np.shape
    cells = (nx - 1) * (ny - 1) * (nz - 1)
    t, kind = _reference_fe_seconds("flyingEdgesOpenMP", vol_np, iso, threads, runs)
    _, phys, logical = host_cores()
    return {"value": round(cells / t / 1e9, 4), "unit": "Gcells/s", "cores": threads, "kind": kind,
            "physical_cores_available": phys, "host_logical_cpus": logical,
            "sample": f"measured on a {nx}x{ny}x{nz} slab of the same volume (a per-cell rate, not a "
                      f"whole-volume run), iso {iso}, median of {runs} runs, {threads} OpenMP threads, "
                      f"OMP_PROC_BIND=close, OMP_PLACES=cores, timer = reference ctor..moveOutput ({t:.3f} s)"}


def bench_c1(threads, runs=5, reps=20):
    """The reference's own CPU case: 128^3, 4 Gaussians (sigma 8-20, seed 0), iso 0.5. The
    reference serial and OpenMP Flying Edges on the host beside one extraction on the GPU (host
    volume already on the device: launch-latency bound, 8.9 MB), and the GPU call with its
    host-to-device copy and the copy of the mesh back."""
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd import synth
    vol = synth.gaussians(128, 4, 0)
    cells = 127 ** 3
    ts, kind = _reference_fe_seconds("flyingEdgesSerial", vol, 0.5, 1, runs)
    to, _ = _reference_fe_seconds("flyingEdgesOpenMP", vol, 0.5, threads, runs)
    mcs = _reference_mc_seconds("mcSerial", vol, 0.5, 1, runs)
    mco = _reference_mc_seconds("mcOpenMP", vol, 0.5, threads, runs)
    algo = m.FlyingEdgesAlgorithm(m.Image3D(vol), 0.5)
    algo.set_timing(())   # no event pairs between the passes in the timed runs
    for _ in range(3):
        algo.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        algo.run()
    torch.cuda.synchronize()
    gpu = (time.perf_counter() - t0) / reps
    m.flying_edges(m.Image3D(vol), 0.5)
    t0 = time.perf_counter()
    for _ in range(5):
        mesh = m.flying_edges(m.Image3D(vol), 0.5)
    whole = (time.perf_counter() - t0) / 5
    return {"nverts": mesh.numberOfVertices(), "ntris": mesh.numberOfTriangles(), "kind": kind,
            "reference_serial_ms": round(ts * 1e3, 3), "reference_openmp_ms": round(to * 1e3, 3),
            "reference_mc_serial_ms": None if mcs is None else round(mcs * 1e3, 3),
            "reference_mc_openmp_ms": None if mco is None else round(mco * 1e3, 3),
            "openmp_threads": threads, "gpu_ms": round(gpu * 1e3, 4),
            "gpu_with_copies_ms": round(whole * 1e3, 3),
            "gpu_gcells_s": round(cells / gpu / 1e9, 2)}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fe_hbm_traffic_1024_iso0.5.txt")
CALIB_BYTES = 1 << 30     # the plain copy that calibrates the FETCH_SIZE / WRITE_SIZE counters


def _pmc_rows(path):
    import csv
    for dirpath, _, files in os.walk(path):
        for f in files:
            if f.endswith("counter_collection.csv"):
                yield from csv.DictReader(open(os.path.join(dirpath, f)))


def summarize_pmc(fetch_dir, write_dir, out_path, command):
    """Turns two rocprofv3 --pmc runs of `bench.py --calibrate` (FETCH_SIZE, then WRITE_SIZE,
    KiB per dispatch) into per-kernel HBM bytes per launch. Each counter is scaled by the ratio of
    the known bytes of a plain 1 GiB device copy, run in the same process, to what the counter
    reads for that copy."""
    per = {}
    for counter, path in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        calib = 0.0
        sums = {}
        rows = sorted((r for r in _pmc_rows(path) if r["Counter_Name"] == counter), key=lambda r: int(r["Dispatch_Id"]))
        for r in rows:
            name, val = r["Kernel_Name"], float(r["Counter_Value"]) * 1024.0
            if "miso::" not in name and sums == {} and ("copyBuffer" in name or "direct_copy" in name):
                calib = val      # the last device copy before the first extraction: the calibration copy
            if "miso::" in name:
                key = name.split("(")[0].replace("void ", "")
                n, t = sums.get(key, (0, 0.0))
                sums[key] = (n + 1, t + val)
        if calib <= 0:
            raise SystemExit(f"no calibration copy found in {path}")
        factor = CALIB_BYTES / calib
        for key, (n, t) in sums.items():
            per.setdefault(key, {})[counter] = (t / n * factor, t / n, factor)
    lines = [f"# HBM traffic per launch, rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy",
             f"# command: {command}"]
    for key, d in sorted(per.items()):
        fb, fr, ff = d.get("FETCH_SIZE", (0, 0, 0))
        wb, wr, wf = d.get("WRITE_SIZE", (0, 0, 0))
        lines.append(f"{key}: fetch_bytes={fb:.0f} write_bytes={wb:.0f} traffic_bytes={fb + wb:.0f} "
                     f"(raw fetch {fr:.0f} x {ff:.3f}, raw write {wr:.0f} x {wf:.3f})")
    open(out_path, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def _csv_rows(path, suffix):
    import csv
    for dirpath, _, files in os.walk(path):
        for f in files:
            if f.endswith(suffix):
                yield from csv.DictReader(open(os.path.join(dirpath, f)))


def summarize_trace(trace_dir, out_path, command, title):
    """Turns a `rocprofv3 --kernel-trace --output-format csv` run of this benchmark into a short
    text summary: per kernel, calls and average / min / max duration (from the kernel trace)."""
    rows = list(_csv_rows(trace_dir, "kernel_trace.csv"))
    if not rows:
        raise SystemExit(f"no kernel trace in {trace_dir}")
    per = {}
    for r in rows:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        per.setdefault(r["Kernel_Name"], []).append(d)
    lines = [f"# {title}", f"# command: {command}",
             "# per kernel: calls, average / min / max duration in us (kernel trace), share of GPU time"]
    total = sum(sum(v) for v in per.values())
    for name, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        short = name if len(name) <= 90 else name[:87] + "..."
        lines.append(f"{short:<90} calls={len(v):5d} avg_us={sum(v) / len(v):9.1f} min_us={min(v):9.1f} "
                     f"max_us={max(v):9.1f} pct={100 * sum(v) / total:6.2f}")
    open(out_path, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def summarize_counters(pmc_dir, out_path, command, title):
    """Per-kernel averages of a `rocprofv3 --pmc ...` run of this benchmark (miso kernels only),
    with two derived figures where their counters are present: the texture addresser's busy
    fraction (TA_BUSY_avr / GRBM_GUI_ACTIVE) and the L1's average read latency to the L2 in
    cycles (TCP_TCC_READ_REQ_LATENCY_sum / TCP_TCC_READ_REQ_sum)."""
    per = {}
    for r in _csv_rows(pmc_dir, "counter_collection.csv"):
        name = r["Kernel_Name"]
        if "miso::" not in name:
            continue
        key = name.split("(")[0].replace("void ", "")
        d = per.setdefault(key, {})
        d.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    lines = [f"# {title}", f"# command: {command}",
             "# per kernel, the average per dispatch of each counter; derived: ta_busy_frac = TA_BUSY_avr / "
             "GRBM_GUI_ACTIVE, l2_read_latency = TCP_TCC_READ_REQ_LATENCY_sum / TCP_TCC_READ_REQ_sum (cycles)"]
    for key, d in sorted(per.items()):
        avg = {c: sum(v) / len(v) for c, v in d.items()}
        parts = [f"{c}={avg[c]:.0f}" for c in sorted(avg)]
        if "TA_BUSY_avr" in avg and avg.get("GRBM_GUI_ACTIVE"):
            parts.append(f"ta_busy_frac={avg['TA_BUSY_avr'] / avg['GRBM_GUI_ACTIVE']:.3f}")
        if avg.get("TCP_TCC_READ_REQ_sum") and "TCP_TCC_READ_REQ_LATENCY_sum" in avg:
            parts.append(f"l2_read_latency={avg['TCP_TCC_READ_REQ_LATENCY_sum'] / avg['TCP_TCC_READ_REQ_sum']:.0f}")
        lines.append(f"{key}: " + " ".join(parts))
    open(out_path, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary, or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith(kernel + "<") or line.startswith(kernel + ":"):
            for tok in line.split():
                if tok.startswith("traffic_bytes="):
                    return float(tok.split("=")[1])
    return None


def bench_c2(reps=10, cpu_threads=None):
    """512^3 gyroid (8 periods) + U(-0.05, 0.05) noise (seed 0), iso 0: Flying Edges and Marching
    Cubes (first-visit order, and edge order as the duplicate-free variant numbers vertices),
    volume resident in HBM, one complete extraction per step."""
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd import synth

    g = synth.gyroid(512, seed=0, device="cuda")
    dv = m.DeviceVolume.from_tensor(g)
    n = 512
    cells = (n - 1) ** 3
    out = {}
    for name, algo in (("flying_edges", m.FlyingEdgesAlgorithm(dv, 0.0)), ("marching_cubes", m._MC(dv, 0.0)),
                       ("marching_cubes_edge_order", m._MC(dv, 0.0, order="edge"))):
        if name == "flying_edges":
            algo.set_timing(())   # no event pairs between the passes in the timed runs
        for _ in range(3):
            algo.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            algo.run()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        V, T = algo.counts()
        b_alg = 4 * n ** 3 + 24 * V + 12 * T
        out[name] = {"ms_per_step": round(wall * 1e3, 4), "gcells_s": round(cells / wall / 1e9, 2),
                     "hbm_frac_wall": round(b_alg / wall / 1e9 / HBM_PEAK_GBS, 4), "nverts": V, "ntris": T}
        if name == "flying_edges":
            out[name]["pass_ms"] = pass_breakdown(algo, reps)
    if cpu_threads:
        # the reference's OpenMP drivers on the same volume, on the host (one full run each, the
        # median of 3; Marching Cubes with its default sections: one z-section per thread)
        host = g.cpu().numpy()
        tf, kind = _reference_fe_seconds("flyingEdgesOpenMP", host, 0.0, cpu_threads, 3)
        tm = _reference_mc_seconds("mcOpenMP", host, 0.0, cpu_threads, 3)
        out["cpu_baseline"] = {"unit": "Gcells/s", "cores": cpu_threads, "kind": kind,
                               "flying_edges_openmp": round(cells / tf / 1e9, 4),
                               "marching_cubes_openmp": None if tm is None else round(cells / tm / 1e9, 4),
                               "sample": f"the whole 512^3 volume, median of 3 runs each, {cpu_threads} OpenMP "
                                         f"threads, timer = the drivers' own ({tf:.3f} s FE, "
                                         f"{'-' if tm is None else f'{tm:.3f}'} s MC)"}
    return out


def stream_copy_gbs(nbytes=2 << 30, reps=10):
    """Measured stream peak beside the published 8 TB/s: bytes read plus written per second by a
    plain device-to-device copy of 2 GiB (torch), timed with events on the current stream."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (t0.elapsed_time(t1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def pass_breakdown(algo, reps):
    """Average device time of each pass over `reps` runs with all four passes timed (HIP event
    pairs around each kernel), then timing left off; the runs are not part of any timed region."""
    import numpy as np

    algo.set_timing((1, 2, 3, 4))
    ms = np.zeros(4)
    for _ in range(reps):
        algo.run()
        ms += np.array(algo.pass_times_ms())
    algo.set_timing(())
    return [round(float(x), 4) for x in ms / max(reps, 1)]


def bench_slabs(vol, iso, planes=128, reps=3):
    """The benchmark volume extracted in z-slabs of `planes` grid planes (FlyingEdgesSlabs): read
    in place from HBM (the cost of slabbing against the whole-volume run), and streamed from
    pinned host memory one slab ahead of the computation (PCIe-inclusive; parts stay on the
    device), and the same with every part copied to host memory."""
    import numpy as np
    import torch

    import miniisosurface_amd as m

    nz, ny, nx = vol.shape
    cells = (nx - 1) * (ny - 1) * (nz - 1)

    def run(s, copy):
        s.restart(iso)
        n = 0
        while s.next():
            if copy:
                s.copy_part()
            n += 1
        p = s.part()
        return n, p["vbase"] + p["nverts"], p["tbase"] + p["ntris"]

    def timed(image, copy):
        # one handle (device buffers allocated once), restarted for every extraction
        s = m.FlyingEdgesSlabs(image, iso, planes)
        run(s, copy)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = run(s, copy)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, res

    out = {"planes": planes}
    wall, (nslabs, V, T) = timed(m.DeviceVolume.from_tensor(vol), False)
    out.update(slabs=nslabs, nverts=V, ntris=T)
    out["device_volume"] = {"ms_per_step": round(wall * 1e3, 3), "gcells_s": round(cells / wall / 1e9, 2)}
    host = torch.empty(vol.shape, dtype=torch.float32, pin_memory=True)
    host.copy_(vol)
    image = m.Image3D(host.numpy())
    wall, _ = timed(image, False)
    out["pinned_host_volume"] = {"ms_per_step": round(wall * 1e3, 3), "gcells_s": round(cells / wall / 1e9, 2),
                                 "h2d_gb_s": round(4 * vol.numel() / wall / 1e9, 2)}
    wall, _ = timed(image, True)
    out["pinned_host_volume_parts_to_host"] = {"ms_per_step": round(wall * 1e3, 3),
                                               "gcells_s": round(cells / wall / 1e9, 2)}
    del image, host
    return out


def bench_large(reps=5):
    """2048^3 gyroid (16 periods, no noise), iso 0.1: 8.6 G voxels (32 GiB), volume resident in HBM,
    one complete extraction per step (the scale check beyond 2^32 voxels)."""
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd import synth

    n = 2048
    vol = synth.gyroid(n, seed=0, periods=16, noise=0.0, device="cuda")
    algo = m.FlyingEdgesAlgorithm(m.DeviceVolume.from_tensor(vol), 0.1)
    algo.set_timing(())
    algo.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        algo.run()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    V, T = algo.counts()
    b_alg = 4 * n ** 3 + 24 * V + 12 * T
    out = {"ms_per_step": round(wall * 1e3, 3), "gcells_s": round((n - 1) ** 3 / wall / 1e9, 2),
           "hbm_frac_wall": round(b_alg / wall / 1e9 / HBM_PEAK_GBS, 4), "nverts": V, "ntris": T,
           "pass_ms": pass_breakdown(algo, 2)}
    del algo, vol
    torch.cuda.empty_cache()
    return out


def bench_host_scope(vol, iso, reps=2):
    """The reference driver's timed scope (ctor .. moveOutput, flyingEdges/serial/main.cpp:110-158)
    from a host volume: the host-to-device copy of the volume, the four passes and the mesh copied
    back into host arrays with 64-bit indices (the reference's size_t). PCIe-inclusive; not `value`."""
    import torch

    import miniisosurface_amd as m
    host = vol.cpu().numpy()
    nz, ny, nx = host.shape
    cells = (nx - 1) * (ny - 1) * (nz - 1)
    image = m.Image3D(host)
    m.flying_edges(image, iso)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        mesh = m.flying_edges(image, iso)
    wall = (time.perf_counter() - t0) / reps
    return {"ms_per_step": round(wall * 1e3, 2), "gcells_s": round(cells / wall / 1e9, 3),
            "nverts": mesh.numberOfVertices(), "ntris": mesh.numberOfTriangles(),
            "volume_bytes": int(host.nbytes), "mesh_bytes": int(mesh.points.nbytes * 2 + mesh.triangles.nbytes)}


def bench_batched(reps, cpu_threads=None):
    """256 mixed-size Gaussian volumes (dims log-uniform in [32, 256], 1-6 blobs, iso U[0.2, 0.8],
    seed 0), and one 512^3 gyroid at 16 isovalues in [-1, 1]: the batch (one persistent launch per
    pass, graph replay, autotuned shape) against the same jobs run one extraction at a time."""
    import numpy as np
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd import synth
    from miniisosurface_amd.batch import BatchExtraction

    def compare(plain_meshes, b):
        for q, ref in enumerate(plain_meshes):
            got = b.mesh(q)
            if not (np.array_equal(got.triangles, ref.triangles)
                    and np.array_equal(got.points.view(np.uint32), ref.points.view(np.uint32))
                    and np.array_equal(got.normals.view(np.uint32), ref.normals.view(np.uint32))):
                return False
        return True

    def run_set(name, jobs, algos, set_iso=None):
        # one extraction at a time: each runs its four passes with its own launches and sync
        # (no pass timing, so no event pairs between its kernels)
        for a in algos:
            a.set_timing(())
        for a in algos[:2]:
            a.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for q, a in enumerate(algos):
                if set_iso is not None:
                    a.set_isoval(set_iso[q])
                a.run()
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t0) / reps
        plain_meshes = []
        for q, a in enumerate(algos):
            if set_iso is not None:
                a.set_isoval(set_iso[q])
                a.run()
            plain_meshes.append(a.moveOutput())
        b = BatchExtraction(jobs)
        tune_ms = b.autotune(reps=2)
        b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        torch.cuda.synchronize()
        batch_s = (time.perf_counter() - t0) / reps
        V, T = b.totals()
        same = compare(plain_meshes, b)
        def wall_ms():
            b.run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                b.run()
            torch.cuda.synchronize()
            return round((time.perf_counter() - t0) / reps * 1e3, 3)

        # at the tuned shape: pass 4's queue order (one queue, or one per XCD), then the launch
        # (six kernels from one graph, or one cooperative kernel)
        chosen, by_order = b.queue_order(), {}
        for order in ("dispatch", "xcd"):
            b.set_queue_order(order)
            by_order[order] = wall_ms()
        b.set_queue_order(chosen)
        launch, by_launch = b.launch(), {}
        for mode in ("graphs", "fused"):
            b.set_launch(mode)
            by_launch[mode] = wall_ms()
        b.set_launch(launch)
        out = {"jobs": len(jobs), "plain_ms": round(plain_s * 1e3, 3), "batched_ms": round(batch_s * 1e3, 3),
               "batched_device_ms": round(b.time_ms(), 3), "speedup": round(plain_s / batch_s, 2),
               "plain_jobs_per_s": round(len(jobs) / plain_s, 1), "batched_jobs_per_s": round(len(jobs) / batch_s, 1),
               "nverts": V, "ntris": T, "tuning": list(b.tuning()), "queue_order": chosen,
               "ms_by_queue_order": by_order, "launch": launch, "ms_by_launch": by_launch,
               "autotune_best_ms": round(tune_ms, 3),
               "identical_output": same}
        b.close()
        return out

    vols = synth.batch_volumes(256, seed=0, device="cuda")
    dvs = [m.DeviceVolume.from_tensor(t) for t, _ in vols]
    jobs = [(dv, iso) for dv, (_, iso) in zip(dvs, vols)]
    algos = [m.FlyingEdgesAlgorithm(dv, iso) for dv, iso in jobs]
    voxels = sum(int(t.numel()) for t, _ in vols)
    mixed = run_set("mixed", jobs, algos)
    mixed["voxels"] = voxels
    del algos
    if cpu_threads:
        # the same volumes on the host: the reference's OpenMP Flying Edges, one volume after
        # another on all threads, by its own timer, on a bounded sample (every 8th volume)
        sample = vols[::8]
        secs, kind = 0.0, None
        for t, iso in sample:
            sec, kind = _reference_fe_seconds("flyingEdgesOpenMP", t.cpu().numpy(), iso, cpu_threads, 1)
            secs += sec
        mixed["cpu_baseline"] = {"value": round(len(sample) / secs, 1), "unit": "volumes/s", "cores": cpu_threads,
                                 "kind": kind, "sample": f"every 8th of the 256 volumes ({len(sample)}), one run "
                                 f"each, OMP_PROC_BIND=close, OMP_PLACES=cores ({secs:.3f} s)"}
    g = synth.gyroid(512, seed=0, device="cuda")
    dv = m.DeviceVolume.from_tensor(g)
    isos = [float(x) for x in np.linspace(-1.0, 1.0, 16)]
    one = m.FlyingEdgesAlgorithm(dv, isos[0])
    sweep = run_set("sweep", [(dv, i) for i in isos], [one] * 16, set_iso=isos)
    return {"mixed_256": mixed, "gyroid512_16_isovalues": sweep}


def bench_serving(n_requests, rate=200.0, seed=0, saturated_rate=20000.0):
    """Poisson arrivals (rate 200/s, seeded) over a resident pool of 16 volumes with cubic dims
    drawn from {64, 128, 256, 384, 512} (Gaussian blobs or gyroid), isovalue uniform over each
    field's [min, max], 200 GB admission budget. Latency runs from arrival to delivery: with
    host delivery (the headline figures) until the request's mesh has landed in pinned host
    memory; with device hand-over (`device_handoff`) until pass 4 has written it, the requester
    then holding its slice of the device output (released at once here)."""
    import numpy as np
    import torch

    from miniisosurface_amd import DeviceVolume, synth
    from miniisosurface_amd.serve import IsosurfaceServer, Request, latency_stats, poisson_requests

    pool = synth.serving_pool(16, seed=seed, device="cuda")
    dvs = [DeviceVolume.from_tensor(t) for t, _, _ in pool]
    ranges = [r for _, r, _ in pool]
    dims = [n for _, _, n in pool]
    torch.cuda.synchronize()

    def play(deliver, reqs, on_done=None):
        srv = IsosurfaceServer(dvs, budget_bytes=200e9, max_jobs=64, deliver=deliver)
        # warm-up: kernels, slot buffers, and pinned host blocks of every size class the stream
        # needs (one request per pool volume near the middle of its range, then a short burst)
        srv.run([Request(volume=q, iso=float(np.float32((lo + hi) / 2)), arrival=0.0, rid=q)
                 for q, (lo, hi) in enumerate(ranges)], on_done=on_done)
        srv.run(poisson_requests(8, 1e4, ranges, seed=seed + 2), on_done=on_done)
        srv.steps, srv.step_sizes, srv.peak_live_bytes = 0, [], 0
        done = srv.run(reqs, on_done=on_done)
        st = latency_stats(done)
        st.update({"rate_per_s": round(len(reqs) / max(reqs[-1].arrival, 1e-9), 1), "steps": srv.steps,
                   "mean_batch": round(float(np.mean(srv.step_sizes)), 2) if srv.step_sizes else 0,
                   "max_batch": int(max(srv.step_sizes)) if srv.step_sizes else 0,
                   "peak_live_gb": round(srv.peak_live_bytes / 1e9, 3), "deliver": deliver})
        delivered = sum(r.nverts * 24 + r.ntris * 12 for r in done if not r.rejected)
        st["delivered_gb"] = round(delivered / 1e9, 3)
        for r in done:
            r.result = None
        srv.close()
        return st

    def consume(r):
        # the requester takes its mesh and lets go of it (the pinned blocks are then reused)
        r.result = None

    st = play("host", poisson_requests(n_requests, rate, ranges, seed=seed + 1), on_done=consume)
    st.update({"pool_dims": dims, "budget_gb": 200, "offered_rate_per_s": rate})
    st["device_handoff"] = play("device", poisson_requests(n_requests, rate, ranges, seed=seed + 1),
                                on_done=lambda r: r.release())
    # the same pool and request mix offered faster than one-at-a-time service keeps up with:
    # arrivals queue while the device works and are served in batches
    st["saturated"] = play("host", poisson_requests(2 * n_requests, saturated_rate, ranges, seed=seed + 3),
                           on_done=consume)
    st["saturated"]["offered_rate_per_s"] = saturated_rate
    st["saturated_device_handoff"] = play("device", poisson_requests(2 * n_requests, saturated_rate, ranges,
                                                                     seed=seed + 3), on_done=lambda r: r.release())
    st["saturated_device_handoff"]["offered_rate_per_s"] = saturated_rate
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--iso", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sparse", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-slabs", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-host-scope", action="store_true")
    ap.add_argument("--c2-only", action="store_true",
                    help="only the 512^3 gyroid section (Flying Edges and Marching Cubes), for profiling")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--calibrate", action="store_true",
                    help="run one plain 1 GiB device copy first (calibrates rocprofv3 byte counters)")
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT"),
                    help="write the per-kernel HBM traffic summary of two rocprofv3 --pmc runs and exit")
    ap.add_argument("--pmc-command", default="", help="the profiled command, recorded in the summary")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("TRACE_DIR", "OUT"),
                    help="write the per-kernel summary of a rocprofv3 --kernel-trace run and exit")
    ap.add_argument("--trace-title", default="", help="first line of the --summarize-trace/-counters summary")
    ap.add_argument("--summarize-counters", nargs=2, metavar=("PMC_DIR", "OUT"),
                    help="write the per-kernel counter averages of a rocprofv3 --pmc run and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, command=args.pmc_command)
        return
    if args.summarize_trace:
        summarize_trace(*args.summarize_trace, command=args.pmc_command, title=args.trace_title)
        return
    if args.summarize_counters:
        summarize_counters(*args.summarize_counters, command=args.pmc_command, title=args.trace_title)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (volumes would shard by z-slabs)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd import synth

    if args.c2_only:
        print(json.dumps({"c2_gyroid_512_iso0": bench_c2(reps=max(args.steps, 1))}))
        return
    n = args.n
    torch.cuda.init()
    # the stream the library's kernels run on, and the events that time them (default priority)
    stream = torch.cuda.Stream()
    t0 = time.time()
    vol = synth.value_noise(n, seed=0, device="cuda")
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    cells = (n - 1) ** 3
    N = n ** 3

    algo = m.FlyingEdgesAlgorithm(m.DeviceVolume.from_tensor(vol), args.iso, stream=stream)
    if args.calibrate:
        src = torch.empty(CALIB_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def run_steps(k):
        ms = np.zeros(4)
        for _ in range(k):
            algo.run()
            ms += np.array(algo.pass_times_ms())
        return ms / max(k, 1)

    # the per-pass breakdown, every pass timed with HIP events (an event pair around each kernel)
    algo.set_timing((1, 2, 3, 4))
    run_steps(args.warmup)
    pass_ms = run_steps(args.steps)
    # the timed region: no event pairs between the kernels, timed on the host and with a HIP
    # event pair around the steps on the stream the library runs on
    algo.set_timing(())
    run_steps(1)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        algo.run()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t1) / args.steps
    dev_ms = ev0.elapsed_time(ev1) / args.steps
    V, T = algo.counts()
    _, _, _, bits, _, _ = algo.device_output()

    # compulsory HBM bytes of one extraction and of each pass (uint32 or uint64 triangle ids)
    tri_b = (bits // 8) * 3 * T
    b_alg = 4 * N + 24 * V + tri_b
    R = n * n
    W = (n + 63) // 64
    mask_b = 8 * W * R
    per_kernel = {
        # the volume once, the sign mask written
        "miso::k_fe_classify": {"ms": pass_ms[0], "bytes": 4 * N + mask_b},
        # the mask read, counts and trims written (16 + 8 bytes per grid row)
        "miso::k_fe_count_rows": {"ms": pass_ms[1], "bytes": mask_b + 24 * R},
        # counts read, vertex and triangle offsets written
        "miso::k_fe_scan": {"ms": pass_ms[2], "bytes": 16 * R + 16 * R},
        # points, normals, triangles written; the mask, counts and offsets read. Its reads of
        # the volume near the surface are not compulsory: pass 1 has read it once already
        "miso::k_fe_emit_rows": {"ms": pass_ms[3], "bytes": 24 * V + tri_b + mask_b + 32 * R},
    }
    kern_total = float(pass_ms.sum())
    traffic_total = 0.0
    kernels = []
    for name, d in per_kernel.items():
        tr = read_traffic(name) if (n == 1024 and args.iso == 0.5) else None
        if tr is None:
            traffic_total = None
        elif traffic_total is not None:
            traffic_total += tr
        kernels.append({"kernel": name, "ms": round(float(d["ms"]), 4), "alg_bytes": int(d["bytes"]),
                        "achieved_gb_s": round(d["bytes"] / (d["ms"] * 1e-3) / 1e9, 1),
                        "frac": round(d["bytes"] / (d["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "traffic_bytes": None if tr is None else int(tr)})
    achieved = b_alg / (dev_ms * 1e-3) / 1e9
    copy_gbs = stream_copy_gbs()

    # the sparse isovalue of the same config
    sparse = None
    if not args.no_sparse:
        algo.set_isoval(0.95)
        sp_ms = np.array(pass_breakdown(algo, 5))
        algo.set_timing(())
        run_steps(1)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        run_steps(5)
        torch.cuda.synchronize()
        sparse_wall = (time.perf_counter() - t2) / 5
        sV, sT = algo.counts()
        sparse = {"ms_per_step": round(sparse_wall * 1e3, 4), "nverts": sV, "ntris": sT,
                  "kernel_ms": round(float(sp_ms.sum()), 4), "gcells_s": round(cells / sparse_wall / 1e9, 3)}

    out = {
        "metric": "FE Gcells/s & %HBM peak at 1024^3",
        "value": round(cells / wall / 1e9, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"flying edges, {n}^3 float32 3-octave value noise (seed 0), iso {args.iso}",
                   "nverts": V, "ntris": T, "index_bits": bits},
        "roofline": {"bound": "hbm",
                     "kernel": "the whole extraction: k_fe_classify, k_fe_count_rows, k_fe_scan, k_fe_emit_rows",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic_total is None else int(traffic_total),
                     "alg_bytes": int(b_alg), "extraction_ms_events": round(dev_ms, 4),
                     "traffic_over_alg": None if traffic_total is None else round(traffic_total / b_alg, 3),
                     "kernels": kernels},
        "extraction": {"bytes_alg": b_alg, "device_ms_events": round(dev_ms, 4),
                       "kernel_ms": round(kern_total, 4),
                       "hbm_frac_kernels": round(b_alg / (kern_total * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "hbm_frac_wall": round(b_alg / wall / 1e9 / HBM_PEAK_GBS, 4),
                       "pass_ms": [round(float(x), 4) for x in pass_ms],
                       "stream_copy_gb_s": round(copy_gbs, 1),
                       "frac_of_stream_copy_wall": round(b_alg / wall / 1e9 / copy_gbs, 4)},
        "sparse_iso_0.95": sparse,
        "volume_gen_s": round(gen_s, 2),
    }
    if not args.no_host_scope:
        out["reference_scope_host_volume"] = bench_host_scope(vol, args.iso)
    if not args.no_slabs:
        out["slabs_128_planes"] = bench_slabs(vol, args.iso)
    if not args.no_large:
        out["large_2048_gyroid_iso0.1"] = bench_large()
    threads = host_cores()[0]
    if not args.no_c2:
        out["c2_gyroid_512_iso0"] = bench_c2(cpu_threads=None if args.no_cpu else threads)
    if not args.no_batch:
        out["batched"] = bench_batched(reps=3, cpu_threads=None if args.no_cpu else threads)
    if not args.no_serve:
        out["serving"] = bench_serving(args.serve_requests)
    if not args.no_cpu:
        slab = vol[: max(2, n // 4)].cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(slab, args.iso, threads)
        out["cpu_baseline_1_thread"] = cpu_baseline(slab[: max(2, n // 16)], args.iso, 1)
        out["c1_gaussians_128_iso0.5"] = bench_c1(threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
