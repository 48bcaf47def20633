"""Benchmark: Flying Edges on one MI355X at 1024^3 float32 (3-octave value noise, iso 0.5).

Contract: `python bench.py [--gpus 1] [--steps K] [--warmup W]` prints ONE JSON line.
A step is one complete extraction (passes 1-4: classify, count, scan, emit, with the host
synchronisation after the scan that sizes the output) of the volume, already resident in HBM.
value = cells / (wall time per step), in Gcells/s. The roofline object reports the dominant
kernel's algorithmic bytes over its average duration, measured with HIP events recorded by the
library on the stream the kernels run on.
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


def cpu_baseline(vol_np, iso, threads):
    """The reference's OpenMP Flying Edges (compiled from its sources into oracle/_ref) on a
    bounded slab of the same volume; falls back to the oracle's OpenMP restatement."""
    from oracle import ora
    from miniisosurface_amd import vtkio
    nz, ny, nx = vol_np.shape
    cells = (nx - 1) * (ny - 1) * (nz - 1)
    exe = ora.ref_binary("flyingEdgesOpenMP")
    times = []
    if exe:
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
            vtkio.save_image(os.path.join(tmp, "v.vtk"), vtkio.Image3D(vol_np))
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
            for _ in range(3):
                out = subprocess.run([exe, "-i", "v.vtk", "-o", "m.vtk", "-v", repr(float(iso)),
                                      "-y", os.path.join(tmp, "y", "r.yaml")], cwd=tmp, env=env,
                                     check=True, capture_output=True, text=True).stdout
                for line in out.splitlines():
                    if "Total Program WALL Time (seconds)" in line:
                        times.append(float(line.split(":")[-1]))
        kind = "reference"
    else:
        for _ in range(3):
            r = ora.FEResult(vol_np, iso, nthreads=threads, want_debug=False)
            times.append(r.times[5])
        kind = "port"
    t = sorted(times)[len(times) // 2]
    return {"value": round(cells / t / 1e9, 4), "unit": "Gcells/s", "cores": threads, "kind": kind,
            "sample": f"{nx}x{ny}x{nz} slab of the same volume, iso {iso}, median of 3 runs, "
                      f"timer = reference ctor..moveOutput ({t:.3f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--iso", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sparse", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import miniisosurface_amd as m
    from miniisosurface_amd import synth

    n = args.n
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    t0 = time.time()
    vol = synth.value_noise(n, seed=0, device="cuda")
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    cells = (n - 1) ** 3
    N = n ** 3

    algo = m.FlyingEdgesAlgorithm(m.DeviceVolume.from_tensor(vol), args.iso, stream=stream)

    def run_steps(k):
        ms = np.zeros(4)
        for _ in range(k):
            algo.run()
            ms += np.array(algo.pass_times_ms())
        return ms / max(k, 1)

    run_steps(args.warmup)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pass_ms = run_steps(args.steps)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t1) / args.steps
    V, T = algo.counts()
    _, _, _, bits, _, _ = algo.device_output()

    # algorithmic bytes (compulsory traffic) of one extraction and of each pass
    b_alg = 4 * N + 24 * V + (bits // 8) * 3 * T
    R = n * n
    W = (n + 63) // 64
    mask_b = 8 * W * R
    per_kernel = {
        "fe_classify": {"ms": pass_ms[0], "bytes": 4 * N + mask_b},
        "fe_count": {"ms": pass_ms[1], "bytes": mask_b + 24 * R},
        "fe_scan": {"ms": pass_ms[2], "bytes": 16 * R + 16 * R},
        "fe_emit": {"ms": pass_ms[3], "bytes": 24 * V + (bits // 8) * 3 * T + mask_b},
    }
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms"])
    dk = per_kernel[dom]
    achieved = dk["bytes"] / (dk["ms"] * 1e-3) / 1e9
    kern_total = float(pass_ms.sum())

    # the sparse isovalue of the same config
    sparse = None
    if not args.no_sparse:
        algo.set_isoval(0.95)
        run_steps(1)
        t2 = time.perf_counter()
        sp_ms = run_steps(5)
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
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "extraction": {"bytes_alg": b_alg, "kernel_ms": round(kern_total, 4),
                       "hbm_frac_kernels": round(b_alg / (kern_total * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "hbm_frac_wall": round(b_alg / wall / 1e9 / HBM_PEAK_GBS, 4),
                       "pass_ms": [round(float(x), 4) for x in pass_ms]},
        "sparse_iso_0.95": sparse,
        "volume_gen_s": round(gen_s, 2),
    }
    if not args.no_cpu:
        slab = vol[: max(2, n // 4)].cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(slab, args.iso, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
