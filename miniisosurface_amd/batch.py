"""Batched Flying Edges: many device volumes (or one volume at many isovalues) per launch.

    b = BatchExtraction([(DeviceVolume, iso), ...])
    b.run()                      # passes 1-4 for every job, replayed from captured HIP graphs
    b.autotune()                 # picks the persistent kernels' launch shape for this batch
    mesh = b.mesh(j)             # job j as vtkio.TriangleMesh (uint64 indices, as the reference)
    pts, nrm, tris = b.device_output(j)   # device pointers into the shared output

Each job's result is bit-identical to FlyingEdgesAlgorithm on the same volume and isovalue. The
reference has no batched path; the per-volume algorithm it replaces is
flyingEdges/serial/FlyingEdgesAlgorithm.h:27-49.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import Volume, check, lib
from .vtkio import TriangleMesh


class BatchExtraction:
    def __init__(self, jobs, stream=None):
        from . import DeviceVolume, _stream_ptr, _volume_struct
        jobs = list(jobs)
        if not jobs:
            raise ValueError("empty batch")
        arr = (Volume * len(jobs))()
        self._keep = []
        for q, (vol, _) in enumerate(jobs):
            if not isinstance(vol, DeviceVolume):
                raise ValueError("batch volumes must be DeviceVolume (device-resident)")
            v, keep = _volume_struct(vol)
            arr[q] = v
            self._keep.append(keep)
        isos = (ctypes.c_float * len(jobs))(*[float(i) for i in (iso for _, iso in jobs)])
        self.njobs = len(jobs)
        self.dims = [tuple(int(d) for d in vol.dims) for vol, _ in jobs]
        h = ctypes.c_void_p()
        check(lib().miso_batch_create(arr, isos, self.njobs, _stream_ptr(stream, True), ctypes.byref(h)),
              "BatchExtraction")
        self._h = h
        self._destroy = lib().miso_batch_destroy

    def run(self):
        check(lib().miso_batch_run(self._h), "batch run")
        return self

    def totals(self):
        v, t = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().miso_batch_totals(self._h, ctypes.byref(v), ctypes.byref(t)), "totals")
        return v.value, t.value

    def counts(self, job):
        v, t = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().miso_batch_job_counts(self._h, job, ctypes.byref(v), ctypes.byref(t)), "job counts")
        return v.value, t.value

    def mesh(self, job) -> TriangleMesh:
        v, t = self.counts(job)
        pts = np.empty((v, 3), np.float32)
        nrm = np.empty((v, 3), np.float32)
        tris = np.empty((t, 3), np.uint64)
        check(lib().miso_batch_copy_out(self._h, job, pts.ctypes.data, nrm.ctypes.data, tris.ctypes.data), "copy_out")
        return TriangleMesh(pts, nrm, tris)

    def device_output(self, job):
        p, n, t = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().miso_batch_job_output(self._h, job, ctypes.byref(p), ctypes.byref(n), ctypes.byref(t)), "output")
        return p.value, n.value, t.value

    def set_tuning(self, blocks_per_cu, chunk):
        check(lib().miso_batch_set_tuning(self._h, int(blocks_per_cu), int(chunk)), "set_tuning")

    def tuning(self):
        b, c = ctypes.c_int(), ctypes.c_int()
        check(lib().miso_batch_get_tuning(self._h, ctypes.byref(b), ctypes.byref(c)), "get_tuning")
        return b.value, c.value

    def set_queue_order(self, order):
        """Pass 4's work queue: "dispatch" (one queue) or "xcd" (one per XCD, tiled rows)."""
        check(lib().miso_batch_set_queue_order(self._h, {"dispatch": 0, "xcd": 1}[order]), "set_queue_order")

    def queue_order(self):
        o = ctypes.c_int()
        check(lib().miso_batch_get_queue_order(self._h, ctypes.byref(o)), "get_queue_order")
        return ("dispatch", "xcd")[o.value]

    def set_launch(self, mode):
        """How run() launches once the output size is known: "graphs" (six kernels replayed from
        one HIP graph) or "fused" (one cooperative kernel, a grid barrier between the passes)."""
        check(lib().miso_batch_set_launch(self._h, {"graphs": 0, "fused": 1}[mode]), "set_launch")

    def launch(self):
        o = ctypes.c_int()
        check(lib().miso_batch_get_launch(self._h, ctypes.byref(o)), "get_launch")
        return ("graphs", "fused")[o.value]

    def autotune(self, reps=3):
        ms = ctypes.c_float()
        check(lib().miso_batch_autotune(self._h, int(reps), ctypes.byref(ms)), "autotune")
        return ms.value

    def time_ms(self):
        ms = ctypes.c_float()
        check(lib().miso_batch_time_ms(self._h, ctypes.byref(ms)), "time")
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
