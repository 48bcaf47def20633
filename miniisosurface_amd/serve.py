"""Serving loop: continuous batching of isosurface requests over volumes resident in HBM.

    server = IsosurfaceServer(pool, budget_bytes=200e9, deliver="host")
    done = server.run(requests)          # requests: Request(volume index, isovalue, arrival s)
    stats = latency_stats(done)

Each slot is a split-phase batch (miso_batch_* in include/miniiso.h) on its own HIP stream. The
loop admits requests as they arrive: an idle slot takes every waiting request (up to max_jobs)
and enqueues passes 1-3 for them; when those counts land, the jobs are admitted in arrival order
while the memory their output needs (24 bytes per vertex, 12 per triangle) fits the budget left
by the outputs still alive, pass 4 of that prefix is enqueued into one freshly allocated output
arena, and the rest go back to the head of the queue. With two slots, the counts of newly
arrived requests run on the device while the previous step's triangles are being written.

A request is done when its mesh is delivered, and its latency runs from arrival to delivery:
  deliver="host"    the step's output is copied into pinned host memory on the slot's stream
                    right behind pass 4; the request is done when that copy has landed, and
                    holds its mesh as host arrays (Request.result, Request.mesh); the device
                    arena is released then;
  deliver="device"  the request is done when pass 4 has finished, and holds device views of its
                    slice of the arena (Request.result); the arena's bytes stay counted against
                    the budget until every request of the step has called Request.release().

The reference has no serving path; each request's result is bit-identical to
FlyingEdgesAlgorithm on its volume and isovalue (flyingEdges/serial/FlyingEdgesAlgorithm.h:27-49).
"""
from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import MisoError, Volume, check, lib
from .vtkio import TriangleMesh

BYTES_PER_VERTEX = 24      # float32 point + float32 normal
BYTES_PER_TRIANGLE = 12    # three uint32 indices


@dataclass
class Request:
    volume: int                  # index into the server's pool
    iso: float
    arrival: float = 0.0         # seconds after the start of run()
    rid: int = 0
    admitted: float | None = None
    done: float | None = None
    nverts: int = 0
    ntris: int = 0
    rejected: bool = False
    step: int = -1
    result: tuple | None = None      # (points (V, 3), normals (V, 3), triangles (T, 3) uint32)
    _arena: object = field(default=None, repr=False)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival

    @property
    def mesh(self) -> TriangleMesh | None:
        """The delivered mesh with the reference's 64-bit triangle indices (host delivery)."""
        if self.result is None:
            return None
        p, n, t = self.result
        if not isinstance(p, np.ndarray):
            p, n, t = p.cpu().numpy(), n.cpu().numpy(), t.cpu().numpy()
        return TriangleMesh(p, n, t.view(np.uint32).astype(np.uint64))

    def release(self):
        """Gives back a device-delivered result: the step's arena is freed once all of its
        requests have released theirs."""
        a, self._arena = self._arena, None
        self.result = None
        if a is not None:
            a.release()


class _Arena:
    """Device output of one step: points, normals, triangles of its admitted jobs."""

    def __init__(self, nv, nt, device, server):
        import torch
        self.pts = torch.empty(max(3 * nv, 1), dtype=torch.float32, device=device)
        self.nrm = torch.empty(max(3 * nv, 1), dtype=torch.float32, device=device)
        self.tris = torch.empty(max(3 * nt, 1), dtype=torch.int32, device=device)
        self.bytes = nv * BYTES_PER_VERTEX + nt * BYTES_PER_TRIANGLE
        self.nv, self.nt = nv, nt
        self.server = server
        self.holders = 0
        self.host = None

    def release(self):
        """One holder less; the last one returns the arena's bytes to the budget."""
        self.holders -= 1
        if self.holders <= 0 and self.server is not None:
            self.server.live_bytes -= self.bytes
            self.server._released = True
            self.pts = self.nrm = self.tris = None
            self.server = None


class _Slot:
    def __init__(self):
        import torch
        # a stream of the default priority, shared by the batch's kernels and the delivery copies
        self.stream = torch.cuda.Stream()
        h = ctypes.c_void_p()
        check(lib().miso_batch_create_empty(ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)),
              "serving slot")
        self.h = h
        self.state = "idle"
        self.reqs = []
        self.sizes = []
        self.arena = None
        self.step = -1

    def query(self):
        rc = lib().miso_batch_query(self.h)
        if rc < 0:
            check(rc, "query")
        return rc == 1

    def close(self):
        if self.h:
            lib().miso_batch_destroy(self.h)
            self.h = None


class IsosurfaceServer:
    def __init__(self, pool, budget_bytes=200e9, max_jobs=64, slots=2, deliver="host", device="cuda"):
        """pool: list of DeviceVolume. deliver: "host" or "device" (see the module text)."""
        from . import DeviceVolume
        if deliver not in ("host", "device"):
            raise ValueError('deliver is "host" or "device"')
        self.deliver = deliver
        self.pool = list(pool)
        for v in self.pool:
            if not isinstance(v, DeviceVolume):
                raise ValueError("the pool holds DeviceVolume entries")
        # the slots run on streams of their own: the resident volumes must be complete first
        import torch
        torch.cuda.synchronize()
        self.budget = int(budget_bytes)
        self.max_jobs = int(max_jobs)
        self.device = device
        self._released = False
        self.slots = [_Slot() for _ in range(slots)]
        self.live_bytes = 0
        self.peak_live_bytes = 0
        self.steps = 0
        self.step_sizes = []

    def close(self):
        for s in self.slots:
            s.close()

    # ---------------------------------------------------------------- slot transitions
    def _start_count(self, slot, reqs):
        from . import _volume_struct
        n = len(reqs)
        arr = (Volume * n)()
        for q, r in enumerate(reqs):
            arr[q] = _volume_struct(self.pool[r.volume])[0]
        isos = (ctypes.c_float * n)(*[r.iso for r in reqs])
        check(lib().miso_batch_set_jobs(slot.h, arr, isos, n), "set_jobs")
        check(lib().miso_batch_count_async(slot.h), "count")
        slot.reqs = reqs
        slot.state = "counting"

    def _counted(self, slot):
        """Counts are in: the output size of each of the slot's jobs."""
        check(lib().miso_batch_count_wait(slot.h), "count_wait")
        v, t = ctypes.c_uint64(), ctypes.c_uint64()
        slot.sizes = []
        for q in range(len(slot.reqs)):
            check(lib().miso_batch_job_counts(slot.h, q, ctypes.byref(v), ctypes.byref(t)), "job counts")
            slot.sizes.append((v.value, t.value))

    def _admit(self, slot, now, waiting):
        """Admits the longest prefix of the slot's counted jobs whose output fits the budget left,
        and enqueues its pass 4; the other jobs return to the head of the queue. With nothing
        admissible the slot waits (state "blocked") for live outputs to be released; a job whose
        output alone exceeds the whole budget is rejected. Returns the rejected requests."""
        sizes = slot.sizes
        free = self.budget - self.live_bytes
        k, nv, nt = 0, 0, 0
        for (jv, jt) in sizes:
            if (nv + jv) * BYTES_PER_VERTEX + (nt + jt) * BYTES_PER_TRIANGLE > free:
                break
            k, nv, nt = k + 1, nv + jv, nt + jt
        if k == 0:
            first = sizes[0][0] * BYTES_PER_VERTEX + sizes[0][1] * BYTES_PER_TRIANGLE
            if first <= self.budget:
                slot.state = "blocked"
                return []
            r = slot.reqs[0]
            r.rejected, r.done = True, now
            waiting.extendleft(reversed(slot.reqs[1:]))
            slot.reqs, slot.sizes, slot.state = [], [], "idle"
            return [r]
        waiting.extendleft(reversed(slot.reqs[k:]))
        slot.reqs = slot.reqs[:k]
        arena = _Arena(nv, nt, self.device, self)
        self.live_bytes += arena.bytes
        self.peak_live_bytes = max(self.peak_live_bytes, self.live_bytes)
        check(lib().miso_batch_emit_async(slot.h, k, arena.pts.data_ptr(), arena.nrm.data_ptr(),
                                          arena.tris.data_ptr(), max(nv, 1), max(nt, 1)), "emit")
        if self.deliver == "host":
            # delivery: the step's output to pinned host memory, on the slot's stream behind pass 4
            import torch
            with torch.cuda.stream(slot.stream):
                arena.host = tuple(torch.empty(t.numel(), dtype=t.dtype, pin_memory=True)
                                   for t in (arena.pts, arena.nrm, arena.tris))
                for h, d in zip(arena.host, (arena.pts, arena.nrm, arena.tris)):
                    h.copy_(d, non_blocking=True)
        slot.arena = arena
        slot.step = self.steps
        self.steps += 1
        self.step_sizes.append(k)
        for q, r in enumerate(slot.reqs):
            r.admitted = now
            r.nverts, r.ntris = sizes[q]
            r.step = slot.step
        slot.state = "emitting"
        return []

    def _complete(self, slot, now):
        """The step's pass 4 (and, for host delivery, its copy to the host) has finished: each
        request receives its slice of the output."""
        check(lib().miso_batch_wait(slot.h), "wait")
        a = slot.arena
        fv, ft = ctypes.c_uint64(), ctypes.c_uint64()
        src = a.host if self.deliver == "host" else (a.pts, a.nrm, a.tris)
        for q, r in enumerate(slot.reqs):
            check(lib().miso_batch_job_offsets(slot.h, q, ctypes.byref(fv), ctypes.byref(ft)), "offsets")
            v0, t0 = fv.value, ft.value
            pts = src[0][3 * v0:3 * (v0 + r.nverts)].view(-1, 3)
            nrm = src[1][3 * v0:3 * (v0 + r.nverts)].view(-1, 3)
            tri = src[2][3 * t0:3 * (t0 + r.ntris)].view(-1, 3)
            if self.deliver == "host":
                r.result = (pts.numpy(), nrm.numpy(), tri.numpy().view(np.uint32))
            else:
                r.result = (pts, nrm, tri)
                r._arena = a
                a.holders += 1
            r.done = now
        done = slot.reqs
        if self.deliver == "host":
            # delivered to the host: the device arena is no longer needed
            a.holders = 1
            a.release()
        elif a.holders == 0:
            a.holders = 1
            a.release()
        slot.arena, slot.reqs, slot.state = None, [], "idle"
        return done

    # ---------------------------------------------------------------- event loop
    def run(self, requests, clock=time.perf_counter, timeout_s=600.0, on_done=None):
        """Plays the requests at their arrival times; returns them completed, in completion order.
        on_done(request) is called as each request is delivered (a device-delivered request that
        it releases frees its share of the budget at once)."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        finished = []
        total = len(pending)
        t0 = clock()
        while len(finished) < total:
            now = clock() - t0
            if now > timeout_s:
                raise MisoError(f"serving loop exceeded {timeout_s} s with {total - len(finished)} requests left")
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            progressed = False
            for slot in self.slots:
                if slot.state == "emitting" and slot.query():
                    got = self._complete(slot, clock() - t0)
                    finished.extend(got)
                    if on_done is not None:
                        for r in got:
                            on_done(r)
                    progressed = True
            released, self._released = self._released, False
            for slot in self.slots:
                if slot.state == "counting" and slot.query():
                    self._counted(slot)
                    finished.extend(self._admit(slot, clock() - t0, waiting))
                    progressed = True
                elif slot.state == "blocked" and released:
                    finished.extend(self._admit(slot, clock() - t0, waiting))
                    progressed = progressed or slot.state != "blocked"
            for slot in self.slots:
                if slot.state == "idle" and waiting:
                    take = [waiting.popleft() for _ in range(min(self.max_jobs, len(waiting)))]
                    self._start_count(slot, take)
                    progressed = True
            if not progressed:
                if not waiting and pending and all(s.state == "idle" for s in self.slots):
                    gap = pending[0].arrival - (clock() - t0)
                    if gap > 2e-4:
                        time.sleep(gap - 1e-4)
                else:
                    time.sleep(0)
        return finished


def poisson_requests(n, rate, pool_ranges, seed=0):
    """n requests with exponential inter-arrival gaps (mean 1/rate s), a pool volume drawn
    uniformly, and an isovalue uniform over that volume's [min, max]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for q in range(n):
        v = int(rng.integers(0, len(pool_ranges)))
        lo, hi = pool_ranges[v]
        out.append(Request(volume=v, iso=float(np.float32(rng.uniform(lo, hi))), arrival=float(t[q]), rid=q))
    return out


def latency_stats(done):
    lat = np.array([r.latency for r in done if not r.rejected], dtype=np.float64)
    tris = sum(r.ntris for r in done if not r.rejected)
    first = min(r.arrival for r in done)
    last = max(r.done for r in done)
    span = max(last - first, 1e-9)
    return {"requests": len(done), "rejected": sum(r.rejected for r in done),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3) if lat.size else None,
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3) if lat.size else None,
            "mean_ms": round(float(lat.mean()) * 1e3, 3) if lat.size else None,
            "triangles_per_s": round(tris / span, 1), "requests_per_s": round(len(done) / span, 2),
            "span_s": round(span, 3)}
