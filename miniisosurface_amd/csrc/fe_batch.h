/*
 * fe_batch.h -- Flying Edges over many volumes (or one volume at many isovalues) in one launch
 * per pass.
 *
 * The jobs' grid rows are concatenated into one row space: the per-row arrays (sign masks,
 * counts, trims, offsets) of job j start at row_base / mask_base. Each pass is one persistent
 * kernel whose blocks take chunks of work items from a device ticket counter; an item is one
 * block of the single-volume kernels (4 wavefronts of lane rows for counting and emission) of
 * one job, found by a binary search of the jobs' item prefix. Classification works on groups of
 * consecutive jobs that read the same volume (one volume at several isovalues): an item is 4
 * grid rows of the group (16 when rows hold at most 256 voxels), loaded once for all its jobs.
 * Pass 3 is the single-volume scan over the concatenated rows, so job j's output is the
 * contiguous range [voff[row_base_j], voff[row_base_{j+1}]) of one shared buffer; its triangles
 * index its own vertices from 0. k_batch_fused runs the same queues and block functions for all
 * four passes in one cooperative launch, with grid barriers between the passes.
 */
#ifndef MINIISO_FE_BATCH_H
#define MINIISO_FE_BATCH_H

#include <hip/hip_cooperative_groups.h>

#include "fe_rows.h"

namespace miso {

constexpr int kClassifyGroup = 16;   // jobs of one volume classified together (pass 1 of a batch)

/* rows per pass 1 work item of a batch job (k_batch_classify) */
inline int64_t batch_classify_rows(const Geom& g) { return g.W <= 4 ? 4 * kShortRows : kRowsPerBlock; }

struct BatchJob {
    Geom g;
    const float* vol;
    float iso;
    int unit;            // unit spacing and zero origin (the exact fast path of pass 4)
    int64_t row_base;    // first row of the job in the concatenated row arrays
    int64_t mask_base;   // first sign-mask word
};

/* the job whose item range holds `item`: the largest j < njobs with prefix[j] <= item */
__device__ __forceinline__ int find_job(const int64_t* __restrict__ prefix, int njobs, int64_t item)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

/* Persistent loop: the block takes `chunk` consecutive items per ticket until the queue is
 * drained (every block exits once a ticket lands past nitems); body(job, item - prefix[job]). */
template <class Body>
__device__ __forceinline__ void run_queue(uint64_t* __restrict__ ticket, const int64_t* __restrict__ prefix,
                                          int njobs, int64_t nitems, int chunk, Body body)
{
    __shared__ int64_t s_base;
    __shared__ int s_job;
    for (;;) {
        __syncthreads();   // every thread has read the previous ticket
        if (threadIdx.x == 0) {
            const int64_t b = (int64_t)__hip_atomic_fetch_add((unsigned long long*)ticket, (unsigned long long)chunk,
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_base = b;
            s_job = b < nitems ? find_job(prefix, njobs, b) : 0;
        }
        __syncthreads();
        const int64_t base = s_base;
        if (base >= nitems) break;
        int j = __builtin_amdgcn_readfirstlane(s_job);
        const int64_t end = base + chunk < nitems ? base + chunk : nitems;
        for (int64_t it = base; it < end; ++it) {
            while (it >= prefix[j + 1]) ++j;
            body(j, it - prefix[j]);
        }
    }
}

/* The XCD-aware variant of run_queue for pass 4. A persistent block b runs on XCD b mod 8 (the
 * dispatcher hands workgroups to the 8 XCDs in turn), and each XCD takes its items from a queue of
 * its own (ticket x): within job j, XCD x owns the dispatch indices b = 8 t + x < n_j, which
 * emit_block_order maps to the job's row blocks as it maps the single-volume kernel's blocks, so
 * each XCD's L2 sees tiles of neighbouring rows, and all XCDs move through a job's planes
 * together. xprefix[x stride + j]: XCD x's items before job j (jobs [0, njobs) are drawn from).
 * body(job, dispatch index). */
template <class Body>
__device__ __forceinline__ void run_queue_xcd(uint64_t* __restrict__ tickets, const int64_t* __restrict__ xprefix,
                                              int64_t stride, int njobs, int chunk, Body body)
{
    __shared__ int64_t s_base;
    __shared__ int s_job;
    const int x = (int)(blockIdx.x & 7u);
    const int64_t* __restrict__ pre = xprefix + (int64_t)x * stride;
    const int64_t nitems = pre[njobs];
    for (;;) {
        __syncthreads();   // every thread has read the previous ticket
        if (threadIdx.x == 0) {
            const int64_t b = (int64_t)__hip_atomic_fetch_add((unsigned long long*)&tickets[x],
                                                              (unsigned long long)chunk, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
            s_base = b;
            s_job = b < nitems ? find_job(pre, njobs, b) : 0;
        }
        __syncthreads();
        const int64_t base = s_base;
        if (base >= nitems) break;
        int j = __builtin_amdgcn_readfirstlane(s_job);
        const int64_t end = base + chunk < nitems ? base + chunk : nitems;
        for (int64_t it = base; it < end; ++it) {
            while (it >= pre[j + 1]) ++j;
            body(j, 8 * (it - pre[j]) + x);
        }
    }
}

/* clears the queue tickets and the scan's tile status (a kernel, so it can sit in a graph) */
__global__ __launch_bounds__(256) void k_batch_zero(uint64_t* __restrict__ words, int64_t n)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        words[t] = 0;
}

/* Pass 1 over groups of jobs: group q is jobs [group_job[q], group_job[q + 1]), consecutive jobs
 * that read the same volume with the same dimensions (one volume at several isovalues); an item
 * loads a block of its rows once and writes every job's sign-mask words. */
/* Pass 1 item blk of classify group q (the jobs [group_job[q], group_job[q + 1]) on one volume). */
__device__ __forceinline__ void batch_classify_item(const BatchJob* __restrict__ jobs,
                                                    const int* __restrict__ group_job, uint64_t* __restrict__ mask,
                                                    int q, int64_t blk)
{
    const int j0 = group_job[q], n = group_job[q + 1] - j0;
    const BatchJob& J = jobs[j0];
    auto sink = [&](int t, float& iso, uint64_t*& m) {
        iso = jobs[j0 + t].iso;
        m = mask + jobs[j0 + t].mask_base;
    };
    // an item is 4 kShortRows rows of a short-row job (batch_classify_rows), else 4 rows
    if (J.g.W <= 4)
        fe_classify_short_rows_n(J.vol, J.g, n, sink, blk);
    else
        fe_classify_block_n(J.vol, J.g, n, sink, blk);
}

__global__ __launch_bounds__(256) void k_batch_classify(const BatchJob* __restrict__ jobs,
                                                        const int* __restrict__ group_job,
                                                        const int64_t* __restrict__ prefix, int ngroups,
                                                        int64_t nitems, int chunk, uint64_t* __restrict__ mask,
                                                        uint64_t* __restrict__ ticket)
{
    run_queue(ticket, prefix, ngroups, nitems, chunk,
              [&](int q, int64_t blk) { batch_classify_item(jobs, group_job, mask, q, blk); });
}

__global__ __launch_bounds__(256) void k_batch_count(const BatchJob* __restrict__ jobs,
                                                     const int64_t* __restrict__ prefix, int njobs, int64_t nitems,
                                                     int chunk, const uint64_t* __restrict__ mask,
                                                     RowCounts* __restrict__ counts, RowTrim* __restrict__ trims,
                                                     uint64_t* __restrict__ ticket)
{
    run_queue(ticket, prefix, njobs, nitems, chunk, [&](int j, int64_t blk) {
        const BatchJob& J = jobs[j];
        fe_count_rows_block(mask + J.mask_base, J.g, counts + J.row_base, trims + J.row_base, blk);
    });
}

/* Job j's first vertex, first triangle, vertex and triangle counts (from the scan). */
__device__ __forceinline__ void batch_job_counts_one(const BatchJob* __restrict__ jobs, int njobs, int64_t R,
                                                     const uint64_t* __restrict__ voff,
                                                     const uint64_t* __restrict__ toff,
                                                     const uint64_t* __restrict__ totals, uint64_t* __restrict__ out,
                                                     int j)
{
    const int64_t b = jobs[j].row_base;
    const int64_t e = j + 1 < njobs ? jobs[j + 1].row_base : R;
    const uint64_t vb = b < R ? voff[b] : totals[0], tb = b < R ? toff[b] : totals[1];
    const uint64_t ve = e < R ? voff[e] : totals[0], te = e < R ? toff[e] : totals[1];
    out[4 * j + 0] = vb;
    out[4 * j + 1] = tb;
    out[4 * j + 2] = ve - vb;
    out[4 * j + 3] = te - tb;
}

__global__ __launch_bounds__(256) void k_batch_job_counts(const BatchJob* __restrict__ jobs, int njobs, int64_t R,
                                                          const uint64_t* __restrict__ voff,
                                                          const uint64_t* __restrict__ toff,
                                                          const uint64_t* __restrict__ totals,
                                                          uint64_t* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs) batch_job_counts_one(jobs, njobs, R, voff, toff, totals, out, j);
}

/* Pass 4 of job j's row block blk into the shared output; the job's triangles index its own
 * vertices, from its first one, voff at its first row. */
__device__ __forceinline__ void batch_emit_item(const uint64_t* __restrict__ s_case, EmitRowsLds* s_w,
                                                const BatchJob* __restrict__ jobs, const uint64_t* __restrict__ mask,
                                                const RowCounts* __restrict__ counts,
                                                const uint64_t* __restrict__ voff, const uint64_t* __restrict__ toff,
                                                float* __restrict__ pts, float* __restrict__ nrm,
                                                uint32_t* __restrict__ tris, int j, int64_t blk)
{
    const BatchJob& J = jobs[j];
    const uint64_t vbias = voff[J.row_base];
    if (J.unit)
        fe_emit_rows_block<uint32_t, true>(s_case, s_w, J.vol, mask + J.mask_base, J.g, J.iso, counts + J.row_base,
                                           voff + J.row_base, toff + J.row_base, pts, nrm, tris, vbias, blk);
    else
        fe_emit_rows_block<uint32_t, false>(s_case, s_w, J.vol, mask + J.mask_base, J.g, J.iso, counts + J.row_base,
                                            voff + J.row_base, toff + J.row_base, pts, nrm, tris, vbias, blk);
}

/* Pass 4 for the first nemit jobs, from the one queue (items [0, nitems), tickets[0]) or, with
 * xprefix, the XCD queues (run_queue_xcd, tickets[0 .. 8)). */
__device__ __forceinline__ void batch_emit_queue(const uint64_t* __restrict__ s_case, EmitRowsLds* s_w,
                                                 const BatchJob* __restrict__ jobs, const int64_t* __restrict__ prefix,
                                                 int njobs, int64_t nitems, int chunk,
                                                 const uint64_t* __restrict__ mask,
                                                 const RowCounts* __restrict__ counts,
                                                 const uint64_t* __restrict__ voff, const uint64_t* __restrict__ toff,
                                                 float* __restrict__ pts, float* __restrict__ nrm,
                                                 uint32_t* __restrict__ tris, uint64_t* __restrict__ tickets,
                                                 const int64_t* __restrict__ xprefix, int nemit)
{
    if (xprefix)
        run_queue_xcd(tickets, xprefix, (int64_t)njobs + 1, nemit, chunk, [&](int j, int64_t b) {
            batch_emit_item(s_case, s_w, jobs, mask, counts, voff, toff, pts, nrm, tris, j,
                            emit_block_order(b, prefix[j + 1] - prefix[j], jobs[j].g));
        });
    else
        run_queue(tickets, prefix, njobs, nitems, chunk, [&](int j, int64_t blk) {
            batch_emit_item(s_case, s_w, jobs, mask, counts, voff, toff, pts, nrm, tris, j, blk);
        });
}

/* Pass 4 for the first jobs (items [0, nitems)); writes nothing when their output exceeds the
 * capacity (the host then grows the buffers and runs this pass again). */
__global__ __launch_bounds__(256) void k_batch_emit(const BatchJob* __restrict__ jobs,
                                                    const int64_t* __restrict__ prefix, int njobs, int64_t nitems,
                                                    int chunk, const uint64_t* __restrict__ mask,
                                                    const RowCounts* __restrict__ counts,
                                                    const uint64_t* __restrict__ voff,
                                                    const uint64_t* __restrict__ toff,
                                                    const uint64_t* __restrict__ need, uint64_t cap_v,
                                                    uint64_t cap_t, float* __restrict__ pts, float* __restrict__ nrm,
                                                    uint32_t* __restrict__ tris, uint64_t* __restrict__ tickets,
                                                    const int64_t* __restrict__ xprefix, int nemit)
{
    __shared__ uint64_t s_case[256];
    __shared__ EmitRowsLds s_w[4];
    // need: vertices and triangles of the emitted jobs (the totals, or the first vertex and
    // triangle of the first job not emitted)
    if (need[0] > cap_v || need[1] > cap_t) return;
    s_case[threadIdx.x] = kCaseEdges[threadIdx.x];
    batch_emit_queue(s_case, s_w, jobs, prefix, njobs, nitems, chunk, mask, counts, voff, toff, pts, nrm, tris,
                     tickets, xprefix, nemit);
}

/* Everything batch_run does on the device after clearing the tickets and the scan state, in one
 * cooperative launch: the grid (every block resident at once) runs pass 1 over the classify
 * groups, pass 2, the scan (tiles taken in order from its ticket) and pass 4 (when the totals
 * fit the output), with a grid-wide barrier between the phases, which also orders every block's
 * writes before the next phase's reads; the per-job counts for the host are written beside
 * pass 4, which does not read them. The queues and block functions are those of the separate
 * kernels, so the output is theirs. state: [0..2] tickets, [3..4] totals, [5..12] the XCD
 * tickets, [13..] the scan state (its ticket, then ntiles vertex and ntiles triangle tile
 * status words). */
__global__ __launch_bounds__(256) void k_batch_fused(const BatchJob* __restrict__ jobs,
                                                     const int* __restrict__ group_job,
                                                     const int64_t* __restrict__ cls_prefix, int ngroups, int64_t ncls,
                                                     const int64_t* __restrict__ lane_prefix, int njobs,
                                                     int64_t nlane, int chunk, uint64_t* __restrict__ mask,
                                                     RowCounts* __restrict__ counts, RowTrim* __restrict__ trims,
                                                     uint64_t* __restrict__ voff, uint64_t* __restrict__ toff,
                                                     uint64_t* __restrict__ state, int64_t R, int64_t ntiles,
                                                     uint64_t* __restrict__ jobout, uint64_t cap_v, uint64_t cap_t,
                                                     float* __restrict__ pts, float* __restrict__ nrm,
                                                     uint32_t* __restrict__ tris, const int64_t* __restrict__ xprefix)
{
    namespace cg = cooperative_groups;
    cg::grid_group grid = cg::this_grid();
    __shared__ uint64_t s_case[256];
    __shared__ EmitRowsLds s_w[4];
    s_case[threadIdx.x] = kCaseEdges[threadIdx.x];
    run_queue(state + 0, cls_prefix, ngroups, ncls, chunk,
              [&](int q, int64_t blk) { batch_classify_item(jobs, group_job, mask, q, blk); });
    grid.sync();
    run_queue(state + 1, lane_prefix, njobs, nlane, chunk, [&](int j, int64_t blk) {
        const BatchJob& J = jobs[j];
        fe_count_rows_block(mask + J.mask_base, J.g, counts + J.row_base, trims + J.row_base, blk);
    });
    grid.sync();
    uint64_t* scan = state + 13;
    uint64_t* totals = state + 3;
    for (;;) {
        const int64_t tile = fe_scan_ticket(scan);
        if (tile >= ntiles) break;
        fe_scan_tile(counts, R, voff, toff, scan, ntiles, totals, tile);
    }
    grid.sync();
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < njobs; j += (int64_t)gridDim.x * blockDim.x)
        batch_job_counts_one(jobs, njobs, R, voff, toff, totals, jobout, (int)j);
    if (totals[0] > cap_v || totals[1] > cap_t) return;
    batch_emit_queue(s_case, s_w, jobs, lane_prefix, njobs, nlane, chunk, mask, counts, voff, toff, pts, nrm, tris,
                     xprefix ? state + 5 : state + 2, xprefix, njobs);
}

}  // namespace miso

#endif
