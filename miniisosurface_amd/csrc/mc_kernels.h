/*
 * mc_kernels.h -- Marching Cubes with the serial reference's vertex numbering, as HIP kernels.
 *
 * The serial reference (marchingCubes/serial/main.cpp:42-192) visits every cell in z, y, x
 * order and numbers a cut edge the first time a cell's triangle list names it (through a hash
 * map keyed by the global edge id, :145-177). Every cut edge is named by each adjacent cell, so
 * the first visitor is the adjacent cell with the smallest (k, j, i): the "owner". A cell owns
 * its edges on the far side of each axis, plus those on the near side where it is the first
 * cell along that axis (i == 0, j == 0 or k == 0). Vertex ids therefore follow (owner cell in
 * linear order, first slot of the edge in the owner's case triangle list) -- no hash map needed.
 *
 *   k_mc_count      owned cut edges per cell row
 *   k_scan_u32      decoupled look-back scan of those counts (vertex offsets per cell row)
 *   k_mc_vertices   owners write point and normal (with the owner's corner coordinates,
 *                   marchingCubes/util/Image3D.cpp:22-65) and record, at the edge's Flying Edges
 *                   rank, the Marching Cubes id
 *   k_mc_triangles  triangles in the same order as Flying Edges, ids read through that map
 */
#ifndef MINIISO_MC_KERNELS_H
#define MINIISO_MC_KERNELS_H

#include "fe_kernels.h"

namespace miso {

/* 12-bit mask of the cube edges a case cuts: the edges its triangles name. */
__device__ __forceinline__ uint32_t case_cut_mask(uint64_t packed)
{
    uint32_t m = 0;
    const int ne = 3 * case_tri_count(packed);
    for (int s = 0; s < ne; ++s) m |= 1u << case_edge(packed, s);
    return m;
}

/* Edges a cell owns under first-visit numbering, given whether it is first along x, y, z. */
__device__ __forceinline__ uint32_t owned_edges(bool i0, bool j0, bool k0)
{
    uint32_t m = (1u << 6) | (1u << 5) | (1u << 11);
    if (k0) m |= (1u << 2) | (1u << 1);
    if (j0) m |= (1u << 4) | (1u << 9);
    if (i0) m |= (1u << 7) | (1u << 10);
    if (j0 && k0) m |= 1u << 0;
    if (i0 && k0) m |= 1u << 3;
    if (i0 && j0) m |= 1u << 8;
    return m;
}

__global__ __launch_bounds__(256) void k_mc_count(const uint64_t* __restrict__ mask, Geom g,
                                                  const RowCounts* __restrict__ counts,
                                                  uint32_t* __restrict__ owned,
                                                  uint64_t* __restrict__ scan_state, int64_t scan_state_words)
{
    __shared__ EmitLds lds_all[kRowsPerBlock];
    __shared__ uint16_t s_cut[256];
    for (int c = threadIdx.x; c < 256; c += blockDim.x) s_cut[c] = (uint16_t)case_cut_mask(kCaseEdges[c]);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < scan_state_words;
         t += (int64_t)gridDim.x * blockDim.x)
        scan_state[t] = 0;
    __syncthreads();
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + wave;
    if (r >= g.ny * g.nz) return;
    const int64_t j = r % g.ny, k = r / g.ny;
    const bool has_y = j < g.ny - 1, has_z = k < g.nz - 1;
    if (!(has_y && has_z) || counts[r].t == 0) {
        if (lane == 0) owned[r] = 0;
        return;
    }
    RowWalker rw(lds_all[wave], mask, g, r, has_y, has_z);
    EmitLds& L = lds_all[wave];
    uint32_t total = 0;
    for (int64_t s0w = 0; s0w < g.W; s0w += kSegWords) {
        rw.stage(s0w);
        const int na = rw.prepare_cells();
        for (int e = lane; e < na; e += 64) {
            const int li = L.list[e];
            const int64_t i = s0w * 64 + li;
            const int cs = rw.cell_case(li);
            total += __popc(s_cut[cs] & owned_edges(i == 0, j == 0, k == 0));
        }
        wave_sync();
    }
    total = wave_sum_u32(total);
    if (lane == 0) owned[r] = total;
}

/* Exclusive scan of a uint32 per grid row (vals[r * stride]) into uint64 offsets (same look-back
 * as pass 3). */
__global__ __launch_bounds__(256) void k_scan_u32(const uint32_t* __restrict__ vals, int64_t R,
                                                  uint64_t* __restrict__ out, uint64_t* __restrict__ state,
                                                  int64_t ntiles, uint64_t* __restrict__ total, int64_t stride = 1)
{
    __shared__ int64_t s_tile;
    __shared__ uint64_t s_w[4];
    __shared__ uint64_t s_e;
    if (threadIdx.x == 0)
        s_tile = (int64_t)__hip_atomic_fetch_add((unsigned long long*)&state[0], 1ull, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int64_t tile = s_tile;
    uint64_t* st = state + 1;
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int64_t r0 = tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems], sv = 0;
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
        const int64_t r = r0 + u;
        v[u] = r < R ? vals[r * stride] : 0;
        sv += v[u];
    }
    const uint64_t iv = wave_incl_scan_u64(sv);
    if (lane == 63) s_w[wave] = iv;
    __syncthreads();
    uint64_t wb = 0, agg = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q < wave) wb += s_w[q];
        agg += s_w[q];
    }
    if (wave == 0) {
        uint64_t ev = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(&st[0], kFlagPre | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(&st[tile], kFlagAgg | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ev = lookback(st, tile);
            if (lane == 0) __hip_atomic_store(&st[tile], kFlagPre | (ev + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_e = ev;
            if (tile == ntiles - 1) total[0] = ev + agg;
        }
    }
    __syncthreads();
    uint64_t p = s_e + wb + iv - sv;
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
        const int64_t r = r0 + u;
        if (r < R) out[r] = p;
        p += v[u];
    }
}

/* EdgeOrder: each owned edge's point goes to fe_to_mc[its Flying Edges id] (an edge-order
 * permutation made beforehand), which this kernel then only reads; moff is not used. */
template <typename Index, bool EdgeOrder = false>
__global__ __launch_bounds__(256) void k_mc_vertices(const float* __restrict__ vol,
                                                     const uint64_t* __restrict__ mask, Geom g, float iso,
                                                     const RowCounts* __restrict__ counts,
                                                     const uint64_t* __restrict__ voff,
                                                     const uint64_t* __restrict__ moff,
                                                     float* __restrict__ pts, float* __restrict__ nrm,
                                                     Index* __restrict__ fe_to_mc)
{
    __shared__ EmitLds lds_all[kRowsPerBlock];
    __shared__ uint16_t s_cut[256];
    for (int c = threadIdx.x; c < 256; c += blockDim.x) s_cut[c] = (uint16_t)case_cut_mask(kCaseEdges[c]);
    __syncthreads();
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + wave;
    if (r >= g.ny * g.nz) return;
    const int64_t j = r % g.ny, k = r / g.ny;
    const bool has_y = j < g.ny - 1, has_z = k < g.nz - 1;
    if (!(has_y && has_z) || counts[r].t == 0) return;
    RowWalker rw(lds_all[wave], mask, g, r, has_y, has_z);
    EmitLds& L = lds_all[wave];
    uint64_t cbase[8];
    combo_bases(counts, voff, rw.rows, cbase);
    uint64_t mrun = EdgeOrder ? 0 : moff[r];
    const float Y = g.org[1] + (float)j * g.sp[1], Z = g.org[2] + (float)k * g.sp[2];
    for (int64_t s0w = 0; s0w < g.W; s0w += kSegWords) {
        rw.stage(s0w);
        const int na = rw.prepare_cells();
        for (int e0 = 0; e0 < na; e0 += 64) {
            const int e = e0 + lane;
            int li = 0, cs = 0;
            uint32_t own = 0;
            if (e < na) {
                li = L.list[e];
                cs = rw.cell_case(li);
                own = s_cut[cs] & owned_edges(s0w * 64 + li == 0, j == 0, k == 0);
            }
            const uint32_t o = (uint32_t)__popc(own);
            const uint32_t incl = wave_incl_scan_u32(o);
            if (own) {
                const int64_t i = s0w * 64 + li;
                const float X = g.org[0] + (float)i * g.sp[0];
                const uint64_t packed = kCaseEdges[cs];
                uint64_t vid = mrun + incl - o;
                uint32_t seen = 0;
                const int ne = 3 * case_tri_count(packed);
                for (int s = 0; s < ne; ++s) {
                    const int ed = case_edge(packed, s);
                    if (!((own >> ed) & 1u) || ((seen >> ed) & 1u)) continue;
                    seen |= 1u << ed;
                    const int a = edge_axis(ed);
                    const int dx = edge_dx(ed), dy = edge_dy(ed), dz = edge_dz(ed);
                    // corner coordinates as the owner cell computes them (getPosCube)
                    float A[3] = {dx ? X + g.sp[0] : X, dy ? Y + g.sp[1] : Y, dz ? Z + g.sp[2] : Z};
                    float B[3] = {A[0], A[1], A[2]};
                    B[a] = A[a] + g.sp[a];
                    float p[3], nn[3];
                    edge_point(vol, g, iso, a, i + dx, j + dy, k + dz, A, B, p, nn);
                    const uint64_t fe = rw.fe_vertex_id(cbase, ed, li);
                    if (EdgeOrder) vid = (uint64_t)fe_to_mc[fe];
                    pts[3 * vid + 0] = p[0];
                    pts[3 * vid + 1] = p[1];
                    pts[3 * vid + 2] = p[2];
                    nrm[3 * vid + 0] = nn[0];
                    nrm[3 * vid + 1] = nn[1];
                    nrm[3 * vid + 2] = nn[2];
                    if (!EdgeOrder) fe_to_mc[fe] = (Index)vid;
                    ++vid;
                }
            }
            mrun += __shfl(incl, 63);
        }
        rw.advance_bases(cbase);
        wave_sync();
    }
}

template <typename Index>
__global__ __launch_bounds__(256) void k_mc_triangles(const uint64_t* __restrict__ mask, Geom g,
                                                      const RowCounts* __restrict__ counts,
                                                      const uint64_t* __restrict__ voff,
                                                      const uint64_t* __restrict__ toff,
                                                      const Index* __restrict__ fe_to_mc,
                                                      Index* __restrict__ tris)
{
    __shared__ EmitLds lds_all[kRowsPerBlock];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + wave;
    if (r >= g.ny * g.nz) return;
    const int64_t j = r % g.ny, k = r / g.ny;
    const bool has_y = j < g.ny - 1, has_z = k < g.nz - 1;
    if (!(has_y && has_z) || counts[r].t == 0) return;
    RowWalker rw(lds_all[wave], mask, g, r, has_y, has_z);
    EmitLds& L = lds_all[wave];
    uint64_t cbase[8];
    combo_bases(counts, voff, rw.rows, cbase);
    uint64_t tri_run = toff[r];
    for (int64_t s0w = 0; s0w < g.W; s0w += kSegWords) {
        rw.stage(s0w);
        const int na = rw.prepare_cells();
        for (int e0 = 0; e0 < na; e0 += 64) {
            const int e = e0 + lane;
            int cs = 0, nt = 0, li = 0;
            if (e < na) {
                li = L.list[e];
                cs = rw.cell_case(li);
                nt = case_tri_count(kCaseEdges[cs]);
            }
            const uint32_t incl = wave_incl_scan_u32((uint32_t)nt);
            const uint64_t slot0 = tri_run + incl - (uint32_t)nt;
            if (nt) {
                const uint64_t packed = kCaseEdges[cs];
                for (int t = 0; t < nt; ++t) {
                    const uint64_t a = fe_to_mc[rw.fe_vertex_id(cbase, case_edge(packed, 3 * t + 0), li)];
                    const uint64_t b = fe_to_mc[rw.fe_vertex_id(cbase, case_edge(packed, 3 * t + 1), li)];
                    const uint64_t c = fe_to_mc[rw.fe_vertex_id(cbase, case_edge(packed, 3 * t + 2), li)];
                    store_tri<Index>(tris, slot0 + t, a, b, c);
                }
            }
            tri_run += __shfl(incl, 63);
        }
        rw.advance_bases(cbase);
        wave_sync();
    }
}

}  // namespace miso

#endif
