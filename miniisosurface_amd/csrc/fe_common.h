/*
 * fe_common.h -- device-side data layout and geometry shared by the Flying Edges and Marching
 * Cubes kernels (gfx950, wave64).
 *
 * HBM layout for one extraction over an nx*ny*nz float32 volume (x fastest):
 *   sign mask   uint64 [ny*nz][W], W = ceil(nx/64): bit b of word w of grid row r=(j,k) is
 *               (v[64w+b, j, k] >= iso). This replaces the reference's one-byte-per-edge
 *               edgeCases (flyingEdges/serial/FlyingEdgesAlgorithm.h:86): 1 bit per voxel.
 *   row counts  RowCounts [ny*nz]: cut x-edges of grid row (j,k), cut y-edges (j,k)->(j+1,k),
 *               cut z-edges (j,k)->(j,k+1), and the triangle count of cell row (j,k) (0 on the
 *               last row or plane, which hold no cells). These are the reference's gridEdge
 *               xstart/ystart/zstart and triCounter after pass 2 (FlyingEdgesAlgorithm.cpp:72-190).
 *   row trims   RowTrim [ny*nz]: xl/xr of pass 1 (FlyingEdgesAlgorithm.cpp:47-65).
 *   offsets     uint64 vertex offset and triangle offset per grid row after the scan (pass 3).
 * Vertex order is the reference's: grid rows k-major then j, and within a row all cut x-edges,
 * then y-edges, then z-edges, each by increasing i (FlyingEdgesAlgorithm.cpp:211-230, 340-484).
 */
#ifndef MINIISO_FE_COMMON_H
#define MINIISO_FE_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tables.h"

#pragma clang fp contract(off)

namespace miso {

/* The order in which the lane-row kernels' blocks take the grid rows (plan_block_order in
 * fe_rows.h): blocks [0, n0) go in tiles of 2^lg_tg blocks by 3 planes of G blocks, U tiles per
 * XCD across a plane; the rest, and every block when n0 is 0, in dispatch order. */
struct BlockOrder {
    uint32_t n0, G, U, lg_tg;
};

struct Geom {
    int64_t nx, ny, nz;
    int64_t W;            // mask words per grid row
    float org[3];
    float sp[3];
    // A z-slab of a larger volume: local plane 0 is plane kofs of a volume of nzg planes (the
    // corner coordinates use the global plane index), and pass 4 writes only the vertices and
    // triangles of rows [emit_lo, emit_hi). A whole volume has kofs 0, nzg nz and all rows.
    int64_t kofs, nzg;
    int64_t emit_lo, emit_hi;
    BlockOrder bo;        // planned for ny, nz and W whenever they are set
};

struct RowCounts {
    uint32_t x, y, z, t;
};

struct RowTrim {
    uint32_t xl, xr;
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

/* Orders this wavefront's LDS writes before its later LDS reads by other lanes. Each wavefront of
 * a block works on its own LDS slice, so no block-wide barrier is needed (and waves that finish
 * early never wait for the others). */
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Unaligned (4-byte aligned) loads of 2, 3 or 4 consecutive voxels. */
struct Vox2 { float x, y; };
struct Vox3 { float x, y, z; };
struct Vox4 { float x, y, z, w; };
template <typename T>
__device__ __forceinline__ T load_vox(const float* __restrict__ p)
{
    T t;
    __builtin_memcpy(&t, p, sizeof(T));
    return t;
}

/* 4 voxels in one 16-byte load marked non-temporal: data streamed once (the whole volume in
 * pass 1) is not kept in the caches. 4-byte alignment suffices. */
typedef float Vox4Stream __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ Vox4 load_vox4_stream(const float* __restrict__ p)
{
    const Vox4Stream q = __builtin_nontemporal_load(reinterpret_cast<const Vox4Stream*>(p));
    return Vox4{q.x, q.y, q.z, q.w};
}

/* bits [0, count) of a 64-bit word, count clipped to [0, 64] */
__device__ __forceinline__ uint64_t low_bits(int64_t count)
{
    return count >= 64 ? ~0ull : (count <= 0 ? 0ull : ((1ull << count) - 1ull));
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v);

/* sum over the 64 lanes, in every lane */
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan_u32(v), 63);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* DPP lane moves (VALU data-parallel primitives, no LDS round trip). row_shr:o moves lane i-o
 * to lane i within each row of 16 lanes (0 where i-o leaves the row); row_bcast:15 / :31 copy
 * lane 15 / 31 of the preceding rows into the rows selected by the row mask. */
template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, RowMask, 0xF, true);
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

/* inclusive prefix sum within each row of 16 lanes */
__device__ __forceinline__ uint32_t row16_incl_scan_u32(uint32_t v)
{
    v += dpp_u32<kDppRowShr1>(v);
    v += dpp_u32<kDppRowShr2>(v);
    v += dpp_u32<kDppRowShr4>(v);
    v += dpp_u32<kDppRowShr8>(v);
    return v;
}

/* inclusive prefix sum across the 64 lanes */
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v)
{
    v = row16_incl_scan_u32(v);
    v += dpp_u32<kDppRowBcast15, 0xA>(v);
    v += dpp_u32<kDppRowBcast31, 0xC>(v);
    return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t n = __shfl_up(v, (unsigned)o);
        if (lane >= o) v += n;
    }
    return v;
}

/* Coordinate of grid index idx along one axis as the reference writes it. A cell computes its
 * lower corner as origin + float(index) * spacing and its upper corner as that plus spacing
 * (flyingEdges/util/Image3D.cpp:45-87), and the vertex on the last index n-1 is only ever written
 * by the cell below it (FlyingEdgesAlgorithm.cpp:376-484), so index n-1 is
 * (origin + float(n-2) * spacing) + spacing. */
__device__ __forceinline__ float axis_coord(float o, float s, int64_t idx, int64_t n)
{
    if (idx < n - 1) return o + (float)idx * s;
    return (o + (float)(n - 2) * s) + s;
}

/* One component of the reference's gradient (flyingEdges/util/Image3D.cpp:116-188): one-sided
 * at the faces, central inside, (lower - upper) / run -- the negative, unnormalised gradient. */
__device__ __forceinline__ float grad_axis(const float* __restrict__ d, int64_t idx, int64_t pos,
                                           int64_t n, int64_t stride, float s)
{
    float hi, lo, run;
    if (pos == 0) {
        hi = d[idx + stride]; lo = d[idx]; run = s;
    } else if (pos == n - 1) {
        hi = d[idx]; lo = d[idx - stride]; run = s;
    } else {
        hi = d[idx + stride]; lo = d[idx - stride]; run = 2.0f * s;
    }
    return (lo - hi) / run;
}

/* Point and normal of the isosurface crossing on the edge from grid vertex (i,j,k) one step
 * along `axis`: w = (iso - v0) / (v1 - v0), p = a + w*(b - a) per component
 * (FlyingEdgesAlgorithm.cpp:623-647), the normal interpolates the two vertex gradients. The
 * corner coordinates a and b are passed in, because Flying Edges and Marching Cubes write a
 * shared vertex from different cells and so from differently rounded corner coordinates. */
__device__ __forceinline__ void edge_point(const float* __restrict__ d, const Geom& g, float iso,
                                           int axis, int64_t i, int64_t j, int64_t k,
                                           const float a[3], const float b[3],
                                           float p[3], float n[3])
{
    const int64_t sxy = g.nx * g.ny;
    const int64_t base = i + g.nx * (j + g.ny * k);
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? g.nx : sxy);
    const float v0 = d[base];
    const float v1 = d[base + stride];
    const float w = (iso - v0) / (v1 - v0);
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = a[c] + w * (b[c] - a[c]);
    const int64_t bb = base + stride;
    const int64_t ib = i + (axis == 0), jb = j + (axis == 1), kb = k + (axis == 2);
    float ga[3], gb[3];
    ga[0] = grad_axis(d, base, i, g.nx, 1, g.sp[0]);
    ga[1] = grad_axis(d, base, j, g.ny, g.nx, g.sp[1]);
    ga[2] = grad_axis(d, base, k, g.nz, sxy, g.sp[2]);
    gb[0] = grad_axis(d, bb, ib, g.nx, 1, g.sp[0]);
    gb[1] = grad_axis(d, bb, jb, g.ny, g.nx, g.sp[1]);
    gb[2] = grad_axis(d, bb, kb, g.nz, sxy, g.sp[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] = ga[c] + w * (gb[c] - ga[c]);
}

/* Flying Edges corner coordinates for the edge at grid vertex (i,j,k) along `axis`. */
__device__ __forceinline__ void fe_edge_corners(const Geom& g, int axis, int64_t i, int64_t j,
                                                int64_t k, float a[3], float b[3])
{
    a[0] = axis_coord(g.org[0], g.sp[0], i, g.nx);
    a[1] = axis_coord(g.org[1], g.sp[1], j, g.ny);
    a[2] = axis_coord(g.org[2], g.sp[2], k + g.kofs, g.nzg);
    b[0] = a[0]; b[1] = a[1]; b[2] = a[2];
    b[axis] = a[axis] + g.sp[axis];
}

/* pass 3's totals {V, T} fit output buffers of cap_v vertices and cap_t triangles of Index ids */
template <typename Index>
__device__ __forceinline__ bool output_fits(const uint64_t* __restrict__ need, uint64_t cap_v, uint64_t cap_t)
{
    const uint64_t v = need[0], t = need[1];
    return v <= cap_v && t <= cap_t && (sizeof(Index) == 8 || v <= 0xFFFFFFFFull);
}

template <typename Index>
__device__ __forceinline__ void store_tri(Index* __restrict__ tris, uint64_t slot, uint64_t a,
                                          uint64_t b, uint64_t c)
{
    // non-temporal: the triangle stream is written once and not read back here
    __builtin_nontemporal_store((Index)a, tris + 3 * slot + 0);
    __builtin_nontemporal_store((Index)b, tris + 3 * slot + 1);
    __builtin_nontemporal_store((Index)c, tris + 3 * slot + 2);
}

}  // namespace miso

#endif
