/*
 * fe_rows.h -- Flying Edges passes 2 and 4 with (grid row, mask word) pairs on the lanes.
 *
 * A wavefront covers RPW consecutive grid rows of W mask words each (RPW = 64 / Wp, Wp the next
 * power of two >= W, so nx <= 4096 voxels per row); lane (q, w) owns word w of row q. Row
 * totals and offsets are segmented wave reductions and prefix sums over the Wp lanes of a row.
 * Consecutive grid rows own consecutive vertex blocks, so a wavefront's vertices form one
 * contiguous range of the output, and its triangles one contiguous range too.
 *
 *   k_fe_count_rows  pass 1b + 2 (trims, x/y/z cut counts, triangle counts per row)
 *   k_fe_emit_rows   pass 4: vertices (cut edges compacted into LDS lists, x-edges then y/z-edges,
 *                    one vertex per lane, written at consecutive addresses), then triangles (active
 *                    cells listed in output order, resolved one triangle per lane, vertex ids from
 *                    row offsets plus popcount ranks held in LDS)
 */
#ifndef MINIISO_FE_ROWS_H
#define MINIISO_FE_ROWS_H

#include "fe_kernels.h"

namespace miso {

struct RowLanes {
    int Wp;       // lanes per row (power of two, >= W)
    int rpw;      // rows per wavefront
    int q;        // this lane's row within the wavefront
    int w;        // this lane's word within the row
};

__device__ __forceinline__ RowLanes row_lanes(int64_t W)
{
    RowLanes L;
    const int lg = W <= 1 ? 0 : 64 - __builtin_clzll((unsigned long long)(W - 1));   // Wp = 2^lg >= W
    L.Wp = 1 << lg;
    L.rpw = 64 >> lg;
    const int lane = lane_id();
    L.q = lane >> lg;
    L.w = lane & (L.Wp - 1);
    return L;
}

/* This wavefront's index within its 256-thread block, in a scalar register. */
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

/* a / b for values that are the same in every lane (scalar code; 32-bit when both fit) */
__host__ __device__ __forceinline__ int64_t uniform_div(int64_t a, int64_t b)
{
    if (((uint64_t)a | (uint64_t)b) >> 32 == 0) return (int64_t)((uint32_t)a / (uint32_t)b);
    return a / b;
}

/* (j, k) of grid row r0 + q, for a wavefront whose first row r0 is the same in every lane and
 * 0 <= q < 64: one division per wavefront, then carries per lane (at most one once ny >= 64). */
struct RowJK {
    int64_t j, k;
};

__host__ __device__ __forceinline__ RowJK split_row(int64_t r0, int q, int64_t ny)
{
    RowJK p;
    p.k = uniform_div(r0, ny);
    p.j = r0 - p.k * ny + q;
    while (p.j >= ny) {
        p.j -= ny;
        ++p.k;
    }
    return p;
}

/* Prefix scans and reductions over the Wp lanes of one row segment (Wp a power of two, the
 * segments aligned), with DPP row shifts and broadcasts. */
struct OpAdd {
    __device__ static uint32_t f(uint32_t a, uint32_t b) { return a + b; }
};
struct OpMin {
    __device__ static uint32_t f(uint32_t a, uint32_t b) { return a < b ? a : b; }
};
struct OpMax {
    __device__ static uint32_t f(uint32_t a, uint32_t b) { return a > b ? a : b; }
};

template <typename Op>
__device__ __forceinline__ uint32_t seg_incl(uint32_t v, int Wp, int w)
{
    uint32_t n;
    n = dpp_u32<kDppRowShr1>(v);
    if (Wp > 1 && w >= 1) v = Op::f(v, n);
    n = dpp_u32<kDppRowShr2>(v);
    if (Wp > 2 && w >= 2) v = Op::f(v, n);
    n = dpp_u32<kDppRowShr4>(v);
    if (Wp > 4 && w >= 4) v = Op::f(v, n);
    n = dpp_u32<kDppRowShr8>(v);
    if (Wp > 8 && w >= 8) v = Op::f(v, n);
    if (Wp > 16) {
        n = dpp_u32<kDppRowBcast15, 0xA>(v);
        if (w & 16) v = Op::f(v, n);
    }
    if (Wp > 32) {
        n = dpp_u32<kDppRowBcast31, 0xC>(v);
        if (w & 32) v = Op::f(v, n);
    }
    return v;
}

__device__ __forceinline__ uint32_t seg_excl_scan(uint32_t v, int Wp, int w) { return seg_incl<OpAdd>(v, Wp, w) - v; }

/* segment total (or min / max), valid in every lane of the segment */
template <typename Op>
__device__ __forceinline__ uint32_t seg_reduce(uint32_t v, int Wp)
{
    const int lane = lane_id();
    const uint32_t inc = seg_incl<Op>(v, Wp, lane & (Wp - 1));
    return (uint32_t)__shfl(inc, (lane & ~(Wp - 1)) + Wp - 1);
}

__device__ __forceinline__ uint32_t seg_sum(uint32_t v, int Wp) { return seg_reduce<OpAdd>(v, Wp); }

/* Sign-mask words around lane (row r, word w): rows (j,k) (j+1,k) (j,k+1) (j+1,k+1), each with
 * the next word for the +1 neighbours. Rows outside the grid read as zero. */
struct LaneWords {
    uint64_t m[4], n[4];
};

__device__ __forceinline__ LaneWords load_lane_words(const uint64_t* __restrict__ mask, const Geom& g, int64_t r,
                                                     int64_t w, bool ok, bool has_y, bool has_z)
{
    LaneWords L;
    const int64_t W = g.W;
    const int64_t rows[4] = {r, r + 1, r + g.ny, r + g.ny + 1};
    const bool rok[4] = {ok, ok && has_y, ok && has_z, ok && has_y && has_z};
    const bool nxt = w + 1 < W;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        L.m[q] = rok[q] ? mask[rows[q] * W + w] : 0ull;
        L.n[q] = (rok[q] && nxt) ? mask[rows[q] * W + w + 1] : 0ull;
    }
    return L;
}

__device__ __forceinline__ uint64_t shifted(uint64_t m, uint64_t next) { return (m >> 1) | (next << 63); }

/* The 8 corner words of a lane: sign bits at i (m) and at i+1 (s) of the 4 rows. */
struct CornerWords {
    uint64_t m[4], s[4];
};

__device__ __forceinline__ CornerWords corner_words(const LaneWords& L)
{
    CornerWords C;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        C.m[q] = L.m[q];
        C.s[q] = shifted(L.m[q], L.n[q]);
    }
    return C;
}

/* A cell's marching-cubes case has corner bit q = v_q, in the reference's order
 * (FlyingEdgesAlgorithm.cpp:138-157): v0..v7 = m0 s0 s1 m1 m2 s2 s3 m3 of CornerWords, i.e. rows
 * (j,k), (j+1,k), (j,k+1), (j+1,k+1) at x = i (m) and i + 1 (s), taken (i,j) (i+1,j) (i+1,j+1)
 * (i,j+1) in plane k, then the same in plane k + 1. */

__device__ __forceinline__ uint64_t active_cells(const CornerWords& C, uint64_t eb)
{
    const uint64_t all = C.m[0] & C.m[1] & C.m[2] & C.m[3] & C.s[0] & C.s[1] & C.s[2] & C.s[3];
    const uint64_t any = C.m[0] | C.m[1] | C.m[2] | C.m[3] | C.s[0] | C.s[1] | C.s[2] | C.s[3];
    return any & ~all & eb;
}

__device__ __forceinline__ uint64_t active_cells(const LaneWords& L, uint64_t eb)
{
    return active_cells(corner_words(L), eb);
}

/* Triangles of the marching-cubes cases of a word's 64 cells, bit-sliced over the cells. For a
 * cell with corners on both sides of the isovalue, the case table's triangle count
 * (numTris, flyingEdges/util/MarchingCubesTables.h:23-41) equals V - 2F - 4S: V the corners
 * >= iso, F the faces whose four corners are >= iso, S = 1 for the six-corner rings (both ends
 * of a body diagonal below iso, the other six corners >= iso). Cells with all corners >= iso are
 * masked out; cells with none contribute 0. */
__device__ __forceinline__ uint32_t word_tri_count(const CornerWords& C, uint64_t eb)
{
    // corners v0..v7 of the cells (the case's bit order)
    const uint64_t v0 = C.m[0], v1 = C.s[0], v2 = C.s[1], v3 = C.m[1];
    const uint64_t v4 = C.m[2], v5 = C.s[2], v6 = C.s[3], v7 = C.m[3];
    const uint64_t a01 = v0 & v1, a23 = v2 & v3, a45 = v4 & v5, a67 = v6 & v7;
    const uint64_t a03 = v0 & v3, a12 = v1 & v2, a56 = v5 & v6, a47 = v4 & v7;
    const uint64_t M = eb & ~(a01 & a23 & a45 & a67);
    uint32_t V = 0;
    V += (uint32_t)__popcll(v0 & M) + (uint32_t)__popcll(v1 & M) + (uint32_t)__popcll(v2 & M) +
         (uint32_t)__popcll(v3 & M);
    V += (uint32_t)__popcll(v4 & M) + (uint32_t)__popcll(v5 & M) + (uint32_t)__popcll(v6 & M) +
         (uint32_t)__popcll(v7 & M);
    const uint32_t F = (uint32_t)__popcll(a01 & a23 & M) + (uint32_t)__popcll(a45 & a67 & M) +
                       (uint32_t)__popcll(a01 & a45 & M) + (uint32_t)__popcll(a23 & a67 & M) +
                       (uint32_t)__popcll(a03 & a47 & M) + (uint32_t)__popcll(a12 & a56 & M);
    // body diagonals (0,6) (1,7) (2,4) (3,5): both ends set, and both ends clear
    const uint64_t pd0 = v0 & v6, pd1 = v1 & v7, pd2 = v2 & v4, pd3 = v3 & v5;
    const uint64_t p01 = pd0 & pd1, p23 = pd2 & pd3;
    const uint32_t S = (uint32_t)__popcll(~(v0 | v6) & pd1 & p23 & eb) + (uint32_t)__popcll(~(v1 | v7) & pd0 & p23 & eb) +
                       (uint32_t)__popcll(~(v2 | v4) & p01 & pd3 & eb) + (uint32_t)__popcll(~(v3 | v5) & p01 & pd2 & eb);
    return V - 2 * F - 4 * S;
}

/* ------------------------------------------------------------------ pass 1b + 2 */
/* per-row work beside the counts (none here; first-visit Marching Cubes adds its owned-edge
 * counts, k_fe_mc_count_rows): called by every lane with the lane's words, row and flags, and
 * `empty` when every word the wavefront loaded is zero */
struct NoRowExtra {
    __device__ void operator()(const LaneWords&, const RowLanes&, int64_t, bool, bool, bool, int64_t, int64_t,
                               bool) const
    {
    }
};

/* the 4 wavefronts of lane-row block blk (rows 4 rpw blk ..) */
template <class Extra = NoRowExtra>
__device__ __forceinline__ void fe_count_rows_block(const uint64_t* __restrict__ mask, const Geom& g,
                                                    RowCounts* __restrict__ counts, RowTrim* __restrict__ trims,
                                                    int64_t blk, Extra extra = Extra())
{
    const RowLanes rl = row_lanes(g.W);
    const int64_t r0 = (blk * 4 + wave_in_block()) * rl.rpw;
    const int64_t r = r0 + rl.q;
    const int64_t R = g.ny * g.nz;
    const bool ok = r < R && rl.w < g.W;
    const RowJK jk = split_row(r0, rl.q, g.ny);
    const int64_t j = jk.j, k = jk.k;
    const bool has_y = j < g.ny - 1, has_z = k < g.nz - 1;
    const LaneWords L = load_lane_words(mask, g, r, rl.w, ok, has_y, has_z);
    // the words and bit 0 of the next words (corner i + 1 of bit 63), as far as the lane loaded them
    const uint64_t any = L.m[0] | L.m[1] | L.m[2] | L.m[3] | ((L.n[0] | L.n[1] | L.n[2] | L.n[3]) & 1ull);
    if (__ballot(any != 0ull) == 0ull) {
        // every voxel the wavefront's rows see is below the isovalue (most wavefronts at a
        // sparse, high isovalue): no cut edge, no triangle, an empty trim
        if (r < R && rl.w == 0) {
            counts[r] = RowCounts{0u, 0u, 0u, 0u};
            trims[r] = RowTrim{(uint32_t)g.nx, 0u};
        }
        extra(L, rl, r, ok, has_y, has_z, j, k, true);
        return;
    }
    const uint64_t vb = low_bits(g.nx - 64 * rl.w), eb = low_bits(g.nx - 1 - 64 * rl.w);
    const uint64_t xcut = (L.m[0] ^ shifted(L.m[0], L.n[0])) & eb;
    const uint64_t ycut = has_y ? ((L.m[0] ^ L.m[1]) & vb) : 0;
    const uint64_t zcut = has_z ? ((L.m[0] ^ L.m[2]) & vb) : 0;
    const uint64_t trim = (xcut | ycut | zcut) & eb;
    // positions as unsigned: "none" is 0xFFFFFFFF for the minimum and 0 (= last + 1) for the maximum
    uint32_t first = trim ? (uint32_t)(64 * rl.w + __builtin_ctzll(trim)) : 0xFFFFFFFFu;
    uint32_t lastp1 = trim ? (uint32_t)(64 * rl.w + 64 - __builtin_clzll(trim)) : 0u;
    // word_tri_count is 0 where every corner of the lane's cells is below the isovalue (`any`
    // zero); a wavefront with only such lanes skips it
    const bool live = has_y && has_z && any != 0ull;
    uint32_t tc = 0u;
    if (__ballot(live) != 0ull) tc = (has_y && has_z) ? word_tri_count(corner_words(L), eb) : 0u;
    // two counts per 32-bit scan: row totals stay below 2^16 (nx <= 4096, <= 5 triangles a cell)
    const uint32_t xy = seg_sum((uint32_t)__popcll(xcut) | (uint32_t)__popcll(ycut) << 16, rl.Wp);
    const uint32_t zt = seg_sum((uint32_t)__popcll(zcut) | tc << 16, rl.Wp);
    const uint32_t xc = xy & 0xFFFFu, yc = xy >> 16, zc = zt & 0xFFFFu;
    tc = zt >> 16;
    first = seg_reduce<OpMin>(first, rl.Wp);
    lastp1 = seg_reduce<OpMax>(lastp1, rl.Wp);
    if (r < R && rl.w == 0) {
        counts[r] = RowCounts{xc, yc, zc, tc};
        trims[r] = RowTrim{lastp1 == 0 ? (uint32_t)g.nx : first, lastp1};
    }
    extra(L, rl, r, ok, has_y, has_z, j, k, false);
}

__global__ __launch_bounds__(256) void k_fe_count_rows(const uint64_t* __restrict__ mask, Geom g,
                                                       RowCounts* __restrict__ counts,
                                                       RowTrim* __restrict__ trims,
                                                       uint64_t* __restrict__ scan_state,
                                                       int64_t scan_state_words)
{
    // the scan of pass 3 runs next on the same stream; clear its tickets and tile status here
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < scan_state_words;
         t += (int64_t)gridDim.x * blockDim.x)
        scan_state[t] = 0;
    fe_count_rows_block(mask, g, counts, trims, blockIdx.x);
}

/* ------------------------------------------------------------------ pass 4: edge points */

/* One gradient component at the grid vertex *v, the reference's formula ((lower - upper) / run,
 * one-sided at the faces; flyingEdges/util/Image3D.cpp:116-188). With unit spacing the divisions
 * by 1 and 2 are exact and written as such. */
template <bool Unit>
__device__ __forceinline__ float grad_at(const float* __restrict__ v, int pos, int n, int stride, float s)
{
    const bool first = pos == 0, last = pos == n - 1;
    const float hi = v[last ? 0 : stride];
    const float lo = v[first ? 0 : -stride];
    const bool one_sided = first || last;
    if (Unit) return one_sided ? (lo - hi) : (lo - hi) * 0.5f;
    return (lo - hi) / (one_sided ? s : 2.0f * s);
}

template <bool Unit>
__device__ __forceinline__ float grad_interior(const float* __restrict__ v, int stride, float s)
{
    const float hi = v[stride], lo = v[-stride];
    if (Unit) return (lo - hi) * 0.5f;
    return (lo - hi) / (2.0f * s);
}

/* Point and normal of the crossing on the edge from grid vertex (i,j,k) along `axis`, with the
 * Flying Edges corner coordinates (fe_edge_corners). `v` points at the voxel of (i,j,k); strides
 * are 32-bit (the launch checks nx * ny < 2^31). Edges whose both ends are interior take a
 * branch without the face tests; it computes the same formulas. */
template <bool Unit>
__device__ __forceinline__ void fe_edge_point(const float* __restrict__ v, const Geom& g, float iso, int axis, int i,
                                              int j, int k, float p[3], float nrm[3])
{
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    const int sy = nx, sz = nx * ny;
    const int step = axis == 0 ? 1 : (axis == 1 ? sy : sz);
    const float* __restrict__ vb = v + step;
    const float v0 = v[0], v1 = vb[0];
    const float w = (iso - v0) / (v1 - v0);
    float A[3], B[3];
    if (Unit) {
        A[0] = (float)i; A[1] = (float)j; A[2] = (float)(k + (int)g.kofs);
        B[0] = A[0]; B[1] = A[1]; B[2] = A[2];
        B[axis] = A[axis] + 1.0f;
    } else {
        fe_edge_corners(g, axis, i, j, k, A, B);
    }
    const int ib = i + (axis == 0), jb = j + (axis == 1), kb = k + (axis == 2);
    float ga0, ga1, ga2, gb0, gb1, gb2;
    if (i >= 1 && j >= 1 && k >= 1 && ib <= nx - 2 && jb <= ny - 2 && kb <= nz - 2) {
        ga0 = grad_interior<Unit>(v, 1, g.sp[0]);
        ga1 = grad_interior<Unit>(v, sy, g.sp[1]);
        ga2 = grad_interior<Unit>(v, sz, g.sp[2]);
        gb0 = grad_interior<Unit>(vb, 1, g.sp[0]);
        gb1 = grad_interior<Unit>(vb, sy, g.sp[1]);
        gb2 = grad_interior<Unit>(vb, sz, g.sp[2]);
    } else {
        ga0 = grad_at<Unit>(v, i, nx, 1, g.sp[0]);
        ga1 = grad_at<Unit>(v, j, ny, sy, g.sp[1]);
        ga2 = grad_at<Unit>(v, k, nz, sz, g.sp[2]);
        gb0 = grad_at<Unit>(vb, ib, nx, 1, g.sp[0]);
        gb1 = grad_at<Unit>(vb, jb, ny, sy, g.sp[1]);
        gb2 = grad_at<Unit>(vb, kb, nz, sz, g.sp[2]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = A[c] + w * (B[c] - A[c]);
    nrm[0] = ga0 + w * (gb0 - ga0);
    nrm[1] = ga1 + w * (gb1 - ga1);
    nrm[2] = ga2 + w * (gb2 - ga2);
}

/* interior central difference (lower - upper) / (2 h), exact as * 0.5 for unit spacing */
template <bool Unit>
__device__ __forceinline__ float central(float lo, float hi, float s)
{
    if (Unit) return (lo - hi) * 0.5f;
    return (lo - hi) / (2.0f * s);
}

template <bool Unit>
__device__ __forceinline__ void edge_finish(const Geom& g, int axis, int i, int j, int k, float w, const float ga[3],
                                            const float gb[3], float p[3], float nrm[3])
{
    float A[3], B[3];
    if (Unit) {
        A[0] = (float)i; A[1] = (float)j; A[2] = (float)(k + (int)g.kofs);
        B[0] = A[0]; B[1] = A[1]; B[2] = A[2];
        B[axis] = A[axis] + 1.0f;
    } else {
        fe_edge_corners(g, axis, i, j, k, A, B);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        p[c] = A[c] + w * (B[c] - A[c]);
        nrm[c] = ga[c] + w * (gb[c] - ga[c]);
    }
}

/* x-edge (i,j,k)-(i+1,j,k): the 12 voxels in 5 loads (the row i-1..i+2, and the pairs i..i+1 in
 * rows j-1, j+1 and planes k-1, k+1); edges touching a face take fe_edge_point. */
template <bool Unit>
__device__ __forceinline__ void vertex_x(const float* __restrict__ v, const Geom& g, float iso, int i, int j, int k,
                                         float p[3], float nrm[3])
{
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    if (!(i >= 1 && i <= nx - 3 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2)) {
        fe_edge_point<Unit>(v, g, iso, 0, i, j, k, p, nrm);
        return;
    }
    const int sy = nx, sz = nx * ny;
    const Vox4 r = load_vox<Vox4>(v - 1);
    const Vox2 ym = load_vox<Vox2>(v - sy), yp = load_vox<Vox2>(v + sy);
    const Vox2 zm = load_vox<Vox2>(v - sz), zp = load_vox<Vox2>(v + sz);
    const float w = (iso - r.y) / (r.z - r.y);
    const float ga[3] = {central<Unit>(r.x, r.z, g.sp[0]), central<Unit>(ym.x, yp.x, g.sp[1]),
                         central<Unit>(zm.x, zp.x, g.sp[2])};
    const float gb[3] = {central<Unit>(r.y, r.w, g.sp[0]), central<Unit>(ym.y, yp.y, g.sp[1]),
                         central<Unit>(zm.y, zp.y, g.sp[2])};
    edge_finish<Unit>(g, 0, i, j, k, w, ga, gb, p, nrm);
}

/* y- or z-edge from (i,j,k) along `axis` (1 or 2): the voxels i-1..i+1 of both end rows in two
 * loads, the two further voxels along the edge, and the four across it (axis c, the third). */
template <bool Unit>
__device__ __forceinline__ void vertex_yz(const float* __restrict__ v, const Geom& g, float iso, int axis, int i,
                                          int j, int k, float p[3], float nrm[3])
{
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    const bool ay = axis == 1;
    const int pa = ay ? j : k, na = ay ? ny : nz, pc = ay ? k : j, nc = ay ? nz : ny;
    if (!(i >= 1 && i <= nx - 2 && pa >= 1 && pa <= na - 3 && pc >= 1 && pc <= nc - 2)) {
        fe_edge_point<Unit>(v, g, iso, axis, i, j, k, p, nrm);
        return;
    }
    const int sy = nx, sz = nx * ny;
    const int sa = ay ? sy : sz, sc = ay ? sz : sy;
    const float* __restrict__ vb = v + sa;
    const Vox3 xa = load_vox<Vox3>(v - 1), xb = load_vox<Vox3>(vb - 1);
    const float am = v[-sa], bp = vb[sa];
    const float cam = v[-sc], cap = v[sc], cbm = vb[-sc], cbp = vb[sc];
    const float w = (iso - xa.y) / (xb.y - xa.y);
    const float spa = ay ? g.sp[1] : g.sp[2], spc = ay ? g.sp[2] : g.sp[1];
    const float gax = central<Unit>(xa.x, xa.z, g.sp[0]), gbx = central<Unit>(xb.x, xb.z, g.sp[0]);
    const float gaa = central<Unit>(am, xb.y, spa), gba = central<Unit>(xa.y, bp, spa);
    const float gac = central<Unit>(cam, cap, spc), gbc = central<Unit>(cbm, cbp, spc);
    const float ga[3] = {gax, ay ? gaa : gac, ay ? gac : gaa};
    const float gb[3] = {gbx, ay ? gba : gbc, ay ? gbc : gba};
    edge_finish<Unit>(g, axis, i, j, k, w, ga, gb, p, nrm);
}

/* A vertex-list entry: output position (wavefront-relative) << 14 | axis << 12 | lane << 6 | bit,
 * the lane (q, w) and bit b naming row q of the wavefront and i = 64 w + b. The row is row j0 + q
 * of plane k0, carried into the following planes. */
template <bool Unit>
__device__ __forceinline__ void emit_vertex(const float* __restrict__ plane, const Geom& g, float iso, uint32_t ent,
                                            int j0, int k0, int wp_log2, bool xgroup, float p[3], float nrm[3])
{
    const int ny = (int)g.ny;
    const int src = (int)((ent >> 6) & 63);
    int j = j0 + (src >> wp_log2), dk = 0;
    while (j >= ny) {
        j -= ny;
        ++dk;
    }
    const int i = ((src & ((1 << wp_log2) - 1)) << 6) | (int)(ent & 63);
    const uint32_t off = (uint32_t)dk * (uint32_t)(g.nx * g.ny) + (uint32_t)j * (uint32_t)g.nx + (uint32_t)i;
    if (xgroup)
        vertex_x<Unit>(plane + off, g, iso, i, j, k0 + dk, p, nrm);
    else
        vertex_yz<Unit>(plane + off, g, iso, (int)((ent >> 12) & 3), i, j, k0 + dk, p, nrm);
}

/* ------------------------------------------------------------------ pass 4: fused emit */
constexpr int kEmitVertCap = 512;   // vertex-list entries per wavefront and round
constexpr int kEmitCellCap = 512;   // active-cell entries per wavefront and round
constexpr int kEmitTriCap = 64 * 5; // triangles of one group of 64 cells (at most 5 each)
constexpr int kStageVerts = 320;    // vertices staged per wavefront and output window

/* LDS of one wavefront. The triangles use the 8 edge-combination cut words and id bases of each
 * lane, the cell list and the triangle map; the vertices before them use the same bytes for their
 * list and staging window (sv), or, without staging, the vertex list (u.vlist). */
struct EmitRowsLds {
    union {
        struct {
            uint64_t cut[8][64];
            uint32_t base[8][64];
            union {
                uint32_t vlist[kEmitVertCap];
                struct {
                    uint32_t cells[kEmitCellCap];
                    uint32_t map[kEmitTriCap];
                } t;
            } u;
        };
        struct {
            uint32_t list[kStageVerts];
            float stage[6 * kStageVerts];   // points, then normals, of the window's vertices
        } sv;
    };
};
static_assert(sizeof(EmitRowsLds) == 8 * 64 * 8 + 8 * 64 * 4 + (kEmitCellCap + kEmitTriCap) * 4,
              "the staging window fits in the triangles' bytes");

/* case of the cell at bit b of a 32-bit half: h[q] holds corner v_q's bits (the case's bit order) */
__device__ __forceinline__ uint32_t case_of_half(const uint32_t h[8], int b)
{
    uint32_t c = __builtin_amdgcn_ubfe(h[0], b, 1);
#pragma unroll
    for (int q = 1; q < 8; ++q) c |= __builtin_amdgcn_ubfe(h[q], b, 1) << q;
    return c;
}

/* Pass 4 for the (row, word) lane layout in one launch: each wavefront writes the vertices of its
 * rows (one vertex per lane from LDS lists), then the triangles of its cell rows (active cells
 * listed per lane, resolved one cell and then one triangle per lane). */

/* Per-wavefront state of pass 4 in the lane-row layout: rows r0 .. r0 + rpw - 1, lane (q, w)
 * holding word w of row r0 + q and the words of its three neighbour rows. */
struct EmitRows {
    RowLanes rl;
    int lane, wv;
    int64_t r0, r;
    bool ok, emit, has_y, has_z;   // emit: the row is in the range pass 4 writes
    LaneWords L;
    uint64_t v0, t0, vb, eb;
    int64_t k0;   // the wavefront's first row is (j0, k0)
    int j0;
};

__device__ __forceinline__ bool emit_rows_setup(EmitRows& E, const uint64_t* __restrict__ mask, const Geom& g,
                                                const RowCounts* __restrict__ counts,
                                                const uint64_t* __restrict__ voff, const uint64_t* __restrict__ toff,
                                                int64_t blk)
{
    E.lane = lane_id();
    E.wv = wave_in_block();
    E.rl = row_lanes(g.W);
    const int64_t R = g.ny * g.nz;
    E.r0 = (blk * 4 + E.wv) * E.rl.rpw;
    if (E.r0 >= R) return false;
    E.r = E.r0 + E.rl.q;
    E.ok = E.r < R && E.rl.w < g.W;
    {
        // a wavefront whose rows have no vertex and no triangle leaves before reading mask words
        // (most of them at a sparse isovalue)
        bool some = false;
        if (E.ok && E.rl.w == 0) {
            const RowCounts c = counts[E.r];
            some = (c.x | c.y | c.z | c.t) != 0u;
        }
        if (__ballot(some) == 0ull) return false;
    }
    E.emit = E.ok && E.r >= g.emit_lo && E.r < g.emit_hi;
    const RowJK jk = split_row(E.r0, E.rl.q, g.ny);
    E.has_y = jk.j < g.ny - 1;
    E.has_z = jk.k < g.nz - 1;
    E.L = load_lane_words(mask, g, E.r, E.rl.w, E.ok, E.has_y, E.has_z);
    E.v0 = voff[E.r0];
    E.t0 = toff[E.r0];
    E.vb = low_bits(g.nx - 64 * E.rl.w);
    E.eb = low_bits(g.nx - 1 - 64 * E.rl.w);
    E.k0 = uniform_div(E.r0, g.ny);
    E.j0 = (int)(E.r0 - E.k0 * g.ny);
    return true;
}

/* Vertices of the wavefront's rows, [x][y][z] cuts per row, through an LDS list of kEmitVertCap
 * entries. Voxel offsets are 32-bit from plane k0. With vmap, vertex id v is written at vmap[v]. */
template <bool Unit, typename Index = uint32_t>
__device__ __forceinline__ void emit_rows_vertices(const EmitRows& E, uint32_t* list, const float* __restrict__ vol,
                                                   const Geom& g, float iso, float* __restrict__ pts,
                                                   float* __restrict__ nrm, const Index* __restrict__ vmap = nullptr)
{
    const RowLanes& rl = E.rl;
    const LaneWords& L = E.L;
    const int lane = E.lane;
    const uint64_t v0 = E.v0, vb = E.vb, eb = E.eb;
    const bool has_y = E.has_y, has_z = E.has_z;
    const int j0 = E.j0;
    const int64_t k0 = E.k0;
    const float* __restrict__ plane = vol + k0 * (g.nx * g.ny);
    const uint64_t live = E.emit ? ~0ull : 0ull;
    const uint64_t cut[3] = {(L.m[0] ^ shifted(L.m[0], L.n[0])) & eb & live,
                             has_y ? ((L.m[0] ^ L.m[1]) & vb & live) : 0ull,
                             has_z ? ((L.m[0] ^ L.m[2]) & vb & live) : 0ull};
    const uint32_t c0 = __popcll(cut[0]), c1 = __popcll(cut[1]), c2 = __popcll(cut[2]);
    // x and y counts share one segmented scan (row totals below 2^16)
    const uint32_t p01 = c0 | c1 << 16;
    const uint32_t inc01 = seg_incl<OpAdd>(p01, rl.Wp, rl.w);
    const uint32_t tot01 = (uint32_t)__shfl((int)inc01, (lane & ~(rl.Wp - 1)) + rl.Wp - 1);
    const uint32_t ex01 = inc01 - p01;
    const uint32_t X = tot01 & 0xFFFFu, Y = tot01 >> 16;
    const uint32_t lane_tot = c0 + c1 + c2;
    const uint32_t incl = wave_incl_scan_u32(lane_tot);
    const uint32_t before_row = __shfl(incl - lane_tot, rl.q * rl.Wp);
    const uint32_t pos0 = before_row + (ex01 & 0xFFFFu);
    const uint32_t pos1 = before_row + X + (ex01 >> 16);
    const uint32_t pos2 = before_row + X + Y + seg_excl_scan(c2, rl.Wp, rl.w);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    // Two lists, x-edges and then y/z-edges, so that the lanes of one step take the same load
    // pattern; each entry carries its output position. The x-edges are listed in lane order,
    // which is output order within each row, so that a step stores runs of consecutive vertices
    // (they mostly fit one step, whose gathers then touch the same lines in any order; 1024^3:
    // 1.594 against 1.615 ms). For rows of 16 words and more, the y/z-edges run word-major (word
    // w of all the wavefront's rows, then word w + 1), so that each of their steps gathers from a
    // narrow x-range of neighbouring rows and shares cache lines (1024^3: 1.65 against 1.69 ms);
    // for shorter rows in output order, so that one step's stores are contiguous (512^3 gyroid:
    // 0.425 against 0.44 ms; 256^3 and 384^3 likewise).
    const bool word_major = rl.Wp >= 16;
    const int wp_log2 = __builtin_ctz((unsigned)rl.Wp), rpw_log2 = 6 - wp_log2;
    const int tl = word_major ? rl.w * rl.rpw + rl.q : lane;                  // y/z list position of this lane
    const int from = word_major ? ((lane & (rl.rpw - 1)) << wp_log2) + (lane >> rpw_log2) : lane;   // lane at position `lane`
    const uint32_t c12t = (uint32_t)__shfl((int)(c1 + c2), from);
    const uint32_t xincl = wave_incl_scan_u32(c0), yzincl_t = wave_incl_scan_u32(c12t);
    const uint32_t xtotal = (uint32_t)__builtin_amdgcn_readlane((int)xincl, 63);
    const uint32_t xfirst = xincl - c0;
    const uint32_t yzfirst = (uint32_t)__shfl((int)(yzincl_t - c12t), tl);
#pragma unroll
    for (int grp = 0; grp < 2; ++grp) {
        const uint32_t gtotal = grp ? total - xtotal : xtotal;
        const uint32_t gfirst = grp ? yzfirst : xfirst;
        for (uint32_t win = 0; win < gtotal; win += kEmitVertCap) {
            const uint32_t wend = win + kEmitVertCap;
            uint32_t idx = gfirst;
#pragma unroll
            for (int a = grp; a < (grp ? 3 : 1); ++a) {
                uint64_t bits = cut[a];
                uint32_t p = a == 0 ? pos0 : (a == 1 ? pos1 : pos2);
                if (idx + (uint32_t)__popcll(bits) <= win || idx >= wend) {
                    idx += (uint32_t)__popcll(bits);
                    continue;
                }
                while (bits && idx < wend) {
                    const int b = __builtin_ctzll(bits);
                    bits &= bits - 1;
                    if (idx >= win) list[idx - win] = p << 14 | (uint32_t)a << 12 | (uint32_t)lane << 6 | (uint32_t)b;
                    ++idx;
                    ++p;
                }
                idx += (uint32_t)__popcll(bits);
            }
            wave_sync();
            const uint32_t n = min((uint32_t)kEmitVertCap, gtotal - win);
            for (uint32_t e = lane; e < n; e += 64) {
                const uint32_t ent = list[e];
                float pp[3], nn[3];
                emit_vertex<Unit>(plane, g, iso, ent, j0, (int)k0, wp_log2, grp == 0, pp, nn);
                uint64_t vid = v0 + (ent >> 14);
                if (vmap) vid = (uint64_t)vmap[vid];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    pts[3 * vid + c] = pp[c];
                    nrm[3 * vid + c] = nn[c];
                }
            }
            wave_sync();
        }
    }
}

/* Vertices of the wavefront's rows through an LDS staging window. The wavefront's vertices form
 * one contiguous output range [v0, v0 + total), written kStageVerts vertices at a time: in each
 * window the lanes list their cut edges whose output position falls in it (x-edges in lane order,
 * y/z-edges word-major for rows of 16 words and more, the order in which the gathers share cache
 * lines best), compute one vertex each, and put its point and normal at the vertex's place in the
 * staging buffer; the window then goes out as consecutive dwords, every store of the wavefront
 * covering 256 contiguous bytes. Stored straight from the lanes, a wavefront's vertices land in
 * runs of a few vertices (one row's cut edges of one axis, or fewer in word-major order), and
 * cache lines at the runs' ends are written partially, several times over. */
template <bool Unit>
__device__ __forceinline__ void emit_rows_vertices_staged(const EmitRows& E, EmitRowsLds& S,
                                                          const float* __restrict__ vol, const Geom& g, float iso,
                                                          float* __restrict__ pts, float* __restrict__ nrm)
{
    const RowLanes& rl = E.rl;
    const LaneWords& L = E.L;
    const int lane = E.lane;
    const uint64_t v0 = E.v0, vb = E.vb, eb = E.eb;
    const int j0 = E.j0;
    const int64_t k0 = E.k0;
    const float* __restrict__ plane = vol + k0 * (g.nx * g.ny);
    const uint64_t live = E.emit ? ~0ull : 0ull;
    const uint64_t cut[3] = {(L.m[0] ^ shifted(L.m[0], L.n[0])) & eb & live,
                             E.has_y ? ((L.m[0] ^ L.m[1]) & vb & live) : 0ull,
                             E.has_z ? ((L.m[0] ^ L.m[2]) & vb & live) : 0ull};
    const uint32_t c0 = __popcll(cut[0]), c1 = __popcll(cut[1]), c2 = __popcll(cut[2]);
    // output positions (wavefront-relative) of the lane's first x-, y- and z-edge vertex: rows in
    // order, [x][y][z] within a row, words in order within an axis
    const uint32_t p01 = c0 | c1 << 16;
    const uint32_t inc01 = seg_incl<OpAdd>(p01, rl.Wp, rl.w);
    const uint32_t tot01 = (uint32_t)__shfl((int)inc01, (lane & ~(rl.Wp - 1)) + rl.Wp - 1);
    const uint32_t ex01 = inc01 - p01;
    const uint32_t X = tot01 & 0xFFFFu, Y = tot01 >> 16;
    const uint32_t lane_tot = c0 + c1 + c2;
    const uint32_t incl = wave_incl_scan_u32(lane_tot);
    const uint32_t before_row = __shfl(incl - lane_tot, rl.q * rl.Wp);
    const uint32_t pos[3] = {before_row + (ex01 & 0xFFFFu), before_row + X + (ex01 >> 16),
                             before_row + X + Y + seg_excl_scan(c2, rl.Wp, rl.w)};
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    const bool word_major = rl.Wp >= 16;
    const int wp_log2 = __builtin_ctz((unsigned)rl.Wp), rpw_log2 = 6 - wp_log2;
    const int tl = word_major ? rl.w * rl.rpw + rl.q : lane;
    const int from = word_major ? ((lane & (rl.rpw - 1)) << wp_log2) + (lane >> rpw_log2) : lane;
    uint32_t* list = S.sv.list;
    float* stage = S.sv.stage;
    for (uint32_t w0 = 0; w0 < total; w0 += kStageVerts) {
        const uint32_t w1 = min(w0 + (uint32_t)kStageVerts, total);
        uint32_t n[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint32_t lo = max(pos[a], w0), hi = min(pos[a] + (a == 0 ? c0 : (a == 1 ? c1 : c2)), w1);
            n[a] = hi > lo ? hi - lo : 0u;
        }
        // list slots: the window's x-edges [0, xtot) in lane order, then its y/z-edges
        const uint32_t xincl = wave_incl_scan_u32(n[0]);
        const uint32_t xtot = (uint32_t)__builtin_amdgcn_readlane((int)xincl, 63);
        const uint32_t n12t = (uint32_t)__shfl((int)(n[1] + n[2]), from);
        const uint32_t yzincl_t = wave_incl_scan_u32(n12t);
        const uint32_t wn = (uint32_t)__builtin_amdgcn_readlane((int)yzincl_t, 63) + xtot;
        uint32_t slot[3];
        slot[0] = xincl - n[0];
        slot[1] = xtot + (uint32_t)__shfl((int)(yzincl_t - n12t), tl);
        slot[2] = slot[1] + n[1];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!n[a]) continue;
            uint64_t bits = cut[a];
            uint32_t p = pos[a], s = slot[a];
            while (bits && p < w1) {
                const int b = __builtin_ctzll(bits);
                bits &= bits - 1;
                if (p >= w0) list[s++] = (p - w0) << 14 | (uint32_t)a << 12 | (uint32_t)lane << 6 | (uint32_t)b;
                ++p;
            }
        }
        wave_sync();
#pragma unroll
        for (int grp = 0; grp < 2; ++grp) {
            for (uint32_t e = (grp ? xtot : 0u) + (uint32_t)lane; e < (grp ? wn : xtot); e += 64) {
                const uint32_t ent = list[e];
                float pp[3], nn[3];
                emit_vertex<Unit>(plane, g, iso, ent, j0, (int)k0, wp_log2, grp == 0, pp, nn);
                const uint32_t at = 3 * (ent >> 14);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    stage[at + c] = pp[c];
                    stage[3 * kStageVerts + at + c] = nn[c];
                }
            }
        }
        wave_sync();
        const uint32_t nf = 3 * (w1 - w0);
        float* __restrict__ dp = pts + 3 * (v0 + w0);
        float* __restrict__ dn = nrm + 3 * (v0 + w0);
        for (uint32_t f = (uint32_t)lane; f < nf; f += 64) {
            dp[f] = stage[f];
            dn[f] = stage[3 * kStageVerts + f];
        }
        wave_sync();
    }
}

/* Triangles of the wavefront's cell rows: active cells listed per lane in output order, then
 * resolved one triangle per lane. Ids are written less vbias, or through remap. */
template <typename Index, bool Remap>
__device__ __forceinline__ void emit_rows_triangles(const EmitRows& E, EmitRowsLds& S, const uint64_t* __restrict__ s_case,
                                                    const Geom& g, const RowCounts* __restrict__ counts,
                                                    const uint64_t* __restrict__ voff, Index* __restrict__ tris,
                                                    uint64_t vbias, const Index* __restrict__ remap)
{
    const RowLanes& rl = E.rl;
    const LaneWords& L = E.L;
    const int lane = E.lane;
    const int64_t r = E.r;
    const bool ok = E.ok, has_y = E.has_y, has_z = E.has_z;
    const uint64_t v0 = E.v0, t0 = E.t0, vb = E.vb, eb = E.eb;
    const bool cells = ok && E.emit && has_y && has_z;
    const CornerWords C = corner_words(L);
    const uint64_t act = cells ? active_cells(C, eb) : 0ull;
    const uint32_t ncell = (uint32_t)__popcll(act);
    const uint32_t cincl = wave_incl_scan_u32(ncell);
    const uint32_t ctotal = (uint32_t)__builtin_amdgcn_readlane((int)cincl, 63);
    if (ctotal == 0) return;
    {
        uint64_t cw[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) cw[q] = (L.m[q] ^ shifted(L.m[q], L.n[q])) & eb;
        cw[4] = (L.m[0] ^ L.m[1]) & vb;
        cw[5] = (L.m[2] ^ L.m[3]) & vb;
        cw[6] = (L.m[0] ^ L.m[2]) & vb;
        cw[7] = (L.m[1] ^ L.m[3]) & vb;
        uint64_t cb[8] = {v0, v0, v0, v0, v0, v0, v0, v0};
        if (act) {
            const int64_t rows[4] = {r, r + 1, r + g.ny, r + g.ny + 1};
            combo_bases(counts, voff, rows, cb);
        }
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            // two combinations per segmented scan (row totals below 2^16)
            const uint32_t ex = seg_excl_scan((uint32_t)__popcll(cw[q]) | (uint32_t)__popcll(cw[q + 1]) << 16, rl.Wp,
                                              rl.w);
            S.cut[q][lane] = cw[q];
            S.cut[q + 1][lane] = cw[q + 1];
            S.base[q][lane] = (uint32_t)(cb[q] - v0) + (ex & 0xFFFFu);
            S.base[q + 1][lane] = (uint32_t)(cb[q + 1] - v0) + (ex >> 16);
        }
    }
    uint32_t lo[8], hi[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        lo[2 * q] = (uint32_t)C.m[q];
        hi[2 * q] = (uint32_t)(C.m[q] >> 32);
        lo[2 * q + 1] = (uint32_t)C.s[q];
        hi[2 * q + 1] = (uint32_t)(C.s[q] >> 32);
    }
    // corner order v0..v7 = m0 s0 s1 m1 m2 s2 s3 m3
    const uint32_t clo[8] = {lo[0], lo[1], lo[3], lo[2], lo[4], lo[5], lo[7], lo[6]};
    const uint32_t chi[8] = {hi[0], hi[1], hi[3], hi[2], hi[4], hi[5], hi[7], hi[6]};
    const uint32_t cfirst = cincl - ncell;
    const uint64_t vrel = v0 - vbias;
    uint64_t slot = t0;
    for (uint32_t win = 0; win < ctotal; win += kEmitCellCap) {
        const uint32_t wend = win + kEmitCellCap;
        // each lane lists the positions (lane, bit) of its active cells
        if (ncell && cfirst < wend && cfirst + ncell > win) {
            uint32_t p = cfirst;
            for (uint64_t a = act; a && p < wend;) {
                const int b = __builtin_ctzll(a);
                a &= a - 1;
                if (p >= win) S.u.t.cells[p - win] = (uint32_t)lane | (uint32_t)b << 6;
                ++p;
            }
        }
        wave_sync();
        const uint32_t n = min((uint32_t)kEmitCellCap, ctotal - win);
        for (uint32_t g0 = 0; g0 < n; g0 += 64) {
            // one cell per lane: its case from the owner lane's corner words (read across lanes)
            const uint32_t idx = g0 + (uint32_t)lane;
            uint32_t ent = idx < n ? S.u.t.cells[idx] : 0u;
            {
                const int src = (int)(ent & 63), b = (int)((ent >> 6) & 63);
                uint32_t half[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const uint32_t l = (uint32_t)__shfl((int)clo[q], src), h = (uint32_t)__shfl((int)chi[q], src);
                    half[q] = b >= 32 ? h : l;
                }
                ent |= case_of_half(half, b & 31) << 12;
            }
            const uint32_t nt = idx < n ? (uint32_t)case_tri_count(s_case[ent >> 12]) : 0u;
            const uint32_t tincl = wave_incl_scan_u32(nt);
            const uint32_t ttot = (uint32_t)__builtin_amdgcn_readlane((int)tincl, 63);
            for (uint32_t tt = 0; tt < nt; ++tt) S.u.t.map[tincl - nt + tt] = ent | tt << 20;
            wave_sync();
            for (uint32_t e = lane; e < ttot; e += 64) {
                const uint32_t me = S.u.t.map[e];
                const int owner = (int)(me & 63);
                const int b = (int)((me >> 6) & 63);
                const uint64_t packed = s_case[(me >> 12) & 0xFF];
                const int tt = (int)(me >> 20);
                uint32_t id[3];
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    const int ed = case_edge(packed, 3 * tt + s);
                    const int q = combo_of_edge(ed);
                    const int p = b + edge_dx(ed);
                    const uint64_t cwq = S.cut[q][owner];
                    const uint64_t below = p == 64 ? cwq : (cwq & ((1ull << p) - 1ull));
                    id[s] = S.base[q][owner] + (uint32_t)__popcll(below);
                }
                if (Remap)
                    store_tri<Index>(tris, slot + e, remap[vrel + id[0]], remap[vrel + id[1]], remap[vrel + id[2]]);
                else
                    store_tri<Index>(tris, slot + e, vrel + id[0], vrel + id[1], vrel + id[2]);
            }
            slot += ttot;
            wave_sync();
        }
    }
}

/* Lane-row block blk. s_case holds kCaseEdges; vertex ids are written less vbias (the first
 * vertex of this volume's output when several volumes share one buffer). With Remap, only the
 * triangles are written, each Flying Edges vertex id replaced by remap[id]. */
template <typename Index, bool Unit, bool Remap = false>
__device__ __forceinline__ void fe_emit_rows_block(const uint64_t* __restrict__ s_case, EmitRowsLds* s_w,
                                                   const float* __restrict__ vol, const uint64_t* __restrict__ mask,
                                                   const Geom& g, float iso, const RowCounts* __restrict__ counts,
                                                   const uint64_t* __restrict__ voff,
                                                   const uint64_t* __restrict__ toff, float* __restrict__ pts,
                                                   float* __restrict__ nrm, Index* __restrict__ tris,
                                                   uint64_t vbias, int64_t blk, const Index* __restrict__ remap = nullptr)
{
    EmitRows E;
    if (!emit_rows_setup(E, mask, g, counts, voff, toff, blk)) return;
    if (!Remap) emit_rows_vertices_staged<Unit>(E, s_w[E.wv], vol, g, iso, pts, nrm);
    emit_rows_triangles<Index, Remap>(E, s_w[E.wv], s_case, g, counts, voff, tris, vbias, remap);
}

/* The order in which pass 4's blocks take the rows, across the 8 XCDs. The hardware hands blocks
 * to the XCDs round-robin (block b to XCD b mod 8); this gives each XCD tiles of TG consecutive
 * row blocks (about 64 grid rows) by TP = 3 planes instead of every eighth block, so that the
 * rows around a wavefront's own that its vertices read (j - 1 .. j + 2, k - 1 .. k + 2) are
 * mostly in the same XCD's L2, while all XCDs still move through the planes together (they share
 * the MALL). At 1024^3 pass 4 takes 1.63 ms in this order against 1.84 ms in dispatch order.
 * Identity when the blocks do not split the planes evenly. A bijection on [0, n) otherwise:
 * block b is the i-th (i = b / 8) of XCD x = b % 8, which takes tile t = i / (TG TP) of its
 * column of tiles, TP planes by TG blocks, planes first.
 * The plan (plan_block_order) is made on the host with the geometry, and a block maps its index
 * with shifts, multiplications and one 32-bit division. That matters at a sparse isovalue, where
 * most blocks write nothing: with the plan worked out in every block in 64-bit arithmetic, pass 4
 * at 1024^3, iso 0.95 took 0.15 ms against 0.08 ms. */
constexpr uint32_t kBlockOrderTP = 3;

inline BlockOrder plan_block_order(int64_t ny, int64_t nz, int64_t W)
{
    BlockOrder o{0u, 1u, 1u, 0u};
    if (W < 1 || W > 64 || ny >= (1ll << 30)) return o;
    int64_t wp = 1;
    while (wp < W) wp <<= 1;
    const int64_t per_block = 4 * (64 / wp);                // 4 wavefronts of lane rows: 4 .. 256
    const int64_t n = (ny * nz + per_block - 1) / per_block;
    if (ny % per_block || n > 0xFFFFFFFFll) return o;
    const int64_t G = ny / per_block;                       // blocks per plane
    int64_t TG = per_block >= 64 ? 1 : 64 / per_block;
    while (TG > 1 && G % (8 * TG)) TG >>= 1;
    if (G % (8 * TG)) return o;
    o.n0 = (uint32_t)(n / (G * kBlockOrderTP) * (G * kBlockOrderTP));   // whole tile rows
    o.G = (uint32_t)G;
    o.U = (uint32_t)(G / (8 * TG));
    o.lg_tg = (uint32_t)__builtin_ctzll((unsigned long long)TG);
    return o;
}

__host__ __device__ __forceinline__ uint32_t apply_block_order(const BlockOrder& o, uint32_t b)
{
    constexpr uint32_t TP = kBlockOrderTP;
    const uint32_t TG = 1u << o.lg_tg;
    const uint32_t x = b & 7u, i = b >> 3;
    const uint32_t q = (i >> o.lg_tg) / TP;                 // i / (TG TP): the XCD's tile
    const uint32_t t = i - q * (TG * TP), c = q / o.U, u = q - c * o.U;
    return (c * TP + t % TP) * o.G + (x + 8 * u) * TG + t / TP;
}

__device__ __forceinline__ int64_t emit_block_order(int64_t b, int64_t n, const Geom& g)
{
    // a bijection on [0, n0) and the identity above it, so also on any grid of at least n0 blocks
    if (b >= (int64_t)g.bo.n0 || n < (int64_t)g.bo.n0) return b;
    return apply_block_order(g.bo, (uint32_t)b);
}

template <typename Index, bool Unit>
__global__ __launch_bounds__(256) void k_fe_emit_rows(const float* __restrict__ vol,
                                                      const uint64_t* __restrict__ mask, Geom g, float iso,
                                                      const RowCounts* __restrict__ counts,
                                                      const uint64_t* __restrict__ voff,
                                                      const uint64_t* __restrict__ toff,
                                                      float* __restrict__ pts, float* __restrict__ nrm,
                                                      Index* __restrict__ tris, uint64_t vbias = 0,
                                                      const uint64_t* __restrict__ need = nullptr,
                                                      uint64_t cap_v = 0, uint64_t cap_t = 0)
{
    // with need (pass 3's totals, on the device), write nothing unless they fit the buffers and
    // the index width (the host then grows the buffers and runs this pass again)
    if (need && !output_fits<Index>(need, cap_v, cap_t)) return;
    __shared__ uint64_t s_case[256];
    __shared__ EmitRowsLds s_w[4];
    s_case[threadIdx.x] = kCaseEdges[threadIdx.x];
    __syncthreads();
    const int64_t blk = emit_block_order(blockIdx.x, gridDim.x, g);
    fe_emit_rows_block<Index, Unit>(s_case, s_w, vol, mask, g, iso, counts, voff, toff, pts, nrm, tris, vbias, blk);
}

/* Marching Cubes for unit spacing and zero origin, where a Marching Cubes vertex has the bits of
 * the Flying Edges vertex on the same edge: the Flying Edges pass 4 with every vertex written at
 * fe_to_mc[its id] and every triangle's ids read through the same map. */
template <typename Index>
__global__ __launch_bounds__(256) void k_mc_emit_rows(const float* __restrict__ vol,
                                                      const uint64_t* __restrict__ mask, Geom g, float iso,
                                                      const RowCounts* __restrict__ counts,
                                                      const uint64_t* __restrict__ voff,
                                                      const uint64_t* __restrict__ toff,
                                                      const Index* __restrict__ fe_to_mc, float* __restrict__ pts,
                                                      float* __restrict__ nrm, Index* __restrict__ tris)
{
    __shared__ uint64_t s_case[256];
    __shared__ EmitRowsLds s_w[4];
    s_case[threadIdx.x] = kCaseEdges[threadIdx.x];
    __syncthreads();
    const int64_t blk = emit_block_order(blockIdx.x, gridDim.x, g);
    EmitRows E;
    if (!emit_rows_setup(E, mask, g, counts, voff, toff, blk)) return;
    emit_rows_vertices<true, Index>(E, s_w[E.wv].u.vlist, vol, g, iso, pts, nrm, fe_to_mc);
    emit_rows_triangles<Index, true>(E, s_w[E.wv], s_case, g, counts, voff, tris, 0, fe_to_mc);
}

/* Marching Cubes triangles: the Flying Edges triangle stream through the id map. */
template <typename Index>
__global__ __launch_bounds__(256) void k_mc_triangles_rows(const uint64_t* __restrict__ mask, Geom g,
                                                           const RowCounts* __restrict__ counts,
                                                           const uint64_t* __restrict__ voff,
                                                           const uint64_t* __restrict__ toff,
                                                           const Index* __restrict__ fe_to_mc,
                                                           Index* __restrict__ tris)
{
    __shared__ uint64_t s_case[256];
    __shared__ EmitRowsLds s_w[4];
    s_case[threadIdx.x] = kCaseEdges[threadIdx.x];
    __syncthreads();
    const int64_t blk = emit_block_order(blockIdx.x, gridDim.x, g);
    fe_emit_rows_block<Index, true, true>(s_case, s_w, nullptr, mask, g, 0.0f, counts, voff, toff, nullptr, nullptr,
                                          tris, 0, blk, fe_to_mc);
}

}  // namespace miso

#endif
