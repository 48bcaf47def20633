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
 *   k_fe_emit_rows   pass 4: vertices (cut edges compacted into an LDS list in output order, one
 *                    vertex per lane, written at consecutive addresses), then triangles (active
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
    int wp = 1;
    while (wp < W) wp <<= 1;
    L.Wp = wp;
    L.rpw = 64 / wp;
    const int lane = lane_id();
    L.q = lane / wp;
    L.w = lane % wp;
    return L;
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
    const uint32_t inc = seg_incl<Op>(v, Wp, lane % Wp);
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

/* marching-cubes case of the cell at bit b: corner bits v0..v7 (FlyingEdgesAlgorithm.cpp:138-157) */
__device__ __forceinline__ int case_of(const CornerWords& C, int b)
{
    const int sh = b & 31;
    const bool hi = b >= 32;
    auto half = [&](uint64_t w) -> uint32_t { return hi ? (uint32_t)(w >> 32) : (uint32_t)w; };
    const uint32_t c = ((half(C.m[0]) >> sh) & 1u) | ((half(C.s[0]) >> sh) & 1u) << 1 |
                       ((half(C.s[1]) >> sh) & 1u) << 2 | ((half(C.m[1]) >> sh) & 1u) << 3 |
                       ((half(C.m[2]) >> sh) & 1u) << 4 | ((half(C.s[2]) >> sh) & 1u) << 5 |
                       ((half(C.s[3]) >> sh) & 1u) << 6 | ((half(C.m[3]) >> sh) & 1u) << 7;
    return (int)c;
}

__device__ __forceinline__ int case_at(const LaneWords& L, int b) { return case_of(corner_words(L), b); }

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

/* ------------------------------------------------------------------ pass 1b + 2 */
__global__ __launch_bounds__(256) void k_fe_count_rows(const uint64_t* __restrict__ mask, Geom g,
                                                       RowCounts* __restrict__ counts,
                                                       RowTrim* __restrict__ trims,
                                                       uint64_t* __restrict__ scan_state,
                                                       int64_t scan_state_words)
{
    __shared__ uint8_t s_ntri[256];
    for (int c = threadIdx.x; c < 256; c += blockDim.x) s_ntri[c] = (uint8_t)case_tri_count(kCaseEdges[c]);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < scan_state_words;
         t += (int64_t)gridDim.x * blockDim.x)
        scan_state[t] = 0;
    __syncthreads();

    const RowLanes rl = row_lanes(g.W);
    const int64_t wave_id = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t r = wave_id * rl.rpw + rl.q;
    const int64_t R = g.ny * g.nz;
    const bool ok = r < R && rl.w < g.W;
    const int64_t j = r % g.ny, k = r / g.ny;
    const bool has_y = j < g.ny - 1, has_z = k < g.nz - 1;
    const LaneWords L = load_lane_words(mask, g, r, rl.w, ok, has_y, has_z);
    const uint64_t vb = low_bits(g.nx - 64 * rl.w), eb = low_bits(g.nx - 1 - 64 * rl.w);
    const uint64_t xcut = (L.m[0] ^ shifted(L.m[0], L.n[0])) & eb;
    const uint64_t ycut = has_y ? ((L.m[0] ^ L.m[1]) & vb) : 0;
    const uint64_t zcut = has_z ? ((L.m[0] ^ L.m[2]) & vb) : 0;
    const uint64_t trim = (xcut | ycut | zcut) & eb;
    // positions as unsigned: "none" is 0xFFFFFFFF for the minimum and 0 (= last + 1) for the maximum
    uint32_t first = trim ? (uint32_t)(64 * rl.w + __builtin_ctzll(trim)) : 0xFFFFFFFFu;
    uint32_t lastp1 = trim ? (uint32_t)(64 * rl.w + 64 - __builtin_clzll(trim)) : 0u;
    uint32_t tc = 0;
    if (has_y && has_z) {
        const CornerWords C = corner_words(L);
        uint64_t act = active_cells(C, eb);
        while (act) {
            const int b = __builtin_ctzll(act);
            act &= act - 1;
            tc += s_ntri[case_of(C, b)];
        }
    }
    const uint32_t xc = seg_sum((uint32_t)__popcll(xcut), rl.Wp);
    const uint32_t yc = seg_sum((uint32_t)__popcll(ycut), rl.Wp);
    const uint32_t zc = seg_sum((uint32_t)__popcll(zcut), rl.Wp);
    tc = seg_sum(tc, rl.Wp);
    first = seg_reduce<OpMin>(first, rl.Wp);
    lastp1 = seg_reduce<OpMax>(lastp1, rl.Wp);
    if (r < R && rl.w == 0) {
        counts[r] = RowCounts{xc, yc, zc, tc};
        trims[r] = RowTrim{lastp1 == 0 ? (uint32_t)g.nx : first, lastp1};
    }
}

/* ------------------------------------------------------------------ pass 4: edge points */

/* One gradient component, the reference's formula ((lower - upper) / run, one-sided at the
 * faces; flyingEdges/util/Image3D.cpp:116-188). With unit spacing the divisions by 1 and 2 are
 * exact and written as such. */
template <bool Unit>
__device__ __forceinline__ float grad_axis_at(const float* __restrict__ d, int64_t idx, int pos, int n, int64_t stride,
                                              float s)
{
    float hi, lo;
    const bool first = pos == 0, last = pos == n - 1;
    hi = last ? d[idx] : d[idx + stride];
    lo = first ? d[idx] : d[idx - stride];
    const bool one_sided = first || last;
    if (Unit) return one_sided ? (lo - hi) : (lo - hi) * 0.5f;
    return (lo - hi) / (one_sided ? s : 2.0f * s);
}

/* Point and normal of the crossing on the edge from grid vertex (i,j,k) along `axis`, with the
 * Flying Edges corner coordinates (fe_edge_corners); 32-bit indices, one 64-bit base. */
template <bool Unit>
__device__ __forceinline__ void fe_edge_point(const float* __restrict__ d, const Geom& g, float iso, int axis, int i,
                                              int j, int k, float p[3], float nrm[3])
{
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
    const int64_t base = (int64_t)i + (int64_t)j * sy + (int64_t)k * sz;
    const int64_t step = axis == 0 ? sx : (axis == 1 ? sy : sz);
    const float v0 = d[base], v1 = d[base + step];
    const float w = (iso - v0) / (v1 - v0);
    float A[3], B[3];
    if (Unit) {
        A[0] = (float)i; A[1] = (float)j; A[2] = (float)k;
        B[0] = A[0]; B[1] = A[1]; B[2] = A[2];
        B[axis] = A[axis] + 1.0f;
    } else {
        fe_edge_corners(g, axis, i, j, k, A, B);
    }
    const int ib = i + (axis == 0), jb = j + (axis == 1), kb = k + (axis == 2);
    const int64_t bb = base + step;
    const float ga0 = grad_axis_at<Unit>(d, base, i, nx, sx, g.sp[0]);
    const float ga1 = grad_axis_at<Unit>(d, base, j, ny, sy, g.sp[1]);
    const float ga2 = grad_axis_at<Unit>(d, base, k, nz, sz, g.sp[2]);
    const float gb0 = grad_axis_at<Unit>(d, bb, ib, nx, sx, g.sp[0]);
    const float gb1 = grad_axis_at<Unit>(d, bb, jb, ny, sy, g.sp[1]);
    const float gb2 = grad_axis_at<Unit>(d, bb, kb, nz, sz, g.sp[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = A[c] + w * (B[c] - A[c]);
    nrm[0] = ga0 + w * (gb0 - ga0);
    nrm[1] = ga1 + w * (gb1 - ga1);
    nrm[2] = ga2 + w * (gb2 - ga2);
}

/* ------------------------------------------------------------------ pass 4: fused emit */
constexpr int kEmitVertCap = 512;   // vertex-list entries per wavefront and round
constexpr int kEmitCellCap = 512;   // active-cell entries per wavefront and round
constexpr int kEmitTriCap = 64 * 5; // triangles of one group of 64 cells (at most 5 each)

/* LDS of one wavefront: the vertex list, then (reusing it) the cell list and triangle map, beside
 * the 8 edge-combination cut words and id bases of each lane. */
struct EmitRowsLds {
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

/* case of the cell at bit b of a 32-bit half (corner bits v0..v7 as case_of) */
__device__ __forceinline__ uint32_t case_of_half(const uint32_t h[8], int b)
{
    uint32_t c = __builtin_amdgcn_ubfe(h[0], b, 1);
#pragma unroll
    for (int q = 1; q < 8; ++q) c |= __builtin_amdgcn_ubfe(h[q], b, 1) << q;
    return c;
}

/* Pass 4 for the (row, word) lane layout in one launch: each wavefront writes the vertices of its
 * rows (one vertex per lane from an LDS list in output order, two per lane per step so that the
 * gathers of both are in flight together), then the triangles of its cell rows (active cells
 * listed per lane, resolved one cell and then one triangle per lane). */
template <typename Index, bool Unit>
__global__ __launch_bounds__(256) void k_fe_emit_rows(const float* __restrict__ vol,
                                                      const uint64_t* __restrict__ mask, Geom g, float iso,
                                                      const RowCounts* __restrict__ counts,
                                                      const uint64_t* __restrict__ voff,
                                                      const uint64_t* __restrict__ toff,
                                                      float* __restrict__ pts, float* __restrict__ nrm,
                                                      Index* __restrict__ tris)
{
    __shared__ uint64_t s_case[256];
    __shared__ EmitRowsLds s_w[4];
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wv = t >> 6;
    s_case[t] = kCaseEdges[t];
    __syncthreads();
    EmitRowsLds& S = s_w[wv];
    const RowLanes rl = row_lanes(g.W);
    const int64_t R = g.ny * g.nz;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + wv) * rl.rpw;
    if (r0 >= R) return;
    const int64_t r = r0 + rl.q;
    const bool ok = r < R && rl.w < g.W;
    const int ny = (int)g.ny;
    const int j = (int)(r % ny), k = (int)(r / ny);
    const bool has_y = j < ny - 1, has_z = k < (int)g.nz - 1;
    const LaneWords L = load_lane_words(mask, g, r, rl.w, ok, has_y, has_z);
    const uint64_t v0 = voff[r0];
    const uint64_t t0 = toff[r0];
    const uint64_t vb = low_bits(g.nx - 64 * rl.w), eb = low_bits(g.nx - 1 - 64 * rl.w);

    // ---------------- vertices: rows r0 .. r0 + rpw - 1, [x][y][z] cuts per row
    {
        const uint64_t cut[3] = {(L.m[0] ^ shifted(L.m[0], L.n[0])) & eb, has_y ? ((L.m[0] ^ L.m[1]) & vb) : 0ull,
                                 has_z ? ((L.m[0] ^ L.m[2]) & vb) : 0ull};
        const uint32_t c0 = __popcll(cut[0]), c1 = __popcll(cut[1]), c2 = __popcll(cut[2]);
        const uint32_t X = seg_sum(c0, rl.Wp), Y = seg_sum(c1, rl.Wp);
        const uint32_t lane_tot = c0 + c1 + c2;
        const uint32_t incl = wave_incl_scan_u32(lane_tot);
        const uint32_t before_row = __shfl(incl - lane_tot, rl.q * rl.Wp);
        const uint32_t pos0 = before_row + seg_excl_scan(c0, rl.Wp, rl.w);
        const uint32_t pos1 = before_row + X + seg_excl_scan(c1, rl.Wp, rl.w);
        const uint32_t pos2 = before_row + X + Y + seg_excl_scan(c2, rl.Wp, rl.w);
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        uint32_t* list = S.u.vlist;
        for (uint32_t win = 0; win < total; win += kEmitVertCap) {
            const uint32_t wend = win + kEmitVertCap;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                uint64_t bits = cut[a];
                uint32_t p = a == 0 ? pos0 : (a == 1 ? pos1 : pos2);
                if (p + (uint32_t)__popcll(bits) <= win || p >= wend) continue;
                while (bits && p < wend) {
                    const int b = __builtin_ctzll(bits);
                    bits &= bits - 1;
                    if (p >= win) list[p - win] = ((uint32_t)rl.q << 20) | ((uint32_t)a << 18) | (uint32_t)(rl.w * 64 + b);
                    ++p;
                }
            }
            wave_sync();
            const uint32_t n = min((uint32_t)kEmitVertCap, total - win);
            for (uint32_t e = lane; e < n; e += 128) {
                const bool two = e + 64 < n;
                const uint32_t ent0 = list[e];
                const uint32_t ent1 = two ? list[e + 64] : ent0;
                float p0[3], n0[3], p1[3], n1[3];
                {
                    const int64_t rr = r0 + (ent0 >> 20);
                    fe_edge_point<Unit>(vol, g, iso, (int)((ent0 >> 18) & 3), (int)(ent0 & 0x3FFFF), (int)(rr % ny),
                                        (int)(rr / ny), p0, n0);
                }
                {
                    const int64_t rr = r0 + (ent1 >> 20);
                    fe_edge_point<Unit>(vol, g, iso, (int)((ent1 >> 18) & 3), (int)(ent1 & 0x3FFFF), (int)(rr % ny),
                                        (int)(rr / ny), p1, n1);
                }
                const uint64_t vid = v0 + win + e;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    pts[3 * vid + c] = p0[c];
                    nrm[3 * vid + c] = n0[c];
                }
                if (two) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        pts[3 * (vid + 64) + c] = p1[c];
                        nrm[3 * (vid + 64) + c] = n1[c];
                    }
                }
            }
            wave_sync();
        }
    }

    // ---------------- triangles: cell rows r0 .. r0 + rpw - 1
    const bool cells = ok && has_y && has_z;
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
        for (int q = 0; q < 8; ++q) {
            S.cut[q][lane] = cw[q];
            S.base[q][lane] = (uint32_t)(cb[q] - v0) + seg_excl_scan((uint32_t)__popcll(cw[q]), rl.Wp, rl.w);
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
    uint64_t slot = t0;
    for (uint32_t win = 0; win < ctotal; win += kEmitCellCap) {
        const uint32_t wend = win + kEmitCellCap;
        if (ncell && cfirst < wend && cfirst + ncell > win) {
            uint32_t p = cfirst;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint32_t a = h ? (uint32_t)(act >> 32) : (uint32_t)act;
                while (a && p < wend) {
                    const int b = __builtin_ctz(a);
                    a &= a - 1;
                    if (p >= win) {
                        const uint32_t cs = case_of_half(h ? chi : clo, b);
                        S.u.t.cells[p - win] = (uint32_t)lane | (uint32_t)(b + 32 * h) << 6 | cs << 12;
                    }
                    ++p;
                }
            }
        }
        wave_sync();
        const uint32_t n = min((uint32_t)kEmitCellCap, ctotal - win);
        for (uint32_t g0 = 0; g0 < n; g0 += 64) {
            const uint32_t idx = g0 + (uint32_t)lane;
            const uint32_t ent = idx < n ? S.u.t.cells[idx] : 0u;
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
                store_tri<Index>(tris, slot + e, v0 + id[0], v0 + id[1], v0 + id[2]);
            }
            slot += ttot;
            wave_sync();
        }
    }
}

}  // namespace miso

#endif
