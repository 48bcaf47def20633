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
 *   k_fe_vertices    pass 4, vertices: cut edges compacted into an LDS list in output order,
 *                    one vertex per lane, written at consecutive addresses
 *   k_fe_triangles   pass 4, triangles: per active cell, case-table order, vertex ids from row
 *                    offsets plus popcount ranks held in LDS
 */
#ifndef MINIISO_FE_ROWS_H
#define MINIISO_FE_ROWS_H

#include "fe_kernels.h"

namespace miso {

constexpr int kListCap = 1024;      // vertex-list entries per wavefront and round (4 KB of LDS)
constexpr int kXcds = 8;            // MI355X: 8 accelerator complex dies, each with its own L2

/* Blocks are dispatched round-robin over the XCDs (block b runs on XCD b % 8). Renumber them so
 * that each XCD walks one contiguous 1/8 of the rows: the neighbour rows a block reads (j+1 and
 * the next plane) were then brought into the same XCD's L2 by the block just before it. */
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb)
{
    const int64_t per = nb / kXcds, extra = nb % kXcds;
    const int64_t x = b % kXcds, i = b / kXcds;
    return x * per + min(x, extra) + i;
}

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

/* ------------------------------------------------------------------ pass 4: vertices */

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

template <bool Unit>
__global__ __launch_bounds__(256) void k_fe_vertices(const float* __restrict__ vol,
                                                     const uint64_t* __restrict__ mask, Geom g, float iso,
                                                     const uint64_t* __restrict__ voff,
                                                     float* __restrict__ pts, float* __restrict__ nrm)
{
    __shared__ uint32_t s_list[4][kListCap];
    const int lane = lane_id();
    const int wv = threadIdx.x >> 6;
    uint32_t* list = s_list[wv];
    const RowLanes rl = row_lanes(g.W);
    const int64_t R = g.ny * g.nz;
    const int64_t wave_id = (int64_t)blockIdx.x * 4 + wv;
    const int64_t r0 = wave_id * rl.rpw;
    if (r0 >= R) return;
    const int64_t r = r0 + rl.q;
    const bool ok = r < R && rl.w < g.W;
    const int ny = (int)g.ny;
    const int j = (int)(r % ny), k = (int)(r / ny);
    const bool has_y = j < ny - 1, has_z = k < (int)g.nz - 1;
    uint64_t m0 = 0, m0n = 0, my = 0, mz = 0;
    if (ok) {
        const uint64_t* __restrict__ mr = mask + r * g.W + rl.w;
        m0 = mr[0];
        m0n = rl.w + 1 < g.W ? mr[1] : 0ull;
        my = has_y ? mr[g.W] : 0ull;
        mz = has_z ? mr[g.ny * g.W] : 0ull;
    }
    const uint64_t vb = low_bits(g.nx - 64 * rl.w), eb = low_bits(g.nx - 1 - 64 * rl.w);
    const uint64_t cut[3] = {(m0 ^ shifted(m0, m0n)) & eb, has_y ? ((m0 ^ my) & vb) : 0ull,
                             has_z ? ((m0 ^ mz) & vb) : 0ull};
    const uint32_t c0 = __popcll(cut[0]), c1 = __popcll(cut[1]), c2 = __popcll(cut[2]);
    // position of this lane's first x / y / z entry in the wavefront's vertex sequence
    const uint32_t X = seg_sum(c0, rl.Wp), Y = seg_sum(c1, rl.Wp);
    const uint32_t lane_tot = c0 + c1 + c2;
    const uint32_t incl = wave_incl_scan_u32(lane_tot);
    const int seg_first_lane = rl.q * rl.Wp;
    const uint32_t before_row = __shfl(incl - lane_tot, seg_first_lane);   // vertices of earlier rows
    const uint32_t pos0 = before_row + seg_excl_scan(c0, rl.Wp, rl.w);
    const uint32_t pos1 = before_row + X + seg_excl_scan(c1, rl.Wp, rl.w);
    const uint32_t pos2 = before_row + X + Y + seg_excl_scan(c2, rl.Wp, rl.w);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    if (total == 0) return;
    const uint64_t vbase = voff[r0];

    for (uint32_t win = 0; win < total; win += kListCap) {
        // write the entries whose position falls in [win, win + kListCap): row q, axis a, bit i
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint64_t bits = cut[a];
            uint32_t p = a == 0 ? pos0 : (a == 1 ? pos1 : pos2);
            if (p + (uint32_t)__popcll(bits) <= win || p >= win + kListCap) continue;
            while (bits) {
                const int b = __builtin_ctzll(bits);
                bits &= bits - 1;
                if (p >= win && p < win + kListCap)
                    list[p - win] = ((uint32_t)rl.q << 20) | ((uint32_t)a << 18) | (uint32_t)(rl.w * 64 + b);
                ++p;
            }
        }
        wave_sync();
        const uint32_t n = min((uint32_t)kListCap, total - win);
        for (uint32_t e = lane; e < n; e += 64) {
            const uint32_t ent = list[e];
            const int64_t rr = r0 + (ent >> 20);
            const int a = (int)((ent >> 18) & 3);
            const int i = (int)(ent & 0x3FFFF);
            const int jj = (int)(rr % ny), kk = (int)(rr / ny);
            const uint64_t vid = vbase + win + e;
            float p[3], nn[3];
            fe_edge_point<Unit>(vol, g, iso, a, i, jj, kk, p, nn);
            pts[3 * vid + 0] = p[0];
            pts[3 * vid + 1] = p[1];
            pts[3 * vid + 2] = p[2];
            nrm[3 * vid + 0] = nn[0];
            nrm[3 * vid + 1] = nn[1];
            nrm[3 * vid + 2] = nn[2];
        }
        wave_sync();
    }
}

/* ------------------------------------------------------------------ pass 4: triangles */
constexpr int kTriMapCap = 512;     // triangle-map entries per wavefront and round (2 KB of LDS)

/* The wavefront's triangles are one contiguous range of the output (its rows are consecutive).
 * Each lane lists the triangles of its word's active cells into an LDS map, in output order
 * (cells in increasing i, then case-table order); the wavefront then resolves the map one
 * triangle per lane, so the index stores of neighbouring lanes are contiguous. Vertex ids come
 * from the owning lane's 8 edge-combination cut words and their id bases, kept in LDS relative
 * to the wavefront's first vertex (the launch checks that these offsets fit 32 bits). */
template <typename Index>
__global__ __launch_bounds__(256) void k_fe_triangles(const uint64_t* __restrict__ mask, Geom g,
                                                      const RowCounts* __restrict__ counts,
                                                      const uint64_t* __restrict__ voff,
                                                      const uint64_t* __restrict__ toff,
                                                      Index* __restrict__ tris)
{
    __shared__ uint64_t s_cut[8][256];
    __shared__ uint32_t s_base[8][256];
    __shared__ uint64_t s_case[256];
    __shared__ uint32_t s_map[4][kTriMapCap];
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wv = t >> 6;
    s_case[t] = kCaseEdges[t];
    __syncthreads();
    const RowLanes rl = row_lanes(g.W);
    const int64_t R = g.ny * g.nz;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + wv) * rl.rpw;
    if (r0 >= R) return;
    const int64_t r = r0 + rl.q;
    const int64_t j = r % g.ny, k = r / g.ny;
    const bool cells = r < R && j < g.ny - 1 && k < g.nz - 1 && rl.w < g.W;
    const uint32_t trow = cells ? counts[r].t : 0u;
    if (__ballot(trow != 0) == 0) return;
    const bool live = cells && trow != 0;
    const LaneWords L = load_lane_words(mask, g, r, rl.w, live, true, true);
    const uint64_t vb = low_bits(g.nx - 64 * rl.w), eb = low_bits(g.nx - 1 - 64 * rl.w);
    uint64_t cw[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) cw[q] = (L.m[q] ^ shifted(L.m[q], L.n[q])) & eb;
    cw[4] = (L.m[0] ^ L.m[1]) & vb;
    cw[5] = (L.m[2] ^ L.m[3]) & vb;
    cw[6] = (L.m[0] ^ L.m[2]) & vb;
    cw[7] = (L.m[1] ^ L.m[3]) & vb;
    const uint64_t v0 = voff[r0];
    const uint64_t t0 = toff[r0];
    uint64_t cb[8] = {v0, v0, v0, v0, v0, v0, v0, v0};
    if (live) {
        const int64_t rows[4] = {r, r + 1, r + g.ny, r + g.ny + 1};
        combo_bases(counts, voff, rows, cb);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        s_cut[q][t] = cw[q];
        s_base[q][t] = (uint32_t)(cb[q] - v0) + seg_excl_scan((uint32_t)__popcll(cw[q]), rl.Wp, rl.w);
    }
    const CornerWords C = corner_words(L);
    const uint64_t act = live ? active_cells(C, eb) : 0ull;
    uint32_t nt = 0;
    for (uint64_t a = act; a;) {
        const int b = __builtin_ctzll(a);
        a &= a - 1;
        nt += (uint32_t)case_tri_count(s_case[case_of(C, b)]);
    }
    const uint32_t incl = wave_incl_scan_u32(nt);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    const uint32_t first = incl - nt;
    uint32_t* map = s_map[wv];
    for (uint32_t win = 0; win < total; win += kTriMapCap) {
        const uint32_t wend = win + kTriMapCap;
        if (nt && first < wend && first + nt > win) {
            uint32_t p = first;
            for (uint64_t a = act; a && p < wend;) {
                const int b = __builtin_ctzll(a);
                a &= a - 1;
                const uint32_t cs = (uint32_t)case_of(C, b);
                const int n = case_tri_count(s_case[cs]);
                for (int tt = 0; tt < n; ++tt, ++p)
                    if (p >= win && p < wend) map[p - win] = (uint32_t)lane | (uint32_t)b << 6 | (uint32_t)tt << 12 | cs << 15;
            }
        }
        wave_sync();
        const uint32_t n = min((uint32_t)kTriMapCap, total - win);
        for (uint32_t e = lane; e < n; e += 64) {
            const uint32_t ent = map[e];
            const int owner = wv * 64 + (int)(ent & 63);
            const int b = (int)((ent >> 6) & 63);
            const int tt = (int)((ent >> 12) & 7);
            const uint64_t packed = s_case[ent >> 15];
            uint32_t id[3];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int ed = case_edge(packed, 3 * tt + s);
                const int q = combo_of_edge(ed);
                const int p = b + edge_dx(ed);
                const uint64_t cwq = s_cut[q][owner];
                const uint64_t below = p == 64 ? cwq : (cwq & ((1ull << p) - 1ull));
                id[s] = s_base[q][owner] + (uint32_t)__popcll(below);
            }
            store_tri<Index>(tris, t0 + win + e, v0 + id[0], v0 + id[1], v0 + id[2]);
        }
        wave_sync();
    }
}

}  // namespace miso

#endif
