/*
 * mc_rows.h -- Marching Cubes with the serial reference's vertex numbering, in the (grid row,
 * mask word) lane layout of fe_rows.h.
 *
 * Vertex order (marchingCubes/serial/main.cpp:145-177, first visit through a hash map): the
 * owner of a cut edge is its adjacent cell with the smallest (k, j, i); ids run over owner cells
 * in linear order and, within a cell, over its owned edges in order of first appearance in the
 * cell's case triangle list. Triangles are the Flying Edges triangle stream with each vertex id
 * mapped to its Marching Cubes id.
 *
 *   k_fe_mc_count_rows Flying Edges pass 2 and the owned cut edges per cell row, bit-sliced over
 *                      the 64 cells of a word, in one pass over the mask words
 *   k_mc_vertices_rows owners' points and normals at their Marching Cubes ids, and the map
 *                      from Flying Edges vertex id to Marching Cubes id
 *   k_fe_emit_rows with a map (fe_rows.h) writes the triangles
 */
#ifndef MINIISO_MC_ROWS_H
#define MINIISO_MC_ROWS_H

#include "fe_rows.h"

namespace miso {

/* Edges a cell owns: always 5, 6, 11; with k == 0 also 1, 2; with j == 0 also 4, 9; with
 * i == 0 also 7, 10; and 0 (j, k == 0), 3 (i, k == 0), 8 (i, j == 0). */
__device__ __forceinline__ uint32_t mc_owned_edges(bool i0, bool j0, bool k0)
{
    uint32_t m = (1u << 5) | (1u << 6) | (1u << 11);
    if (k0) m |= (1u << 1) | (1u << 2);
    if (j0) m |= (1u << 4) | (1u << 9);
    if (i0) m |= (1u << 7) | (1u << 10);
    if (j0 && k0) m |= 1u << 0;
    if (i0 && k0) m |= 1u << 3;
    if (i0 && j0) m |= 1u << 8;
    return m;
}

/* corner plane q (v0..v7) of the cells of a word */
__device__ __forceinline__ uint64_t corner_plane(const CornerWords& C, int q)
{
    switch (q) {
    case 0: return C.m[0];
    case 1: return C.s[0];
    case 2: return C.s[1];
    case 3: return C.m[1];
    case 4: return C.m[2];
    case 5: return C.s[2];
    case 6: return C.s[3];
    default: return C.m[3];
    }
}

/* Owned cut edges of the cell row of lane words L (row r = (j, k)), summed over the row's lanes
 * (valid in every lane of the row): bit-sliced over the 64 cells of a word. Lanes without cells
 * (the last row or plane, or past the row) count nothing, whatever their words. */
__device__ __forceinline__ uint32_t mc_owned_row_count(const LaneWords& L, const Geom& g, const RowLanes& rl, bool ok,
                                                       bool has_y, bool has_z, int64_t j, int64_t k)
{
    const CornerWords C = corner_words(L);
    const uint64_t eb = (ok && has_y && has_z) ? low_bits(g.nx - 1 - 64 * rl.w) : 0ull;
    const uint64_t first_cell = rl.w == 0 ? 1ull : 0ull;
    uint32_t cnt = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const uint64_t cut = corner_plane(C, kEdgeLo[e]) ^ corner_plane(C, kEdgeHi[e]);
        const bool need_i = (0x588u >> e) & 1u, need_j = (0x311u >> e) & 1u, need_k = (0x00Fu >> e) & 1u;
        // e: 0(j,k) 1(k) 2(k) 3(i,k) 4(j) 7(i) 8(i,j) 9(j) 10(i); always 5 6 11
        uint64_t own = eb;
        if (need_j && j != 0) own = 0;
        if (need_k && k != 0) own = 0;
        if (need_i) own &= first_cell;
        cnt += (uint32_t)__popcll(cut & own);
    }
    return seg_sum(cnt, rl.Wp);
}

/* Flying Edges pass 2 (k_fe_count_rows) and the first-visit owned-edge counts per cell row in
 * one pass over the mask words, each row's words loaded once. */
__global__ __launch_bounds__(256) void k_fe_mc_count_rows(const uint64_t* __restrict__ mask, Geom g,
                                                          RowCounts* __restrict__ counts,
                                                          RowTrim* __restrict__ trims,
                                                          uint64_t* __restrict__ scan_state,
                                                          int64_t scan_state_words, uint32_t* __restrict__ owned,
                                                          uint64_t* __restrict__ mc_scan_state,
                                                          int64_t mc_scan_state_words)
{
    // the scans of both numberings run next on the same stream; clear their tickets and status
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < scan_state_words;
         t += (int64_t)gridDim.x * blockDim.x)
        scan_state[t] = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < mc_scan_state_words;
         t += (int64_t)gridDim.x * blockDim.x)
        mc_scan_state[t] = 0;
    const int64_t R = g.ny * g.nz;
    fe_count_rows_block(mask, g, counts, trims, blockIdx.x,
                        [&](const LaneWords& L, const RowLanes& rl, int64_t r, bool ok, bool has_y, bool has_z,
                            int64_t j, int64_t k, bool empty) {
                            const uint32_t cnt = empty ? 0u : mc_owned_row_count(L, g, rl, ok, has_y, has_z, j, k);
                            if (r < R && rl.w == 0) owned[r] = cnt;
                        });
}

constexpr int kMcCellCap = 512;
constexpr int kMcVertCap = 64 * 12;

struct McRowsLds {
    uint64_t cut[8][64];
    uint32_t base[8][64];
    uint32_t cells[kMcCellCap];
    uint32_t verts[kMcVertCap];
};

/* Per case: for each cut edge e, its rank in order of first appearance in the case's triangle
 * list (nibble e, bits 4e .. 4e + 3), and the 12-bit mask of cut edges (bits 48 .. 59). */
__device__ __forceinline__ uint64_t first_appearance(uint64_t packed)
{
    uint64_t pos = 0;
    uint32_t seen = 0;
    int n = 0;
    const int ns = 3 * case_tri_count(packed);
    for (int s = 0; s < ns; ++s) {
        const int e = case_edge(packed, s);
        if ((seen >> e) & 1u) continue;
        seen |= 1u << e;
        pos |= (uint64_t)n << (4 * e);
        ++n;
    }
    return pos | (uint64_t)seen << 48;
}

/* Marching Cubes corner coordinates of the owner cell (marchingCubes/util/Image3D.cpp:22-65):
 * origin + index * spacing, plus one spacing on the far side. */
__device__ __forceinline__ void mc_corners(const Geom& g, int64_t i, int64_t j, int64_t k, int e, float A[3],
                                           float B[3])
{
    const float X = g.org[0] + (float)i * g.sp[0], Y = g.org[1] + (float)j * g.sp[1];
    const float Z = g.org[2] + (float)k * g.sp[2];
    A[0] = edge_dx(e) ? X + g.sp[0] : X;
    A[1] = edge_dy(e) ? Y + g.sp[1] : Y;
    A[2] = edge_dz(e) ? Z + g.sp[2] : Z;
    B[0] = A[0];
    B[1] = A[1];
    B[2] = A[2];
    const int a = edge_axis(e);
    B[a] = A[a] + g.sp[a];
}

/* Edge order (the order of util::duplicateRemover, marchingCubes/util/DuplicateRemover.h:34-70:
 * vertices sorted by global edge id, marchingCubes/util/Image3D.cpp:170-253): all cut x-edges in
 * (k, j, i) order, then all y-edges, then all z-edges. perm maps each Flying Edges vertex id
 * (row blocks [x][y][z]) to that order; xyz holds the per-row exclusive scans of the x, y and z
 * counts (R each) and tot their totals. Sixteen lanes per row, four rows per wavefront: the
 * lanes run over the row's ids [x][y][z] together (64-byte store runs). */
template <typename Index>
__global__ __launch_bounds__(256) void k_edge_order_perm(const RowCounts* __restrict__ counts,
                                                         const uint64_t* __restrict__ voff,
                                                         const uint64_t* __restrict__ xyz,
                                                         const uint64_t* __restrict__ tot, int64_t R,
                                                         Index* __restrict__ perm)
{
    const int64_t r = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    if (r >= R) return;
    const uint32_t l = threadIdx.x & 15;
    const RowCounts c = counts[r];
    const uint32_t n = c.x + c.y + c.z;
    if (n == 0) return;
    const uint64_t f0 = voff[r];
    const uint64_t bx = xyz[r], by = tot[0] + xyz[R + r], bz = tot[0] + tot[1] + xyz[2 * R + r];
    const uint32_t cxy = c.x + c.y;
    for (uint32_t t = l; t < n; t += 16) {
        const uint64_t id = t < c.x ? bx + t : (t < cxy ? by + (t - c.x) : bz + (t - cxy));
        perm[f0 + t] = (Index)id;
    }
}

/* EdgeOrder: each owned edge's point goes to fe_to_mc[its Flying Edges id] (an edge-order
 * permutation made beforehand), which this kernel then only reads. */
template <typename Index, bool Unit, bool EdgeOrder = false>
__global__ __launch_bounds__(256) void k_mc_vertices_rows(const float* __restrict__ vol,
                                                          const uint64_t* __restrict__ mask, Geom g, float iso,
                                                          const RowCounts* __restrict__ counts,
                                                          const uint64_t* __restrict__ voff,
                                                          const uint64_t* __restrict__ moff,
                                                          float* __restrict__ pts, float* __restrict__ nrm,
                                                          Index* __restrict__ fe_to_mc)
{
    __shared__ uint64_t s_first[256];
    __shared__ McRowsLds s_w[4];
    s_first[threadIdx.x] = first_appearance(kCaseEdges[threadIdx.x]);
    __syncthreads();
    const int lane = lane_id();
    const int wv = wave_in_block();
    McRowsLds& S = s_w[wv];
    const RowLanes rl = row_lanes(g.W);
    const int64_t R = g.ny * g.nz;
    const int64_t r0 = (emit_block_order(blockIdx.x, gridDim.x, g) * 4 + wv) * rl.rpw;   // Flying Edges pass 4's order
    if (r0 >= R) return;
    const int64_t r = r0 + rl.q;
    const bool ok = r < R && rl.w < g.W;
    const RowJK jk = split_row(r0, rl.q, g.ny);
    const int64_t j = jk.j, k = jk.k;
    const bool cells = ok && j < g.ny - 1 && k < g.nz - 1;
    const LaneWords L = load_lane_words(mask, g, r, rl.w, cells, true, true);
    const uint64_t vb = low_bits(g.nx - 64 * rl.w), eb = low_bits(g.nx - 1 - 64 * rl.w);
    const CornerWords C = corner_words(L);
    const uint64_t act = cells ? active_cells(C, eb) : 0ull;
    const uint32_t ncell = (uint32_t)__popcll(act);
    const uint32_t cincl = wave_incl_scan_u32(ncell);
    const uint32_t ctotal = (uint32_t)__builtin_amdgcn_readlane((int)cincl, 63);
    if (ctotal == 0) return;
    const uint64_t v0 = voff[r0];
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
    const uint32_t clo[8] = {lo[0], lo[1], lo[3], lo[2], lo[4], lo[5], lo[7], lo[6]};
    const uint32_t chi[8] = {hi[0], hi[1], hi[3], hi[2], hi[4], hi[5], hi[7], hi[6]};
    const uint32_t cfirst = cincl - ncell;
    const int64_t k0 = uniform_div(r0, g.ny), j0 = r0 - k0 * g.ny;
    uint64_t mid = EdgeOrder ? 0 : moff[r0];   // the wavefront's first Marching Cubes vertex
    for (uint32_t win = 0; win < ctotal; win += kMcCellCap) {
        const uint32_t wend = win + kMcCellCap;
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
                        S.cells[p - win] = (uint32_t)lane | (uint32_t)(b + 32 * h) << 6 | cs << 12;
                    }
                    ++p;
                }
            }
        }
        wave_sync();
        const uint32_t n = min((uint32_t)kMcCellCap, ctotal - win);
        // The owned edges of consecutive 64-cell groups collect in S.verts until the next group
        // might not fit: x-edges from the front, y/z-edges from the back, so that each step of
        // the flush runs one vertex path on 64 lanes. Entry: window cell | edge << 9 | id << 13,
        // the id relative to fbase (the first Marching Cubes vertex of the collection).
        uint32_t fx = 0, fyz = 0;
        uint64_t fbase = mid;
        auto vertex_of = [&](uint32_t ve, bool xpath) {
            const uint32_t cent = S.cells[ve & 511];
            const int e = (int)((ve >> 9) & 15);
            const int src = (int)(cent & 63), b = (int)((cent >> 6) & 63);
            const int q = src >> __builtin_ctz((unsigned)rl.Wp), w = src & (rl.Wp - 1);
            int64_t jj = j0 + q, kk = k0;
            while (jj >= g.ny) {
                jj -= g.ny;
                ++kk;
            }
            const int64_t i = 64 * (int64_t)w + b;
            float pp[3], nn[3];
            if (Unit) {
                // unit spacing, zero origin: the owner's corner coordinates are float(index),
                // as Flying Edges' are, so the Flying Edges vertex code gives the same bits
                const int ia = (int)i + edge_dx(e), ja = (int)jj + edge_dy(e), ka = (int)kk + edge_dz(e);
                const float* __restrict__ v = vol + ((int64_t)ka * g.ny + ja) * g.nx + ia;
                if (xpath)
                    vertex_x<true>(v, g, iso, ia, ja, ka, pp, nn);
                else
                    vertex_yz<true>(v, g, iso, edge_axis(e), ia, ja, ka, pp, nn);
            } else {
                float A[3], B[3];
                mc_corners(g, i, jj, kk, e, A, B);
                edge_point(vol, g, iso, edge_axis(e), i + edge_dx(e), jj + edge_dy(e), kk + edge_dz(e), A, B, pp, nn);
            }
            // the edge's Flying Edges id, as the triangle pass computes it
            const int cq = combo_of_edge(e);
            const int p = b + edge_dx(e);
            const uint64_t cwq = S.cut[cq][src];
            const uint64_t below = p == 64 ? cwq : (cwq & ((1ull << p) - 1ull));
            const uint64_t fe = v0 + S.base[cq][src] + (uint32_t)__popcll(below);
            const uint64_t id = EdgeOrder ? (uint64_t)fe_to_mc[fe] : fbase + (ve >> 13);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pts[3 * id + c] = pp[c];
                nrm[3 * id + c] = nn[c];
            }
            if (!EdgeOrder) fe_to_mc[fe] = (Index)id;
        };
        auto flush = [&]() {
            wave_sync();
            for (uint32_t v = lane; v < fx; v += 64) vertex_of(S.verts[v], true);
            for (uint32_t v = lane; v < fyz; v += 64) vertex_of(S.verts[kMcVertCap - 1 - v], false);
            wave_sync();
            fx = fyz = 0;
            fbase = mid;
        };
        constexpr uint32_t kXEdges = 0x55u;   // edges 0, 2, 4, 6 run along x
        for (uint32_t g0 = 0; g0 < n; g0 += 64) {
            // one cell per lane: its owned cut edges, in first-appearance order
            const uint32_t idx = g0 + (uint32_t)lane;
            uint32_t ent = 0, own = 0;
            uint64_t order = 0;
            if (idx < n) {
                ent = S.cells[idx];
                const int src = (int)(ent & 63), b = (int)((ent >> 6) & 63);
                const int q = src >> __builtin_ctz((unsigned)rl.Wp), w = src & (rl.Wp - 1);
                int64_t jj = j0 + q, kk = k0;
                while (jj >= g.ny) {
                    jj -= g.ny;
                    ++kk;
                }
                order = s_first[ent >> 12];
                own = (uint32_t)(order >> 48) & 0xFFFu & mc_owned_edges(w == 0 && b == 0, jj == 0, kk == 0);
            }
            const uint32_t no = (uint32_t)__popc(own), nox = (uint32_t)__popc(own & kXEdges);
            // both counts in one scan (a group has at most 64 * 12 owned edges)
            const uint32_t inc = wave_incl_scan_u32(no | nox << 16);
            const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
            const uint32_t vtot = tot & 0xFFFFu, xtot = tot >> 16;
            if (fx + fyz + vtot > (uint32_t)kMcVertCap) flush();
            if (own) {
                // first-visit rank of an owned edge: owned edges appearing earlier in the cell's
                // triangles, after the vertices of earlier lanes and groups
                uint32_t posmask = 0;
                for (uint32_t o = own; o; o &= o - 1) posmask |= 1u << ((order >> (4 * __builtin_ctz(o))) & 0xF);
                const uint32_t base = (uint32_t)(mid - fbase) + (inc & 0xFFFFu) - no;
                uint32_t sx = fx + (inc >> 16) - nox;
                uint32_t syz = fyz + ((inc & 0xFFFFu) - no) - ((inc >> 16) - nox);
                for (uint32_t o = own; o; o &= o - 1) {
                    const int e = __builtin_ctz(o);
                    const uint32_t p = (uint32_t)(order >> (4 * e)) & 0xF;
                    const uint32_t rel = base + (uint32_t)__popc(posmask & ((1u << p) - 1u));
                    const uint32_t ve = idx | (uint32_t)e << 9 | rel << 13;
                    if ((kXEdges >> e) & 1u)
                        S.verts[sx++] = ve;
                    else
                        S.verts[kMcVertCap - 1 - syz++] = ve;
                }
            }
            fx += xtot;
            fyz += vtot - xtot;
            mid += vtot;
        }
        flush();
        wave_sync();
    }
}

}  // namespace miso

#endif
