/*
 * fe_kernels.h -- Flying Edges passes 1-4 as HIP kernels for gfx950 (wave64).
 *
 *   k_fe_classify  pass 1a: one wavefront per grid row, coalesced loads, v >= iso through
 *                  __ballot into the 1-bit sign mask (replaces FlyingEdgesAlgorithm.cpp:31-45).
 *   k_fe_count     pass 1b + pass 2: from the masks of rows (j,k), (j+1,k), (j,k+1), (j+1,k+1):
 *                  xl/xr trims (cpp:47-65 with isCutEdge :515-554), cut-edge counts per grid
 *                  row and triangle counts per cell row (cpp:72-190) with popcounts.
 *   k_fe_scan      pass 3: single-pass decoupled look-back exclusive scan of the per-row vertex
 *                  and triangle counts into 64-bit offsets (cpp:196-269).
 *   k_fe_emit      pass 4: per grid row, the cut edges are compacted into an LDS list already in
 *                  output order, so lanes write consecutive vertices; per cell row, the active
 *                  cells are compacted, a wave prefix sum of numTris gives triangle slots, and
 *                  vertex ids come from row offsets plus popcount ranks (cpp:275-497).
 */
#ifndef MINIISO_FE_KERNELS_H
#define MINIISO_FE_KERNELS_H

#include "fe_common.h"

namespace miso {

constexpr int kRowsPerBlock = 4;        // one wavefront per grid row, 4 waves per block
constexpr int kSegWords = 16;           // pass 4 processes a row in segments of 16 words (1024 voxels)
constexpr int kScanItems = 16;
constexpr int kScanTile = 256 * kScanItems;
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

/* ------------------------------------------------------------------ pass 1a: sign mask */
/* Rows 4 blk .. 4 blk + 3, one per wavefront. Each lane loads 4 consecutive voxels (one 16-byte
 * non-temporal load, so a wavefront step covers 256 voxels = 4 mask words); its 4 sign bits are placed at
 * bit 4 * (lane % 16) and OR-reduced over each 16-lane row with DPP, so lane 16 t + 15 ends up
 * holding mask word t of the step. Voxels past nx read as below the isovalue. Rows of at most
 * 256 voxels take one voxel per lane and __ballot instead (fewer instructions per short row).
 * The voxels are loaded once and classified against n isovalues of the same volume:
 * sink(t, iso, mask) names isovalue t and the mask array it writes (n = 1 for one extraction). */
template <class Sink>
__device__ __forceinline__ void fe_classify_block_n(const float* __restrict__ vol, const Geom& g, int n, Sink sink,
                                                    int64_t blk)
{
    const int lane = lane_id();
    const int64_t r = blk * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= g.ny * g.nz) return;
    const float* __restrict__ row = vol + r * g.nx;
    if (g.W <= 4) {
        // short rows: one 4-byte voxel per lane and word, the sign bits straight from __ballot
        float v[4];
        bool in[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = 64 * u + lane;
            in[u] = u < g.W && i < g.nx;
            v[u] = in[u] ? row[i] : 0.0f;
        }
        for (int t = 0; t < n; ++t) {
            float iso;
            uint64_t* mask;
            sink(t, iso, mask);
            uint64_t* __restrict__ mrow = mask + r * g.W;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint64_t b = __ballot(in[u] && v[u] >= iso);
                if (lane == 0 && u < g.W) mrow[u] = b;
            }
        }
        return;
    }
    const int sub = lane & 15;
    constexpr int kSteps = 4;                      // 1024 voxels in flight per wavefront
    for (int64_t c0 = 0; c0 < g.nx; c0 += 256 * kSteps) {
        float v[kSteps][4];
        uint32_t in[kSteps];                       // bit c: voxel i + c lies in the row
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
#pragma unroll
            for (int c = 0; c < 4; ++c) v[u][c] = 0.0f;
            in[u] = 0;
            if (c0 + 256 * u >= g.nx) continue;    // wavefront-uniform: short rows skip whole steps
            const int64_t i = c0 + 256 * u + 4 * lane;
            if (i + 4 <= g.nx) {
                const Vox4 q = load_vox4_stream(row + i);
                v[u][0] = q.x; v[u][1] = q.y; v[u][2] = q.z; v[u][3] = q.w;
                in[u] = 15u;
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    v[u][c] = i + c < g.nx ? row[i + c] : 0.0f;
                    in[u] |= (uint32_t)(i + c < g.nx) << c;
                }
            }
        }
        for (int t = 0; t < n; ++t) {
            float iso;
            uint64_t* mask;
            sink(t, iso, mask);
            uint64_t* __restrict__ mrow = mask + r * g.W;
#pragma unroll
            for (int u = 0; u < kSteps; ++u) {
                if (c0 + 256 * u >= g.nx) break;
                uint32_t nib = 0;
#pragma unroll
                for (int c = 0; c < 4; ++c) nib |= (uint32_t)(v[u][c] >= iso) << c;
                nib &= in[u];
                // bits of lane sub go to 4 sub .. 4 sub + 3 of the (lo, hi) 32-bit halves of the word
                uint32_t lo = sub < 8 ? nib << (4 * sub) : 0u;
                uint32_t hi = sub < 8 ? 0u : nib << (4 * (sub - 8));
                lo |= dpp_u32<kDppRowShr1>(lo);
                hi |= dpp_u32<kDppRowShr1>(hi);
                lo |= dpp_u32<kDppRowShr2>(lo);
                hi |= dpp_u32<kDppRowShr2>(hi);
                lo |= dpp_u32<kDppRowShr4>(lo);
                hi |= dpp_u32<kDppRowShr4>(hi);
                lo |= dpp_u32<kDppRowShr8>(lo);
                hi |= dpp_u32<kDppRowShr8>(hi);
                const int64_t w = (c0 + 256 * u) / 64 + (lane >> 4);
                if (sub == 15 && w < g.W) mrow[w] = (uint64_t)hi << 32 | lo;
            }
        }
    }
}

/* Rows of at most 256 voxels (W <= 4) in blocks of 4 kShortRows rows: each wavefront loads
 * kShortRows consecutive rows (one voxel per lane and word, as the short-row path above) before
 * classifying them, one memory round trip for all of them. The batch's pass 1 takes this for
 * its short-row jobs, where a work item otherwise moves one short row per wavefront. */
constexpr int kShortRows = 4;

template <class Sink>
__device__ __forceinline__ void fe_classify_short_rows_n(const float* __restrict__ vol, const Geom& g, int n,
                                                         Sink sink, int64_t blk)
{
    const int lane = lane_id();
    const int64_t R = g.ny * g.nz;
    const int64_t r0 = (blk * 4 + (threadIdx.x >> 6)) * kShortRows;
    if (r0 >= R) return;
    float v[kShortRows][4];
    bool in[kShortRows][4];
#pragma unroll
    for (int q = 0; q < kShortRows; ++q) {
        const int64_t r = r0 + q;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = 64 * u + lane;
            in[q][u] = r < R && u < g.W && i < g.nx;
            v[q][u] = in[q][u] ? vol[r * g.nx + i] : 0.0f;
        }
    }
    for (int t = 0; t < n; ++t) {
        float iso;
        uint64_t* mask;
        sink(t, iso, mask);
#pragma unroll
        for (int q = 0; q < kShortRows; ++q) {
            const int64_t r = r0 + q;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint64_t b = __ballot(in[q][u] && v[q][u] >= iso);
                if (lane == 0 && u < g.W && r < R) mask[r * g.W + u] = b;
            }
        }
    }
}

__device__ __forceinline__ void fe_classify_block(const float* __restrict__ vol, const Geom& g, float iso,
                                                  uint64_t* __restrict__ mask, int64_t blk)
{
    fe_classify_block_n(vol, g, 1, [&](int, float& is, uint64_t*& m) { is = iso; m = mask; }, blk);
}

__global__ __launch_bounds__(256) void k_fe_classify(const float* __restrict__ vol, Geom g, float iso,
                                                     uint64_t* __restrict__ mask)
{
    fe_classify_block(vol, g, iso, mask, blockIdx.x);
}

/* ------------------------------------------------------------------ pass 1b + 2: counts */
__global__ __launch_bounds__(256) void k_fe_count(const uint64_t* __restrict__ mask, Geom g,
                                                  RowCounts* __restrict__ counts,
                                                  RowTrim* __restrict__ trims,
                                                  uint64_t* __restrict__ scan_state,
                                                  int64_t scan_state_words)
{
    __shared__ uint8_t s_ntri[256];
    for (int c = threadIdx.x; c < 256; c += blockDim.x) s_ntri[c] = (uint8_t)case_tri_count(kCaseEdges[c]);
    // The scan of pass 3 runs next on the same stream; clear its tickets and tile status here.
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < scan_state_words;
         t += (int64_t)gridDim.x * blockDim.x)
        scan_state[t] = 0;
    __syncthreads();

    const int lane = lane_id();
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= g.ny * g.nz) return;
    const int64_t j = r % g.ny, k = r / g.ny;
    const bool has_y = j < g.ny - 1, has_z = k < g.nz - 1, cells = has_y && has_z;
    const int64_t W = g.W;
    const uint64_t* __restrict__ m0p = mask + r * W;
    const uint64_t* __restrict__ m1p = m0p + W;            // row (j+1, k)
    const uint64_t* __restrict__ m2p = m0p + g.ny * W;     // row (j, k+1)
    const uint64_t* __restrict__ m3p = m2p + W;            // row (j+1, k+1)

    uint32_t xc = 0, yc = 0, zc = 0, tc = 0;
    int64_t first = INT64_MAX, last = -1;
    for (int64_t w0 = 0; w0 < W; w0 += 64) {
        const int64_t w = w0 + lane;
        if (w >= W) break;
        const bool nxt = w + 1 < W;
        const uint64_t m0 = m0p[w], m0n = nxt ? m0p[w + 1] : 0;
        const uint64_t my = has_y ? m1p[w] : 0, myn = (has_y && nxt) ? m1p[w + 1] : 0;
        const uint64_t mz = has_z ? m2p[w] : 0, mzn = (has_z && nxt) ? m2p[w + 1] : 0;
        const uint64_t vb = low_bits(g.nx - 64 * w);          // vertices i < nx
        const uint64_t eb = low_bits(g.nx - 1 - 64 * w);      // x-edges / cells i < nx-1
        const uint64_t s0 = (m0 >> 1) | (m0n << 63);           // bit i -> vertex i+1
        const uint64_t xcut = (m0 ^ s0) & eb;
        const uint64_t ycut = has_y ? ((m0 ^ my) & vb) : 0;
        const uint64_t zcut = has_z ? ((m0 ^ mz) & vb) : 0;
        xc += __popcll(xcut);
        yc += __popcll(ycut);
        zc += __popcll(zcut);
        // isCutEdge tests i in [0, nx-2] only (FlyingEdgesAlgorithm.cpp:52-63)
        const uint64_t trim = (xcut | ycut | zcut) & eb;
        if (trim) {
            first = min(first, 64 * w + (int64_t)__builtin_ctzll(trim));
            last = max(last, 64 * w + 63 - (int64_t)__builtin_clzll(trim));
        }
        if (cells) {
            const uint64_t myz = m3p[w], myzn = nxt ? m3p[w + 1] : 0;
            const uint64_t sy = (my >> 1) | (myn << 63);
            const uint64_t sz = (mz >> 1) | (mzn << 63);
            const uint64_t syz = (myz >> 1) | (myzn << 63);
            const uint64_t all = m0 & s0 & sy & my & mz & sz & syz & myz;
            const uint64_t any = m0 | s0 | sy | my | mz | sz | syz | myz;
            uint64_t act = any & ~all & eb;
            while (act) {
                const int b = __builtin_ctzll(act);
                act &= act - 1;
                const int c = (int)((m0 >> b) & 1) | (int)((s0 >> b) & 1) << 1 | (int)((sy >> b) & 1) << 2 |
                              (int)((my >> b) & 1) << 3 | (int)((mz >> b) & 1) << 4 |
                              (int)((sz >> b) & 1) << 5 | (int)((syz >> b) & 1) << 6 |
                              (int)((myz >> b) & 1) << 7;
                tc += s_ntri[c];
            }
        }
    }
    xc = wave_sum_u32(xc);
    yc = wave_sum_u32(yc);
    zc = wave_sum_u32(zc);
    tc = wave_sum_u32(tc);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        first = min(first, (int64_t)__shfl_xor((long long)first, o));
        last = max(last, (int64_t)__shfl_xor((long long)last, o));
    }
    if (lane == 0) {
        counts[r] = RowCounts{xc, yc, zc, tc};
        trims[r] = RowTrim{last < 0 ? (uint32_t)g.nx : (uint32_t)first, last < 0 ? 0u : (uint32_t)(last + 1)};
    }
}

/* ------------------------------------------------------------------ pass 3: scan */

/* Decoupled look-back (Merrill & Garland): one wavefront inspects 64 predecessors at a time,
 * waiting until each has published an aggregate or an inclusive prefix. Returns the exclusive
 * prefix of `tile`. */
__device__ uint64_t lookback(uint64_t* __restrict__ status, int64_t tile)
{
    const int lane = lane_id();
    uint64_t excl = 0;
    int64_t p = tile - 1;
    for (;;) {
        const int64_t q = p - lane;
        uint64_t s = kFlagPre;          // "before the first tile": an inclusive prefix of 0
        if (q >= 0) {
            do {
                s = __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } while ((s >> 62) == 0);
        }
        const uint64_t pre = __ballot((s >> 62) == 2);
        if (pre) {
            const int stop = __ffsll((unsigned long long)pre) - 1;
            excl += wave_sum_u64(lane <= stop ? (s & kValMask) : 0);
            return excl;
        }
        excl += wave_sum_u64(s & kValMask);
        p -= 64;
    }
}

/* One 4096-row tile of the scan (tile numbers are taken from state[0] in launch order, so that a
 * tile's predecessors have started: the look-back waits for them). */
__device__ __forceinline__ void fe_scan_tile(const RowCounts* __restrict__ counts, int64_t R,
                                             uint64_t* __restrict__ voff, uint64_t* __restrict__ toff,
                                             uint64_t* __restrict__ state, int64_t ntiles,
                                             uint64_t* __restrict__ totals, int64_t tile)
{
    __shared__ uint64_t s_wv[4], s_wt[4];
    __shared__ uint64_t s_ev, s_et;
    uint64_t* stv = state + 1;
    uint64_t* stt = state + 1 + ntiles;
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int64_t r0 = tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems], t[kScanItems];
    uint64_t sv = 0, st = 0;
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
        const int64_t r = r0 + u;
        if (r < R) {
            const RowCounts c = counts[r];
            v[u] = (uint64_t)c.x + c.y + c.z;
            t[u] = c.t;
        } else {
            v[u] = 0;
            t[u] = 0;
        }
        sv += v[u];
        st += t[u];
    }
    const uint64_t iv = wave_incl_scan_u64(sv), it = wave_incl_scan_u64(st);
    if (lane == 63) {
        s_wv[wave] = iv;
        s_wt[wave] = it;
    }
    __syncthreads();
    uint64_t wbv = 0, wbt = 0, aggv = 0, aggt = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q < wave) {
            wbv += s_wv[q];
            wbt += s_wt[q];
        }
        aggv += s_wv[q];
        aggt += s_wt[q];
    }
    if (wave == 0) {
        uint64_t ev = 0, et = 0;
        if (tile == 0) {
            if (lane == 0) {
                __hip_atomic_store(&stv[0], kFlagPre | aggv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&stt[0], kFlagPre | aggt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            if (lane == 0) {
                __hip_atomic_store(&stv[tile], kFlagAgg | aggv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&stt[tile], kFlagAgg | aggt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            ev = lookback(stv, tile);
            et = lookback(stt, tile);
            if (lane == 0) {
                __hip_atomic_store(&stv[tile], kFlagPre | (ev + aggv), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&stt[tile], kFlagPre | (et + aggt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) {
            s_ev = ev;
            s_et = et;
            if (tile == ntiles - 1) {
                totals[0] = ev + aggv;
                totals[1] = et + aggt;
            }
        }
    }
    __syncthreads();
    uint64_t pv = s_ev + wbv + iv - sv, pt = s_et + wbt + it - st;
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
        const int64_t r = r0 + u;
        if (r < R) {
            voff[r] = pv;
            toff[r] = pt;
        }
        pv += v[u];
        pt += t[u];
    }
}

/* The next tile number of the scan, for every thread of the block. */
__device__ __forceinline__ int64_t fe_scan_ticket(uint64_t* __restrict__ state)
{
    __shared__ int64_t s_tile;
    __syncthreads();   // every thread has read the previous tile number
    if (threadIdx.x == 0)
        s_tile = (int64_t)__hip_atomic_fetch_add((unsigned long long*)&state[0], 1ull, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_tile;
}

__global__ __launch_bounds__(256) void k_fe_scan(const RowCounts* __restrict__ counts, int64_t R,
                                                 uint64_t* __restrict__ voff, uint64_t* __restrict__ toff,
                                                 uint64_t* __restrict__ state, int64_t ntiles,
                                                 uint64_t* __restrict__ totals)
{
    // state[0]: ticket, state[1 .. ntiles]: vertex status, state[1+ntiles ..]: triangle status
    fe_scan_tile(counts, R, voff, toff, state, ntiles, totals, fe_scan_ticket(state));
}

/* ------------------------------------------------------------------ pass 4: emit */

struct EmitLds {
    uint64_t m[4][kSegWords + 1];         // sign-mask words of rows (j,k) (j+1,k) (j,k+1) (j+1,k+1)
    uint64_t cut[8][kSegWords + 1];       // cut words per (axis,row) combination, see combo_of_edge
    uint32_t pre[8][kSegWords + 1];       // exclusive popcount prefix of cut[q] within the segment
    uint16_t list[3 * 64 * kSegWords];    // compacted cut edges (axis << 14 | i) or active cells
};

/* Vertex id combination of a cube edge: which of the four grid rows holds the edge and along
 * which axis. 0-3: x-edges of rows (j,k) (j+1,k) (j,k+1) (j+1,k+1); 4-5: y-edges of rows (j,k)
 * (j,k+1); 6-7: z-edges of rows (j,k) (j+1,k). */
__device__ __forceinline__ int combo_of_edge(int e)
{
    // e:      0  1  2  3  4  5  6  7  8  9  10 11
    // combo:  0  4  1  4  2  5  3  5  6  6  7  7
    return (int)((0x776653524140ull >> (4 * e)) & 0xF);
}

/* Walks one row of the grid (j,k) segment by segment; holds the four sign-mask rows of the
 * cell row (j,k) in LDS. Each wavefront owns one RowWalker and its EmitLds slice. */
struct RowWalker {
    EmitLds& L;
    const uint64_t* __restrict__ mask;
    const Geom& g;
    int64_t rows[4];
    bool row_ok[4];
    int64_t s0w = 0;   // first word of the current segment
    int nw = 0;        // words in the current segment

    __device__ RowWalker(EmitLds& l, const uint64_t* __restrict__ m, const Geom& gg, int64_t r, bool has_y, bool has_z)
        : L(l), mask(m), g(gg)
    {
        rows[0] = r; rows[1] = r + 1; rows[2] = r + g.ny; rows[3] = r + g.ny + 1;
        row_ok[0] = true; row_ok[1] = has_y; row_ok[2] = has_z; row_ok[3] = has_y && has_z;
    }

    /* stage mask words [s0w, s0w + nw] (one extra word for the +1 neighbours) of the 4 rows */
    __device__ void stage(int64_t seg_w0)
    {
        const int lane = lane_id();
        s0w = seg_w0;
        nw = (int)min((int64_t)kSegWords, g.W - s0w);
        for (int t = lane; t < 4 * (kSegWords + 1); t += 64) {
            const int q = t / (kSegWords + 1), w = t % (kSegWords + 1);
            const int64_t gw = s0w + w;
            L.m[q][w] = (row_ok[q] && w <= nw && gw < g.W) ? mask[rows[q] * g.W + gw] : 0ull;
        }
        wave_sync();
    }

    /* Compacts the cut edges of grid row (j,k) in this segment into L.list: x-edges, then
     * y-edges, then z-edges, each by increasing i. Returns the count; per-axis starts/counts. */
    __device__ int compact_row_edges(bool has_y, bool has_z, int start[3], int cnt[3])
    {
        const int lane = lane_id();
        int n = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            start[a] = n;
            if ((a == 1 && !has_y) || (a == 2 && !has_z)) {
                cnt[a] = 0;
                continue;
            }
            for (int w = 0; w < nw; ++w) {
                const int64_t gw = s0w + w;
                const uint64_t m0 = L.m[0][w];
                uint64_t bits;
                if (a == 0) {
                    const uint64_t sh = (m0 >> 1) | (L.m[0][w + 1] << 63);
                    bits = (m0 ^ sh) & low_bits(g.nx - 1 - 64 * gw);
                } else {
                    bits = (m0 ^ L.m[a][w]) & low_bits(g.nx - 64 * gw);
                }
                if ((bits >> lane) & 1ull)
                    L.list[n + __popcll(bits & __lanemask_lt())] = (uint16_t)((a << 14) | (w * 64 + lane));
                n += __popcll(bits);
            }
            cnt[a] = n - start[a];
        }
        wave_sync();
        return n;
    }

    /* Cut words of the 8 (axis,row) combinations and their popcount prefix for this segment,
     * then the compacted list of active cells (cells with at least one cut edge). */
    __device__ int prepare_cells()
    {
        const int lane = lane_id();
        if (lane <= nw) {
            const int w = lane;
            const int64_t gw = s0w + w;
            const uint64_t vb = low_bits(g.nx - 64 * gw), eb = low_bits(g.nx - 1 - 64 * gw);
            const bool has_next = w < nw;   // word w+1 staged (word nw is only read at bit 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint64_t m = L.m[q][w];
                const uint64_t nb = has_next ? L.m[q][w + 1] : 0ull;
                L.cut[q][w] = (m ^ ((m >> 1) | (nb << 63))) & eb;
            }
            L.cut[4][w] = (L.m[0][w] ^ L.m[1][w]) & vb;
            L.cut[5][w] = (L.m[2][w] ^ L.m[3][w]) & vb;
            L.cut[6][w] = (L.m[0][w] ^ L.m[2][w]) & vb;
            L.cut[7][w] = (L.m[1][w] ^ L.m[3][w]) & vb;
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t pc = lane <= nw ? (uint32_t)__popcll(L.cut[q][lane]) : 0u;
            const uint32_t inc = wave_incl_scan_u32(pc);
            if (lane <= nw) L.pre[q][lane] = inc - pc;
        }
        int na = 0;
        for (int w = 0; w < nw; ++w) {
            const int64_t gw = s0w + w;
            const uint64_t eb = low_bits(g.nx - 1 - 64 * gw);
            const uint64_t a0 = L.m[0][w], a1 = L.m[1][w], a2 = L.m[2][w], a3 = L.m[3][w];
            const uint64_t b0 = (a0 >> 1) | (L.m[0][w + 1] << 63);
            const uint64_t b1 = (a1 >> 1) | (L.m[1][w + 1] << 63);
            const uint64_t b2 = (a2 >> 1) | (L.m[2][w + 1] << 63);
            const uint64_t b3 = (a3 >> 1) | (L.m[3][w + 1] << 63);
            const uint64_t all = a0 & a1 & a2 & a3 & b0 & b1 & b2 & b3;
            const uint64_t any = a0 | a1 | a2 | a3 | b0 | b1 | b2 | b3;
            const uint64_t act = any & ~all & eb;
            if ((act >> lane) & 1ull) L.list[na + __popcll(act & __lanemask_lt())] = (uint16_t)(w * 64 + lane);
            na += __popcll(act);
        }
        wave_sync();
        return na;
    }

    /* marching-cubes case of the cell at local index li of this segment */
    __device__ int cell_case(int li) const
    {
        const int w = li >> 6, b = li & 63;
        auto bit = [&](int q, int dx) -> int {
            const int p = b + dx;
            return (int)((L.m[q][w + (p >> 6)] >> (p & 63)) & 1ull);
        };
        return bit(0, 0) | bit(0, 1) << 1 | bit(1, 1) << 2 | bit(1, 0) << 3 | bit(2, 0) << 4 |
               bit(2, 1) << 5 | bit(3, 1) << 6 | bit(3, 0) << 7;
    }

    /* Flying Edges vertex id of cube edge `e` of the cell at local index li: the row offset plus
     * axis-block offset (cbase) plus the rank of the edge among the row's cut edges. */
    __device__ uint64_t fe_vertex_id(const uint64_t cbase[8], int e, int li) const
    {
        const int q = combo_of_edge(e);
        const int p = li + edge_dx(e);
        const int w = p >> 6, b = p & 63;
        const uint64_t below = L.cut[q][w] & ((1ull << b) - 1ull);
        return cbase[q] + L.pre[q][w] + (uint64_t)__popcll(below);
    }

    __device__ void advance_bases(uint64_t cbase[8]) const
    {
#pragma unroll
        for (int q = 0; q < 8; ++q) cbase[q] += L.pre[q][nw - 1] + __popcll(L.cut[q][nw - 1]);
    }
};

/* Row-offset plus axis-block bases of the 8 combinations for cell row r. */
__device__ __forceinline__ void combo_bases(const RowCounts* __restrict__ counts, const uint64_t* __restrict__ voff,
                                            const int64_t rows[4], uint64_t cbase[8])
{
    RowCounts cq[4];
    uint64_t pq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cq[q] = counts[rows[q]];
        pq[q] = voff[rows[q]];
    }
    cbase[0] = pq[0]; cbase[1] = pq[1]; cbase[2] = pq[2]; cbase[3] = pq[3];
    cbase[4] = pq[0] + cq[0].x;
    cbase[5] = pq[2] + cq[2].x;
    cbase[6] = pq[0] + cq[0].x + cq[0].y;
    cbase[7] = pq[1] + cq[1].x + cq[1].y;
}

template <typename Index>
__global__ __launch_bounds__(256) void k_fe_emit(const float* __restrict__ vol,
                                                 const uint64_t* __restrict__ mask, Geom g, float iso,
                                                 const RowCounts* __restrict__ counts,
                                                 const uint64_t* __restrict__ voff,
                                                 const uint64_t* __restrict__ toff,
                                                 float* __restrict__ pts, float* __restrict__ nrm,
                                                 Index* __restrict__ tris, uint64_t vbias = 0,
                                                 const uint64_t* __restrict__ need = nullptr, uint64_t cap_v = 0,
                                                 uint64_t cap_t = 0)
{
    if (need && !output_fits<Index>(need, cap_v, cap_t)) return;   // as k_fe_emit_rows
    __shared__ EmitLds lds_all[kRowsPerBlock];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + wave;
    if (r >= g.ny * g.nz || r < g.emit_lo || r >= g.emit_hi) return;
    const RowCounts c0 = counts[r];
    if (c0.x + c0.y + c0.z == 0 && c0.t == 0) return;
    const int64_t j = r % g.ny, k = r / g.ny;
    const bool has_y = j < g.ny - 1, has_z = k < g.nz - 1, cells = has_y && has_z && c0.t;
    RowWalker rw(lds_all[wave], mask, g, r, has_y, has_z);
    EmitLds& L = lds_all[wave];

    const uint64_t p0 = voff[r];
    uint64_t vb_run[3] = {p0, p0 + c0.x, p0 + c0.x + c0.y};
    uint64_t cbase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t tri_run = 0;
    if (cells) {
        combo_bases(counts, voff, rw.rows, cbase);
        tri_run = toff[r];
    }

    for (int64_t s0w = 0; s0w < g.W; s0w += kSegWords) {
        rw.stage(s0w);
        // ---- vertices of grid row r, already in output order
        int start[3], cnt[3];
        const int n = rw.compact_row_edges(has_y, has_z, start, cnt);
        for (int e = lane; e < n; e += 64) {
            const uint32_t ent = L.list[e];
            const int a = (int)(ent >> 14);
            const int64_t i = s0w * 64 + (int64_t)(ent & 0x3FFF);
            const uint64_t vid = vb_run[a] + (uint64_t)(e - start[a]);
            float A[3], B[3], p[3], nn[3];
            fe_edge_corners(g, a, i, j, k, A, B);
            edge_point(vol, g, iso, a, i, j, k, A, B, p, nn);
            pts[3 * vid + 0] = p[0];
            pts[3 * vid + 1] = p[1];
            pts[3 * vid + 2] = p[2];
            nrm[3 * vid + 0] = nn[0];
            nrm[3 * vid + 1] = nn[1];
            nrm[3 * vid + 2] = nn[2];
        }
        wave_sync();
        vb_run[0] += cnt[0];
        vb_run[1] += cnt[1];
        vb_run[2] += cnt[2];

        // ---- triangles of cell row (j,k)
        if (cells) {
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
                        const uint64_t a = rw.fe_vertex_id(cbase, case_edge(packed, 3 * t + 0), li);
                        const uint64_t b = rw.fe_vertex_id(cbase, case_edge(packed, 3 * t + 1), li);
                        const uint64_t c = rw.fe_vertex_id(cbase, case_edge(packed, 3 * t + 2), li);
                        store_tri<Index>(tris, slot0 + t, a - vbias, b - vbias, c - vbias);
                    }
                }
                tri_run += __shfl(incl, 63);
            }
            rw.advance_bases(cbase);
        }
        wave_sync();
    }
}

}  // namespace miso

#endif
