/*
 * capi.hip -- host side of libminiiso.so: the C ABI of include/miniiso.h over the HIP kernels.
 *
 * Flying Edges keeps the reference's four-pass state machine (flyingEdges/serial/
 * FlyingEdgesAlgorithm.h:27-49): pass 1 and 2 build the per-row state, pass 3 scans it and sizes
 * the output (the only host synchronisation, as in the reference where pass 3 allocates the
 * output vectors, FlyingEdgesAlgorithm.cpp:266-268), pass 4 writes points, normals and triangles.
 * Device buffers are allocated once per handle and grown only when an output does not fit.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "miniiso.h"
#include "fe_kernels.h"
#include "mc_kernels.h"
#include "fe_rows.h"
#include "fe_batch.h"
#include "mc_rows.h"

using namespace miso;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCALL(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? MISO_ERR_ALLOC : MISO_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                   \
    } while (0)

template <typename T>
int dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(MISO_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
    }
    return MISO_OK;
}

template <typename T>
void dev_free(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int check_volume(const miso_volume* v)
{
    if (!v) return fail(MISO_ERR_ARG, "volume is NULL");
    if (v->nx < 0 || v->ny < 0 || v->nz < 0) return fail(MISO_ERR_ARG, "negative dimension");
    if (v->nx > 0xFFFFFFFFll) return fail(MISO_ERR_ARG, "nx above 2^32 is not supported");
    if (!v->data && v->nx * v->ny * v->nz > 0) return fail(MISO_ERR_ARG, "volume data is NULL");
    return MISO_OK;
}

Geom make_geom(const miso_volume* v)
{
    Geom g{};
    g.nx = v->nx;
    g.ny = v->ny;
    g.nz = v->nz;
    g.W = (v->nx + 63) / 64;
    for (int a = 0; a < 3; ++a) {
        g.org[a] = v->origin[a];
        g.sp[a] = v->spacing[a];
    }
    g.kofs = 0;
    g.nzg = v->nz;
    g.emit_lo = 0;
    g.emit_hi = v->ny * v->nz;
    g.bo = plan_block_order(g.ny, g.nz, g.W);
    return g;
}

inline bool has_cells(const Geom& g) { return g.nx >= 2 && g.ny >= 2 && g.nz >= 2; }

/* Shared per-extraction device state: volume, sign masks, row counts, offsets, scan state. */
struct RowState {
    Geom g{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    const float* d_vol = nullptr;
    float* d_vol_own = nullptr;
    uint64_t* d_mask = nullptr;
    RowCounts* d_counts = nullptr;
    RowTrim* d_trims = nullptr;
    uint64_t* d_voff = nullptr;
    uint64_t* d_toff = nullptr;
    uint64_t* d_scan = nullptr;
    int64_t ntiles = 0;
    uint64_t* d_totals = nullptr;
    uint64_t* h_totals = nullptr;   // pinned

    int init(const miso_volume* v, void* user_stream)
    {
        g = make_geom(v);
        if (user_stream) {
            stream = (hipStream_t)user_stream;
        } else {
            // a blocking stream: ordered after work on the null stream (where callers that
            // produce a device volume without naming a stream usually ran)
            HIPCALL(hipStreamCreateWithFlags(&stream, hipStreamDefault));
            own_stream = true;
        }
        const int64_t N = g.nx * g.ny * g.nz;
        if (v->on_device || N == 0) {
            d_vol = v->data;
        } else {
            if (int rc = dev_alloc(&d_vol_own, (size_t)N)) return rc;
            HIPCALL(hipMemcpyAsync(d_vol_own, v->data, (size_t)N * sizeof(float), hipMemcpyHostToDevice, stream));
            HIPCALL(hipStreamSynchronize(stream));   // the host volume is only borrowed until here
            d_vol = d_vol_own;
        }
        HIPCALL(hipHostMalloc((void**)&h_totals, 2 * sizeof(uint64_t), hipHostMallocDefault));
        h_totals[0] = h_totals[1] = 0;
        if (!has_cells(g)) return MISO_OK;
        const int64_t R = g.ny * g.nz;
        ntiles = (R + kScanTile - 1) / kScanTile;
        int rc = 0;
        if ((rc = dev_alloc(&d_mask, (size_t)(R * g.W)))) return rc;
        if ((rc = dev_alloc(&d_counts, (size_t)R))) return rc;
        if ((rc = dev_alloc(&d_trims, (size_t)R))) return rc;
        if ((rc = dev_alloc(&d_voff, (size_t)R))) return rc;
        if ((rc = dev_alloc(&d_toff, (size_t)R))) return rc;
        if ((rc = dev_alloc(&d_scan, (size_t)(1 + 2 * ntiles)))) return rc;
        if ((rc = dev_alloc(&d_totals, 2))) return rc;
        return MISO_OK;
    }


    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        dev_free(d_vol_own);
        dev_free(d_mask);
        dev_free(d_counts);
        dev_free(d_trims);
        dev_free(d_voff);
        dev_free(d_toff);
        dev_free(d_scan);
        dev_free(d_totals);
        if (h_totals) (void)hipHostFree(h_totals);
        h_totals = nullptr;
        if (own_stream && stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    unsigned row_blocks() const { return (unsigned)((g.ny * g.nz + kRowsPerBlock - 1) / kRowsPerBlock); }

    /* unit spacing, zero origin: coordinates are float(index) and the gradient divisions by 1
     * and 2 are exact (index < 2^24 keeps float(index) exact) */
    bool unit_geometry() const
    {
        return g.sp[0] == 1.0f && g.sp[1] == 1.0f && g.sp[2] == 1.0f && g.org[0] == 0.0f && g.org[1] == 0.0f &&
               g.org[2] == 0.0f && g.nx < (1 << 24) && g.ny < (1 << 24) && g.nzg < (1 << 24);
    }

    /* (row, word) lane kernels: rows of up to 4096 voxels; the vertex ids a wavefront's
     * triangles reference (rows r0 .. r0 + 64 + ny + 1) within 2^32 of its first one; the voxels
     * its vertices read (planes k0 - 1 .. k0 + 64 / ny + 3) within 2^31 of plane k0 */
    bool lane_rows() const
    {
        const int64_t planes = (63 + g.ny) / g.ny + 3;
        return g.W <= 64 && (g.ny + 130) * 3 * g.nx < (int64_t(1) << 32) &&
               planes * g.nx * g.ny < (int64_t(1) << 31);
    }
    int rows_per_wave() const
    {
        int wp = 1;
        while (wp < g.W) wp <<= 1;
        return 64 / wp;
    }
    unsigned lane_row_blocks() const
    {
        const int64_t per_block = 4 * (int64_t)rows_per_wave();
        return (unsigned)((g.ny * g.nz + per_block - 1) / per_block);
    }

    int launch_classify(float iso)
    {
        hipLaunchKernelGGL(k_fe_classify, dim3(row_blocks()), dim3(256), 0, stream, d_vol, g, iso, d_mask);
        HIPCALL(hipGetLastError());
        return MISO_OK;
    }

    int launch_count()
    {
        if (lane_rows())
            hipLaunchKernelGGL(k_fe_count_rows, dim3(lane_row_blocks()), dim3(256), 0, stream,
                               (const uint64_t*)d_mask, g, d_counts, d_trims, d_scan, (int64_t)(1 + 2 * ntiles));
        else
            hipLaunchKernelGGL(k_fe_count, dim3(row_blocks()), dim3(256), 0, stream, (const uint64_t*)d_mask, g,
                               d_counts, d_trims, d_scan, (int64_t)(1 + 2 * ntiles));
        HIPCALL(hipGetLastError());
        return MISO_OK;
    }

    int launch_scan()
    {
        hipLaunchKernelGGL(k_fe_scan, dim3((unsigned)ntiles), dim3(256), 0, stream, (const RowCounts*)d_counts,
                           g.ny * g.nz, d_voff, d_toff, d_scan, ntiles, d_totals);
        HIPCALL(hipGetLastError());
        HIPCALL(hipMemcpyAsync(h_totals, d_totals, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
        return MISO_OK;
    }
};

/* Marching Cubes state on top of the Flying Edges row state. */
struct McState {
    uint32_t* d_owned = nullptr;     // owned cut edges per cell row (grid-row indexed)
    uint64_t* d_moff = nullptr;      // Marching Cubes vertex offset per cell row
    uint64_t* d_scan = nullptr;
    uint64_t* d_total = nullptr;
    int64_t ntiles = 0;
    void* d_map = nullptr;           // Flying Edges vertex id -> Marching Cubes vertex id
    uint64_t map_bytes = 0;
    int order = MISO_MC_ORDER_FIRST_VISIT;
    uint64_t* d_xyz = nullptr;       // edge order: per-row exclusive scans of the x, y, z counts
    uint64_t* d_axstate = nullptr;   // their three scan states, then the three totals

    int init(const Geom& g)
    {
        const int64_t R = g.ny * g.nz;
        ntiles = (R + kScanTile - 1) / kScanTile;
        int rc = 0;
        if ((rc = dev_alloc(&d_owned, (size_t)R))) return rc;
        if ((rc = dev_alloc(&d_moff, (size_t)R))) return rc;
        if ((rc = dev_alloc(&d_scan, (size_t)(1 + ntiles)))) return rc;
        if ((rc = dev_alloc(&d_total, 1))) return rc;
        return MISO_OK;
    }

    /* the edge-order buffers, allocated on first use */
    int init_edge_order(const Geom& g)
    {
        if (d_xyz) return MISO_OK;
        const int64_t R = g.ny * g.nz;
        if (int rc = dev_alloc(&d_xyz, (size_t)(3 * R))) return rc;
        return dev_alloc(&d_axstate, (size_t)(3 * (1 + ntiles) + 3));
    }

    void release()
    {
        dev_free(d_owned);
        dev_free(d_moff);
        dev_free(d_scan);
        dev_free(d_total);
        dev_free(d_xyz);
        dev_free(d_axstate);
        if (d_map) (void)hipFree(d_map);
        d_map = nullptr;
    }

    /* first-visit order with the row-lane kernels: the owned-edge counts come with pass 2
     * (k_fe_mc_count_rows), launched by launch_fe_count */
    bool fused_count(const RowState& s) const { return order != MISO_MC_ORDER_EDGE && s.lane_rows(); }

    /* pass 2 of the Flying Edges row state, with the owned-edge counts where fused_count */
    int launch_fe_count(RowState& s)
    {
        if (!fused_count(s)) return s.launch_count();
        hipLaunchKernelGGL(k_fe_mc_count_rows, dim3(s.lane_row_blocks()), dim3(256), 0, s.stream,
                           (const uint64_t*)s.d_mask, s.g, s.d_counts, s.d_trims, s.d_scan,
                           (int64_t)(1 + 2 * s.ntiles), d_owned, d_scan, (int64_t)(1 + ntiles));
        HIPCALL(hipGetLastError());
        return MISO_OK;
    }

    int launch_count(RowState& s)
    {
        if (order == MISO_MC_ORDER_EDGE) return MISO_OK;   // no first-visit offsets needed
        if (!fused_count(s)) {
            hipLaunchKernelGGL(k_mc_count, dim3(s.row_blocks()), dim3(256), 0, s.stream, (const uint64_t*)s.d_mask,
                               s.g, (const RowCounts*)s.d_counts, d_owned, d_scan, (int64_t)(1 + ntiles));
            HIPCALL(hipGetLastError());
        }
        hipLaunchKernelGGL(k_scan_u32, dim3((unsigned)ntiles), dim3(256), 0, s.stream, (const uint32_t*)d_owned,
                           s.g.ny * s.g.nz, d_moff, d_scan, ntiles, d_total);
        HIPCALL(hipGetLastError());
        return MISO_OK;
    }

    template <typename Index>
    int emit(RowState& s, float iso, float* pts, float* nrm, Index* tris, uint64_t nverts)
    {
        const uint64_t need = (nverts ? nverts : 1) * sizeof(Index);
        if (need > map_bytes) {
            if (d_map) (void)hipFree(d_map);
            d_map = nullptr;
            uint8_t* p = nullptr;
            if (int rc = dev_alloc(&p, (size_t)need)) return rc;
            d_map = p;
            map_bytes = need;
        }
        if (order == MISO_MC_ORDER_EDGE) {
            // vertices sorted by global edge id: the permutation from the three axis scans,
            // then owners write at perm[FE id]; triangles through the same permutation
            const int64_t R = s.g.ny * s.g.nz;
            const int64_t sw = 1 + ntiles;
            hipLaunchKernelGGL(k_batch_zero, dim3(64), dim3(256), 0, s.stream, d_axstate, 3 * sw + 3);
            for (int a = 0; a < 3; ++a)
                hipLaunchKernelGGL(k_scan_u32, dim3((unsigned)ntiles), dim3(256), 0, s.stream,
                                   (const uint32_t*)s.d_counts + a, R, d_xyz + a * R, d_axstate + a * sw, ntiles,
                                   d_axstate + 3 * sw + a, (int64_t)4);
            hipLaunchKernelGGL(k_edge_order_perm<Index>, dim3((unsigned)((R + 15) / 16)), dim3(256), 0, s.stream,
                               (const RowCounts*)s.d_counts, (const uint64_t*)s.d_voff, (const uint64_t*)d_xyz,
                               (const uint64_t*)(d_axstate + 3 * sw), R, (Index*)d_map);
            const uint64_t* no_moff = nullptr;
            if (s.lane_rows() && s.unit_geometry()) {
                // Flying Edges pass 4 through the permutation
                hipLaunchKernelGGL(k_mc_emit_rows<Index>, dim3(s.lane_row_blocks()), dim3(256), 0, s.stream, s.d_vol,
                                   (const uint64_t*)s.d_mask, s.g, iso, (const RowCounts*)s.d_counts,
                                   (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, (const Index*)d_map, pts,
                                   nrm, tris);
                HIPCALL(hipGetLastError());
                return MISO_OK;
            }
            if (!s.lane_rows()) {
                hipLaunchKernelGGL((k_mc_vertices<Index, true>), dim3(s.row_blocks()), dim3(256), 0, s.stream,
                                   s.d_vol, (const uint64_t*)s.d_mask, s.g, iso, (const RowCounts*)s.d_counts,
                                   (const uint64_t*)s.d_voff, no_moff, pts, nrm, (Index*)d_map);
                HIPCALL(hipGetLastError());
                hipLaunchKernelGGL(k_mc_triangles<Index>, dim3(s.row_blocks()), dim3(256), 0, s.stream,
                                   (const uint64_t*)s.d_mask, s.g, (const RowCounts*)s.d_counts,
                                   (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, (const Index*)d_map, tris);
                HIPCALL(hipGetLastError());
                return MISO_OK;
            }
            if (s.unit_geometry())
                hipLaunchKernelGGL((k_mc_vertices_rows<Index, true, true>), dim3(s.lane_row_blocks()), dim3(256), 0,
                                   s.stream, s.d_vol, (const uint64_t*)s.d_mask, s.g, iso,
                                   (const RowCounts*)s.d_counts, (const uint64_t*)s.d_voff, no_moff, pts, nrm,
                                   (Index*)d_map);
            else
                hipLaunchKernelGGL((k_mc_vertices_rows<Index, false, true>), dim3(s.lane_row_blocks()), dim3(256), 0,
                                   s.stream, s.d_vol, (const uint64_t*)s.d_mask, s.g, iso,
                                   (const RowCounts*)s.d_counts, (const uint64_t*)s.d_voff, no_moff, pts, nrm,
                                   (Index*)d_map);
            HIPCALL(hipGetLastError());
            hipLaunchKernelGGL(k_mc_triangles_rows<Index>, dim3(s.lane_row_blocks()), dim3(256), 0, s.stream,
                               (const uint64_t*)s.d_mask, s.g, (const RowCounts*)s.d_counts,
                               (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, (const Index*)d_map, tris);
            HIPCALL(hipGetLastError());
            return MISO_OK;
        }
        if (s.lane_rows()) {
            // (the map alone followed by k_mc_emit_rows, as edge order runs, measures the same here:
            // first-visit ids scatter the Flying Edges emit's vertex writes)
            if (s.unit_geometry())
                hipLaunchKernelGGL((k_mc_vertices_rows<Index, true>), dim3(s.lane_row_blocks()), dim3(256), 0, s.stream,
                                   s.d_vol, (const uint64_t*)s.d_mask, s.g, iso, (const RowCounts*)s.d_counts,
                                   (const uint64_t*)s.d_voff, (const uint64_t*)d_moff, pts, nrm, (Index*)d_map);
            else
                hipLaunchKernelGGL((k_mc_vertices_rows<Index, false>), dim3(s.lane_row_blocks()), dim3(256), 0,
                                   s.stream, s.d_vol, (const uint64_t*)s.d_mask, s.g, iso, (const RowCounts*)s.d_counts,
                                   (const uint64_t*)s.d_voff, (const uint64_t*)d_moff, pts, nrm, (Index*)d_map);
            HIPCALL(hipGetLastError());
            hipLaunchKernelGGL(k_mc_triangles_rows<Index>, dim3(s.lane_row_blocks()), dim3(256), 0, s.stream,
                               (const uint64_t*)s.d_mask, s.g, (const RowCounts*)s.d_counts,
                               (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, (const Index*)d_map, tris);
            HIPCALL(hipGetLastError());
            return MISO_OK;
        }
        hipLaunchKernelGGL(k_mc_vertices<Index>, dim3(s.row_blocks()), dim3(256), 0, s.stream, s.d_vol,
                           (const uint64_t*)s.d_mask, s.g, iso, (const RowCounts*)s.d_counts,
                           (const uint64_t*)s.d_voff, (const uint64_t*)d_moff, pts, nrm, (Index*)d_map);
        HIPCALL(hipGetLastError());
        hipLaunchKernelGGL(k_mc_triangles<Index>, dim3(s.row_blocks()), dim3(256), 0, s.stream,
                           (const uint64_t*)s.d_mask, s.g, (const RowCounts*)s.d_counts, (const uint64_t*)s.d_voff,
                           (const uint64_t*)s.d_toff, (const Index*)d_map, tris);
        HIPCALL(hipGetLastError());
        return MISO_OK;
    }

    int launch_emit(RowState& s, float iso, float* pts, float* nrm, void* tris, int index_bits, uint64_t nverts)
    {
        if (index_bits == 64) return emit<uint64_t>(s, iso, pts, nrm, (uint64_t*)tris, nverts);
        return emit<uint32_t>(s, iso, pts, nrm, (uint32_t*)tris, nverts);
    }
};

/* Output buffers, grown on demand. */
struct MeshBuffers {
    float* d_pts = nullptr;
    float* d_nrm = nullptr;
    void* d_tris = nullptr;
    uint64_t cap_v = 0, cap_t = 0;
    int cap_t_bits = 0;
    uint64_t nverts = 0, ntris = 0;
    int index_bits = 32;
    bool force64 = false;

    int reserve(uint64_t v, uint64_t t)
    {
        nverts = v;
        ntris = t;
        index_bits = (force64 || v > 0xFFFFFFFFull) ? 64 : 32;
        if (v > cap_v || !d_pts) {
            dev_free(d_pts);
            dev_free(d_nrm);
            if (int rc = dev_alloc(&d_pts, (size_t)(3 * v))) return rc;
            if (int rc = dev_alloc(&d_nrm, (size_t)(3 * v))) return rc;
            cap_v = v;
        }
        const uint64_t need = 3 * t * (uint64_t)(index_bits / 8);
        if (need > cap_t * (uint64_t)(cap_t_bits / 8) * 3 || !d_tris) {
            if (d_tris) (void)hipFree(d_tris);
            d_tris = nullptr;
            uint8_t* p = nullptr;
            if (int rc = dev_alloc(&p, (size_t)need)) return rc;
            d_tris = p;
            cap_t = t;
            cap_t_bits = index_bits;
        }
        return MISO_OK;
    }

    void release()
    {
        dev_free(d_pts);
        dev_free(d_nrm);
        if (d_tris) (void)hipFree(d_tris);
        d_tris = nullptr;
    }

    int copy_out(hipStream_t s, float* pts, float* nrm, uint64_t* tris)
    {
        if (pts && nverts) HIPCALL(hipMemcpyAsync(pts, d_pts, 3 * nverts * sizeof(float), hipMemcpyDeviceToHost, s));
        if (nrm && nverts) HIPCALL(hipMemcpyAsync(nrm, d_nrm, 3 * nverts * sizeof(float), hipMemcpyDeviceToHost, s));
        if (tris && ntris) {
            if (index_bits == 64) {
                HIPCALL(hipMemcpyAsync(tris, d_tris, 3 * ntris * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
                HIPCALL(hipStreamSynchronize(s));
            } else {
                // widen 32-bit indices to the reference's size_t in place: copy into the upper
                // half of the caller's buffer, then expand from the front (no overlap hazard)
                uint32_t* tmp = reinterpret_cast<uint32_t*>(tris) + 3 * ntris;
                HIPCALL(hipMemcpyAsync(tmp, d_tris, 3 * ntris * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                HIPCALL(hipStreamSynchronize(s));
                for (uint64_t q = 0; q < 3 * ntris; ++q) tris[q] = tmp[q];
            }
        }
        HIPCALL(hipStreamSynchronize(s));
        return MISO_OK;
    }
};

/* true case of each cell, or the reference's trimmed convention (0 outside [xl, xr)) */
__global__ void k_fe_cube_cases(const uint64_t* __restrict__ mask, Geom g, const RowTrim* __restrict__ trims,
                                uint8_t* __restrict__ out, int all_cells)
{
    const int64_t cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
    const int64_t n = cx * cy * cz;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = c % cx, j = (c / cx) % cy, k = c / (cx * cy);
        const int64_t r = j + g.ny * k;
        const int64_t rows[4] = {r, r + 1, r + g.ny, r + g.ny + 1};
        if (!all_cells) {
            // calcTrimValues (FlyingEdgesAlgorithm.cpp:606-621)
            uint32_t xl = trims[rows[0]].xl, xr = trims[rows[0]].xr;
            for (int q = 1; q < 4; ++q) {
                xl = min(xl, trims[rows[q]].xl);
                xr = max(xr, trims[rows[q]].xr);
            }
            if (xl > xr) xl = xr;
            if (i < (int64_t)xl || i >= (int64_t)xr) {
                out[c] = 0;
                continue;
            }
        }
        auto bit = [&](int q, int64_t x) -> int {
            return (int)((mask[rows[q] * g.W + (x >> 6)] >> (x & 63)) & 1ull);
        };
        out[c] = (uint8_t)(bit(0, i) | bit(0, i + 1) << 1 | bit(1, i + 1) << 2 | bit(1, i) << 3 |
                           bit(2, i) << 4 | bit(2, i + 1) << 5 | bit(3, i + 1) << 6 | bit(3, i) << 7);
    }
}

}  // namespace

/* ====================================================================== Flying Edges */

struct miso_fe {
    RowState st;
    MeshBuffers out;
    float iso = 0.0f;
    int done = 0;                    // last completed pass
    hipEvent_t ev[8] = {};           // start/stop of passes 1..4
    bool events = false;
    unsigned timed = 0u;             // bit p - 1: pass p's device time is recorded with events
    float pass_ms[4] = {0, 0, 0, 0};
};

namespace {

/* need: pass 4 launched before its output size is known on the host (pass 3's totals still on the
 * device), into the buffers of an earlier run; the kernel writes nothing if they do not fit. */
int fe_pass_launch(miso_fe* h, int pass, const uint64_t* need = nullptr)
{
    RowState& s = h->st;
    if (!has_cells(s.g)) return MISO_OK;
    const bool timed = pass >= 1 && pass <= 4 && ((h->timed >> (pass - 1)) & 1u);
    if (timed) HIPCALL(hipEventRecord(h->ev[2 * (pass - 1)], s.stream));
    switch (pass) {
    case 1:
        if (int rc = s.launch_classify(h->iso)) return rc;
        break;
    case 2:
        if (int rc = s.launch_count()) return rc;
        break;
    case 3:
        if (int rc = s.launch_scan()) return rc;
        break;
    case 4: {
        const MeshBuffers& o = h->out;
        const bool work = need || o.nverts || o.ntris;
        const uint64_t cap_t = o.index_bits ? o.cap_t * (uint64_t)o.cap_t_bits / (uint64_t)o.index_bits : 0;
        if (work && s.lane_rows()) {
#define MISO_EMIT_ROWS(IDX, UNIT)                                                                                  \
    hipLaunchKernelGGL((k_fe_emit_rows<IDX, UNIT>), dim3(s.lane_row_blocks()), dim3(256), 0, s.stream, s.d_vol,    \
                       (const uint64_t*)s.d_mask, s.g, h->iso, (const RowCounts*)s.d_counts,                     \
                       (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, o.d_pts, o.d_nrm, (IDX*)o.d_tris,   \
                       (uint64_t)0, need, o.cap_v, cap_t)
            const bool unit = s.unit_geometry();
            if (o.index_bits == 64) {
                if (unit) MISO_EMIT_ROWS(uint64_t, true);
                else MISO_EMIT_ROWS(uint64_t, false);
            } else {
                if (unit) MISO_EMIT_ROWS(uint32_t, true);
                else MISO_EMIT_ROWS(uint32_t, false);
            }
#undef MISO_EMIT_ROWS
            HIPCALL(hipGetLastError());
        } else if (work) {
            if (o.index_bits == 64)
                hipLaunchKernelGGL(k_fe_emit<uint64_t>, dim3(s.row_blocks()), dim3(256), 0, s.stream, s.d_vol,
                                   (const uint64_t*)s.d_mask, s.g, h->iso, (const RowCounts*)s.d_counts,
                                   (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, o.d_pts, o.d_nrm,
                                   (uint64_t*)o.d_tris, (uint64_t)0, need, o.cap_v, cap_t);
            else
                hipLaunchKernelGGL(k_fe_emit<uint32_t>, dim3(s.row_blocks()), dim3(256), 0, s.stream, s.d_vol,
                                   (const uint64_t*)s.d_mask, s.g, h->iso, (const RowCounts*)s.d_counts,
                                   (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, o.d_pts, o.d_nrm,
                                   (uint32_t*)o.d_tris, (uint64_t)0, need, o.cap_v, cap_t);
            HIPCALL(hipGetLastError());
        }
        break;
    }
    default:
        return fail(MISO_ERR_ARG, "pass must be 1..4");
    }
    if (timed) HIPCALL(hipEventRecord(h->ev[2 * (pass - 1) + 1], s.stream));
    return MISO_OK;
}

int fe_after_scan(miso_fe* h)
{
    RowState& s = h->st;
    HIPCALL(hipStreamSynchronize(s.stream));
    uint64_t v = 0, t = 0;
    if (has_cells(s.g)) {
        v = s.h_totals[0];
        t = s.h_totals[1];
    }
    return h->out.reserve(v, t);
}

int fe_collect_times(miso_fe* h)
{
    if (!has_cells(h->st.g)) return MISO_OK;
    for (int p = 0; p < 4; ++p) {
        h->pass_ms[p] = 0.0f;
        if ((h->timed >> p) & 1u) HIPCALL(hipEventElapsedTime(&h->pass_ms[p], h->ev[2 * p], h->ev[2 * p + 1]));
    }
    return MISO_OK;
}

}  // namespace

extern "C" int miso_fe_create(const miso_volume* vol, float isoval, void* stream, miso_fe** out)
{
    if (!out) return fail(MISO_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (int rc = check_volume(vol)) return rc;
    miso_fe* h = new (std::nothrow) miso_fe();
    if (!h) return fail(MISO_ERR_ALLOC, "host allocation failed");
    h->iso = isoval;
    int rc = h->st.init(vol, stream);
    if (rc == MISO_OK) {
        for (int e = 0; e < 8 && rc == MISO_OK; ++e)
            if (hipEventCreate(&h->ev[e]) != hipSuccess) rc = fail(MISO_ERR_HIP, "hipEventCreate failed");
        h->events = rc == MISO_OK;
    }
    if (rc != MISO_OK) {
        miso_fe_destroy(h);
        return rc;
    }
    *out = h;
    return MISO_OK;
}

extern "C" int miso_fe_set_isoval(miso_fe* h, float isoval)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    h->iso = isoval;
    h->done = 0;
    return MISO_OK;
}

extern "C" int miso_fe_pass(miso_fe* h, int pass)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (pass < 1 || pass > 4) return fail(MISO_ERR_ARG, "pass must be 1..4");
    if (pass != h->done + 1 && !(pass == 1)) return fail(MISO_ERR_STATE, "pass %d called after pass %d", pass, h->done);
    if (int rc = fe_pass_launch(h, pass)) return rc;
    if (pass == 3) {
        if (int rc = fe_after_scan(h)) return rc;
    } else {
        HIPCALL(hipStreamSynchronize(h->st.stream));
    }
    if (pass == 4) fe_collect_times(h);
    h->done = pass;
    return MISO_OK;
}

extern "C" int miso_fe_run(miso_fe* h)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    for (int p = 1; p <= 3; ++p)
        if (int rc = fe_pass_launch(h, p)) return rc;
    MeshBuffers& o = h->out;
    if (has_cells(h->st.g) && o.d_pts && o.d_tris && o.index_bits) {
        // Buffers from an earlier run: pass 4 follows pass 3 at once, without a host round trip,
        // and checks on the device that the totals fit. Only if they do not (a larger mesh, or
        // other index width) are the buffers grown and pass 4 run again.
        if (int rc = fe_pass_launch(h, 4, h->st.d_totals)) return rc;
        HIPCALL(hipStreamSynchronize(h->st.stream));
        const uint64_t v = h->st.h_totals[0], t = h->st.h_totals[1];
        const int bits = (o.force64 || v > 0xFFFFFFFFull) ? 64 : 32;
        const uint64_t cap_t = o.cap_t * (uint64_t)o.cap_t_bits / (uint64_t)o.index_bits;
        if (v <= o.cap_v && t <= cap_t && bits == o.index_bits) {
            o.nverts = v;
            o.ntris = t;
            h->done = 4;
            return fe_collect_times(h);
        }
    }
    if (int rc = fe_after_scan(h)) return rc;
    if (int rc = fe_pass_launch(h, 4)) return rc;
    HIPCALL(hipStreamSynchronize(h->st.stream));
    h->done = 4;
    return fe_collect_times(h);
}

extern "C" int miso_fe_counts(const miso_fe* h, uint64_t* nverts, uint64_t* ntris)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (h->done < 3) return fail(MISO_ERR_STATE, "counts are known after pass 3");
    if (nverts) *nverts = h->out.nverts;
    if (ntris) *ntris = h->out.ntris;
    return MISO_OK;
}

extern "C" int miso_fe_copy_out(miso_fe* h, float* points, float* normals, uint64_t* triangles)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (h->done < 4) return fail(MISO_ERR_STATE, "the mesh is complete after pass 4");
    return h->out.copy_out(h->st.stream, points, normals, triangles);
}

extern "C" int miso_fe_device_output(miso_fe* h, const float** points, const float** normals,
                                     const void** triangles, int* index_bits)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (h->done < 4) return fail(MISO_ERR_STATE, "the mesh is complete after pass 4");
    if (points) *points = h->out.d_pts;
    if (normals) *normals = h->out.d_nrm;
    if (triangles) *triangles = h->out.d_tris;
    if (index_bits) *index_bits = h->out.index_bits;
    return MISO_OK;
}

extern "C" int miso_fe_set_index_bits(miso_fe* h, int index_bits)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (index_bits != 32 && index_bits != 64) return fail(MISO_ERR_ARG, "index_bits must be 32 or 64");
    h->out.force64 = index_bits == 64;
    return MISO_OK;
}

extern "C" int miso_fe_set_timing(miso_fe* h, unsigned passes)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (passes > 0xFu) return fail(MISO_ERR_ARG, "passes is a mask of bits 0..3");
    h->timed = passes;
    return MISO_OK;
}

extern "C" int miso_fe_pass_times(const miso_fe* h, float ms[4])
{
    if (!h || !ms) return fail(MISO_ERR_ARG, "NULL argument");
    for (int p = 0; p < 4; ++p) ms[p] = h->pass_ms[p];
    return MISO_OK;
}

extern "C" int miso_fe_rows(const miso_fe* h, uint32_t* xl, uint32_t* xr, uint32_t* xcount, uint32_t* ycount,
                            uint32_t* zcount, uint32_t* tcount, uint64_t* voffset, uint64_t* toffset)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (h->done < 3) return fail(MISO_ERR_STATE, "row state is complete after pass 3");
    const RowState& s = h->st;
    if (!has_cells(s.g)) return MISO_OK;
    const size_t R = (size_t)(s.g.ny * s.g.nz);
    std::vector<RowCounts> c(R);
    std::vector<RowTrim> t(R);
    HIPCALL(hipMemcpy(c.data(), s.d_counts, R * sizeof(RowCounts), hipMemcpyDeviceToHost));
    HIPCALL(hipMemcpy(t.data(), s.d_trims, R * sizeof(RowTrim), hipMemcpyDeviceToHost));
    if (voffset) HIPCALL(hipMemcpy(voffset, s.d_voff, R * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (toffset) HIPCALL(hipMemcpy(toffset, s.d_toff, R * sizeof(uint64_t), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < R; ++r) {
        if (xl) xl[r] = t[r].xl;
        if (xr) xr[r] = t[r].xr;
        if (xcount) xcount[r] = c[r].x;
        if (ycount) ycount[r] = c[r].y;
        if (zcount) zcount[r] = c[r].z;
        if (tcount) tcount[r] = c[r].t;
    }
    return MISO_OK;
}

extern "C" int miso_fe_cube_cases(miso_fe* h, uint8_t* cases, int all_cells)
{
    if (!h || !cases) return fail(MISO_ERR_ARG, "NULL argument");
    if (h->done < 2) return fail(MISO_ERR_STATE, "cube cases are known after pass 2");
    const RowState& s = h->st;
    if (!has_cells(s.g)) return MISO_OK;
    const size_t n = (size_t)((s.g.nx - 1) * (s.g.ny - 1) * (s.g.nz - 1));
    uint8_t* d = nullptr;
    if (int rc = dev_alloc(&d, n)) return rc;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fe_cube_cases, dim3(blocks), dim3(256), 0, s.stream, (const uint64_t*)s.d_mask, s.g,
                       (const RowTrim*)s.d_trims, d, all_cells);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cases, d, n, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(MISO_ERR_HIP, "cube case dump failed: %s", hipGetErrorString(e));
    return MISO_OK;
}

extern "C" void miso_fe_destroy(miso_fe* h)
{
    if (!h) return;
    h->st.release();
    h->out.release();
    for (int e = 0; e < 8; ++e)
        if (h->ev[e]) (void)hipEventDestroy(h->ev[e]);
    delete h;
}

/* ====================================================================== Flying Edges over z-slabs */

/* The volume in slabs of `planes` grid planes (the section decomposition with ghost planes of
 * marchingCubes/mpi/mpiutil.h:87-221, here along z only and run one slab after another on one
 * GPU). A slab owns the grid rows of planes [k0, k1): their vertices and the triangles of their
 * cell rows. It reads planes [k0 - 1, k1 + 2): one ghost plane below for the gradient, and two
 * above, for the z-edges and cells of plane k1 - 1 and the gradient at plane k1. The rows of
 * plane k1 are counted again (the next slab owns them) because the slab's last cell row names
 * their vertices. Ids continue from slab to slab, so the parts concatenate to the whole-volume
 * mesh, bit for bit.
 *
 * Two slots, each with its own row state and stream, alternate: while slab s writes its part,
 * slab s + 1 is classified, counted and scanned on the other slot's stream (and, for a host
 * volume, copied into that slot's staging buffer first). */
struct miso_fe_slabs {
    Geom full{};
    int64_t planes = 0;
    float iso = 0.0f;
    const float* src = nullptr;       // the whole volume, host or device
    bool src_on_device = false;
    RowState st[2];                   // per slot, each sized for planes + 3 data planes
    float* d_stage[2] = {nullptr, nullptr};
    hipEvent_t copied[2] = {};        // a slot's staging buffer holds its slab's planes
    hipEvent_t emitted[2] = {};       // a slot's part is written (its buffers may be reused)
    bool emit_pending[2] = {false, false};
    uint64_t* h_vown = nullptr;       // pinned, one per slot: the slab's owned vertex count
    hipStream_t copy_stream = nullptr;
    MeshBuffers out;                  // the current slab's part (64-bit ids)
    int64_t slot_k0[2] = {0, 0}, slot_k1[2] = {0, 0};
    int64_t counted_k1 = 0;           // planes whose counts have been launched end here
    int64_t slab = -1;                // the current slab (its part is in `out`)
    int64_t k0 = 0, k1 = 0;
    uint64_t vbase = 0, tbase = 0, vnext = 0, tnext = 0;
    bool finished = false;
};

namespace {

/* Rows below lo belong to the previous slab: zeroed, so the slab's offsets start at 0. Rows from
 * hi on keep their vertex counts (the slab's last cell row names those vertices) but their cells
 * belong to the next slab: their triangle counts are zeroed. */
__global__ __launch_bounds__(256) void k_slab_trim_counts(RowCounts* __restrict__ counts, int64_t R, int64_t lo,
                                                          int64_t hi)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < R; r += (int64_t)gridDim.x * 256) {
        if (r < lo) counts[r] = RowCounts{0u, 0u, 0u, 0u};
        else if (r >= hi) counts[r].t = 0u;
    }
}

int64_t slab_data_lo(int64_t k0) { return k0 > 0 ? k0 - 1 : 0; }
int64_t slab_data_hi(const miso_fe_slabs* h, int64_t k1) { return std::min<int64_t>(k1 + 2, h->full.nz); }

/* Passes 1-3 of the slab owning planes [k0, min(k0 + planes, nz)) on slot p, with the D2H copies
 * of its totals; the slot's previous work (the part it last wrote) must have finished. */
int slab_launch_counts(miso_fe_slabs* h, int p, int64_t k0)
{
    RowState& s = h->st[p];
    const Geom& G = h->full;
    const int64_t k1 = std::min<int64_t>(k0 + h->planes, G.nz);
    const int64_t d0 = slab_data_lo(k0), d1 = slab_data_hi(h, k1);
    const int64_t plane = G.nx * G.ny;
    if (h->src_on_device) {
        s.d_vol = h->src + d0 * plane;
    } else {
        HIPCALL(hipMemcpyAsync(h->d_stage[p], h->src + d0 * plane, (size_t)((d1 - d0) * plane) * sizeof(float),
                               hipMemcpyHostToDevice, h->copy_stream));
        HIPCALL(hipEventRecord(h->copied[p], h->copy_stream));
        HIPCALL(hipStreamWaitEvent(s.stream, h->copied[p], 0));
        s.d_vol = h->d_stage[p];
    }
    Geom g = G;
    g.nz = d1 - d0;
    g.kofs = d0;
    g.nzg = G.nz;
    g.emit_lo = (k0 - d0) * G.ny;
    g.emit_hi = (k1 - d0) * G.ny;
    g.bo = plan_block_order(g.ny, g.nz, g.W);
    s.g = g;
    const int64_t R = g.ny * g.nz;
    s.ntiles = (R + kScanTile - 1) / kScanTile;
    if (int rc = s.launch_classify(h->iso)) return rc;
    if (int rc = s.launch_count()) return rc;
    hipLaunchKernelGGL(k_slab_trim_counts, dim3((unsigned)std::min<int64_t>((R + 255) / 256, 4096)), dim3(256), 0,
                       s.stream, s.d_counts, R, g.emit_lo, g.emit_hi);
    HIPCALL(hipGetLastError());
    if (int rc = s.launch_scan()) return rc;
    if (g.emit_hi < R)
        HIPCALL(hipMemcpyAsync(h->h_vown + p, s.d_voff + g.emit_hi, sizeof(uint64_t), hipMemcpyDeviceToHost, s.stream));
    h->slot_k0[p] = k0;
    h->slot_k1[p] = k1;
    h->counted_k1 = k1;
    return MISO_OK;
}

/* Pass 4 of the slab counted on slot p, into `out`, once its counts are back. */
int slab_emit(miso_fe_slabs* h, int p)
{
    RowState& s = h->st[p];
    HIPCALL(hipStreamSynchronize(s.stream));
    const Geom& g = s.g;
    const int64_t R = g.ny * g.nz;
    const uint64_t nv = g.emit_hi < R ? h->h_vown[p] : s.h_totals[0];
    const uint64_t nt = s.h_totals[1];
    // the previous part (written on the other slot) is no longer needed; its kernels must be done
    // before its buffers are reused or reallocated
    if (h->emit_pending[p ^ 1]) {
        HIPCALL(hipEventSynchronize(h->emitted[p ^ 1]));
        h->emit_pending[p ^ 1] = false;
    }
    MeshBuffers& o = h->out;
    if (nv > o.cap_v || nt > o.cap_t || !o.d_pts || !o.d_tris) {
        // grow with slack, so that slabs of similar size do not reallocate each time
        if (int rc = o.reserve(nv + nv / 4 + 64, nt + nt / 4 + 64)) return rc;
    }
    o.nverts = nv;
    o.ntris = nt;
    o.index_bits = 64;
    if (nv || nt) {
        const uint64_t vbias = (uint64_t)0 - h->vnext;   // ids continue from the previous slabs
        if (s.lane_rows()) {
            if (s.unit_geometry())
                hipLaunchKernelGGL((k_fe_emit_rows<uint64_t, true>), dim3(s.lane_row_blocks()), dim3(256), 0,
                                   s.stream, s.d_vol, (const uint64_t*)s.d_mask, s.g, h->iso,
                                   (const RowCounts*)s.d_counts, (const uint64_t*)s.d_voff,
                                   (const uint64_t*)s.d_toff, o.d_pts, o.d_nrm, (uint64_t*)o.d_tris, vbias);
            else
                hipLaunchKernelGGL((k_fe_emit_rows<uint64_t, false>), dim3(s.lane_row_blocks()), dim3(256), 0,
                                   s.stream, s.d_vol, (const uint64_t*)s.d_mask, s.g, h->iso,
                                   (const RowCounts*)s.d_counts, (const uint64_t*)s.d_voff,
                                   (const uint64_t*)s.d_toff, o.d_pts, o.d_nrm, (uint64_t*)o.d_tris, vbias);
        } else {
            hipLaunchKernelGGL(k_fe_emit<uint64_t>, dim3(s.row_blocks()), dim3(256), 0, s.stream, s.d_vol,
                               (const uint64_t*)s.d_mask, s.g, h->iso, (const RowCounts*)s.d_counts,
                               (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, o.d_pts, o.d_nrm,
                               (uint64_t*)o.d_tris, vbias);
        }
        HIPCALL(hipGetLastError());
    }
    HIPCALL(hipEventRecord(h->emitted[p], s.stream));
    h->emit_pending[p] = true;
    h->slab += 1;
    h->k0 = h->slot_k0[p];
    h->k1 = h->slot_k1[p];
    h->vbase = h->vnext;
    h->tbase = h->tnext;
    h->vnext += nv;
    h->tnext += nt;
    return MISO_OK;
}

/* waits for the current part */
int slab_part_ready(miso_fe_slabs* h)
{
    const int p = (int)(h->slab & 1);
    if (h->emit_pending[p]) {
        HIPCALL(hipEventSynchronize(h->emitted[p]));
        h->emit_pending[p] = false;
    }
    return MISO_OK;
}

}  // namespace

extern "C" void miso_fe_slabs_destroy(miso_fe_slabs* h)
{
    if (!h) return;
    if (h->copy_stream) (void)hipStreamSynchronize(h->copy_stream);
    for (int p = 0; p < 2; ++p) h->st[p].release();
    h->out.release();
    for (int p = 0; p < 2; ++p) {
        dev_free(h->d_stage[p]);
        if (h->copied[p]) (void)hipEventDestroy(h->copied[p]);
        if (h->emitted[p]) (void)hipEventDestroy(h->emitted[p]);
    }
    if (h->copy_stream) (void)hipStreamDestroy(h->copy_stream);
    if (h->h_vown) (void)hipHostFree(h->h_vown);
    delete h;
}

extern "C" int miso_fe_slabs_create(const miso_volume* vol, float isoval, int64_t planes, void* stream,
                                    miso_fe_slabs** out)
{
    if (!out) return fail(MISO_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (int rc = check_volume(vol)) return rc;
    if (planes < 1) return fail(MISO_ERR_ARG, "planes per slab must be at least 1");
    miso_fe_slabs* h = new (std::nothrow) miso_fe_slabs();
    if (!h) return fail(MISO_ERR_ALLOC, "host allocation failed");
    h->full = make_geom(vol);
    h->planes = std::min<int64_t>(planes, std::max<int64_t>(vol->nz, 1));
    h->iso = isoval;
    h->src = vol->data;
    h->src_on_device = vol->on_device != 0;
    h->finished = !has_cells(h->full);
    // row state for the largest slab: planes + 3 data planes (never more than the volume)
    miso_volume sv = *vol;
    sv.nz = std::min<int64_t>(h->planes + 3, vol->nz);
    sv.data = nullptr;
    sv.on_device = 1;
    int rc = MISO_OK;
    for (int p = 0; p < 2 && rc == MISO_OK; ++p) {
        rc = h->st[p].init(&sv, stream);
        if (rc == MISO_OK && hipEventCreateWithFlags(&h->emitted[p], hipEventDisableTiming) != hipSuccess)
            rc = fail(MISO_ERR_HIP, "hipEventCreate failed");
    }
    const int64_t stage = sv.nz * vol->nx * vol->ny;
    if (rc == MISO_OK && !h->src_on_device && !h->finished) {
        for (int p = 0; p < 2 && rc == MISO_OK; ++p) {
            rc = dev_alloc(&h->d_stage[p], (size_t)stage);
            if (rc == MISO_OK && hipEventCreateWithFlags(&h->copied[p], hipEventDisableTiming) != hipSuccess)
                rc = fail(MISO_ERR_HIP, "hipEventCreate failed");
        }
        if (rc == MISO_OK && hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking) != hipSuccess)
            rc = fail(MISO_ERR_HIP, "hipStreamCreate failed");
    }
    if (rc == MISO_OK && hipHostMalloc((void**)&h->h_vown, 2 * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess)
        rc = fail(MISO_ERR_ALLOC, "hipHostMalloc failed");
    // the first slab's passes 1-3 start at once
    if (rc == MISO_OK && !h->finished) rc = slab_launch_counts(h, 0, 0);
    if (rc != MISO_OK) {
        miso_fe_slabs_destroy(h);
        return rc;
    }
    h->out.force64 = true;
    *out = h;
    return MISO_OK;
}

extern "C" int miso_fe_slabs_next(miso_fe_slabs* h, int* done)
{
    if (!h || !done) return fail(MISO_ERR_ARG, "NULL argument");
    *done = 0;
    if (h->finished || (h->slab >= 0 && h->k1 >= h->full.nz)) {
        h->finished = true;
        *done = 1;
        return MISO_OK;
    }
    const int p = (int)((h->slab + 1) & 1);      // the slot whose counts are in flight
    if (int rc = slab_emit(h, p)) return rc;
    // the next slab's passes 1-3 on the other slot, behind its last part (already waited for)
    if (h->counted_k1 < h->full.nz)
        if (int rc = slab_launch_counts(h, p ^ 1, h->counted_k1)) return rc;
    return MISO_OK;
}

extern "C" int miso_fe_slabs_restart(miso_fe_slabs* h, float isoval)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    for (int p = 0; p < 2; ++p) HIPCALL(hipStreamSynchronize(h->st[p].stream));
    if (h->copy_stream) HIPCALL(hipStreamSynchronize(h->copy_stream));
    h->emit_pending[0] = h->emit_pending[1] = false;
    h->iso = isoval;
    h->slab = -1;
    h->k0 = h->k1 = 0;
    h->vbase = h->tbase = h->vnext = h->tnext = 0;
    h->counted_k1 = 0;
    h->finished = !has_cells(h->full);
    if (!h->finished)
        if (int rc = slab_launch_counts(h, 0, 0)) return rc;
    return MISO_OK;
}

extern "C" int miso_fe_slabs_part(const miso_fe_slabs* h, int64_t* k0, int64_t* k1, uint64_t* vbase,
                                  uint64_t* nverts, uint64_t* tbase, uint64_t* ntris)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (h->slab < 0) return fail(MISO_ERR_STATE, "call miso_fe_slabs_next first");
    if (k0) *k0 = h->k0;
    if (k1) *k1 = h->k1;
    if (vbase) *vbase = h->vbase;
    if (nverts) *nverts = h->out.nverts;
    if (tbase) *tbase = h->tbase;
    if (ntris) *ntris = h->out.ntris;
    return MISO_OK;
}

extern "C" int miso_fe_slabs_copy_part(miso_fe_slabs* h, float* points, float* normals, uint64_t* triangles)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (h->slab < 0) return fail(MISO_ERR_STATE, "call miso_fe_slabs_next first");
    if (int rc = slab_part_ready(h)) return rc;
    return h->out.copy_out(h->st[h->slab & 1].stream, points, normals, triangles);
}

extern "C" int miso_fe_slabs_device_part(miso_fe_slabs* h, const float** points, const float** normals,
                                         const uint64_t** triangles)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (h->slab < 0) return fail(MISO_ERR_STATE, "call miso_fe_slabs_next first");
    if (int rc = slab_part_ready(h)) return rc;
    if (points) *points = h->out.d_pts;
    if (normals) *normals = h->out.d_nrm;
    if (triangles) *triangles = (const uint64_t*)h->out.d_tris;
    return MISO_OK;
}

/* ====================================================================== Marching Cubes */

struct miso_mc {
    RowState st;
    MeshBuffers out;
    float iso = 0.0f;
    bool done = false;
    McState mc;
    hipEvent_t ev[2] = {};
    float ms = 0.0f;
};

extern "C" int miso_mc_create(const miso_volume* vol, float isoval, void* stream, miso_mc** out)
{
    if (!out) return fail(MISO_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (int rc = check_volume(vol)) return rc;
    miso_mc* h = new (std::nothrow) miso_mc();
    if (!h) return fail(MISO_ERR_ALLOC, "host allocation failed");
    h->iso = isoval;
    int rc = h->st.init(vol, stream);
    if (rc == MISO_OK && has_cells(h->st.g)) rc = h->mc.init(h->st.g);
    for (int e = 0; e < 2 && rc == MISO_OK; ++e)
        if (hipEventCreate(&h->ev[e]) != hipSuccess) rc = fail(MISO_ERR_HIP, "hipEventCreate failed");
    if (rc != MISO_OK) {
        miso_mc_destroy(h);
        return rc;
    }
    *out = h;
    return MISO_OK;
}

extern "C" int miso_mc_run(miso_mc* h)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    RowState& s = h->st;
    if (!has_cells(s.g)) {
        h->done = true;
        return h->out.reserve(0, 0);
    }
    HIPCALL(hipEventRecord(h->ev[0], s.stream));
    if (int rc = s.launch_classify(h->iso)) return rc;
    if (int rc = h->mc.launch_fe_count(s)) return rc;
    if (int rc = s.launch_scan()) return rc;  // FE row offsets: the rank -> id map layout
    if (int rc = h->mc.launch_count(s)) return rc;
    HIPCALL(hipStreamSynchronize(s.stream));
    if (int rc = h->out.reserve(s.h_totals[0], s.h_totals[1])) return rc;
    if (h->out.nverts || h->out.ntris)
        if (int rc = h->mc.launch_emit(s, h->iso, h->out.d_pts, h->out.d_nrm, h->out.d_tris, h->out.index_bits,
                                       h->out.nverts))
            return rc;
    HIPCALL(hipEventRecord(h->ev[1], s.stream));
    HIPCALL(hipStreamSynchronize(s.stream));
    HIPCALL(hipEventElapsedTime(&h->ms, h->ev[0], h->ev[1]));
    h->done = true;
    return MISO_OK;
}

extern "C" int miso_mc_set_order(miso_mc* h, int order)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (order != MISO_MC_ORDER_FIRST_VISIT && order != MISO_MC_ORDER_EDGE)
        return fail(MISO_ERR_ARG, "order must be MISO_MC_ORDER_FIRST_VISIT or MISO_MC_ORDER_EDGE");
    if (order == MISO_MC_ORDER_EDGE && has_cells(h->st.g))
        if (int rc = h->mc.init_edge_order(h->st.g)) return rc;
    h->mc.order = order;
    h->done = false;
    return MISO_OK;
}

extern "C" int miso_mc_counts(const miso_mc* h, uint64_t* nverts, uint64_t* ntris)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (!h->done) return fail(MISO_ERR_STATE, "run first");
    if (nverts) *nverts = h->out.nverts;
    if (ntris) *ntris = h->out.ntris;
    return MISO_OK;
}

extern "C" int miso_mc_copy_out(miso_mc* h, float* points, float* normals, uint64_t* triangles)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (!h->done) return fail(MISO_ERR_STATE, "run first");
    return h->out.copy_out(h->st.stream, points, normals, triangles);
}

extern "C" int miso_mc_device_output(miso_mc* h, const float** points, const float** normals,
                                     const void** triangles, int* index_bits)
{
    if (!h) return fail(MISO_ERR_ARG, "handle is NULL");
    if (!h->done) return fail(MISO_ERR_STATE, "run first");
    if (points) *points = h->out.d_pts;
    if (normals) *normals = h->out.d_nrm;
    if (triangles) *triangles = h->out.d_tris;
    if (index_bits) *index_bits = h->out.index_bits;
    return MISO_OK;
}

extern "C" int miso_mc_time_ms(const miso_mc* h, float* ms)
{
    if (!h || !ms) return fail(MISO_ERR_ARG, "NULL argument");
    *ms = h->ms;
    return MISO_OK;
}

extern "C" void miso_mc_destroy(miso_mc* h)
{
    if (!h) return;
    h->st.release();
    h->mc.release();
    h->out.release();
    for (int e = 0; e < 2; ++e)
        if (h->ev[e]) (void)hipEventDestroy(h->ev[e]);
    delete h;
}

/* ====================================================================== Flying Edges without allocation */

/* miso_fe_run_noalloc keeps all of its per-extraction device state -- the sign mask, the row counts
 * and trims, the row offsets, the scan state and the two totals -- in this file-scope array instead
 * of memory allocated per call. Each GPU holds its own copy from the library's load to the end of
 * the process. In host code `scratch` names no device address; the device address is taken per
 * device with hipGetSymbolAddress (scratch_ptr) and only that is passed to kernels. */
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

/* The device address of `scratch` on the current device, looked up on the first call there. */
int scratch_ptr(void** out)
{
    int dev = 0;
    HIPCALL(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(MISO_ERR_ARG, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MISO_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MISO_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

/* Byte offsets of the row state inside the scratch array for geometry g; the total in `end`. */
struct ScratchLayout {
    size_t mask, counts, trims, voff, toff, scan, totals, end;
    explicit ScratchLayout(const Geom& g)
    {
        const size_t R = (size_t)(g.ny * g.nz);
        const size_t ntiles = (R + kScanTile - 1) / kScanTile;
        mask = 0;
        counts = align256(mask + R * (size_t)g.W * sizeof(uint64_t));
        trims = align256(counts + R * sizeof(RowCounts));
        voff = align256(trims + R * sizeof(RowTrim));
        toff = align256(voff + R * sizeof(uint64_t));
        scan = align256(toff + R * sizeof(uint64_t));
        totals = align256(scan + (1 + 2 * ntiles) * sizeof(uint64_t));
        end = totals + 2 * sizeof(uint64_t);
    }
};

}  // namespace

extern "C" int miso_fe_run_noalloc(const miso_volume* vol, float isoval, void* stream, float* points,
                                   float* normals, uint32_t* triangles, uint64_t cap_vertices,
                                   uint64_t cap_triangles, uint64_t* nverts, uint64_t* ntris)
{
    if (int rc = check_volume(vol)) return rc;
    if (!vol->on_device) return fail(MISO_ERR_ARG, "miso_fe_run_noalloc takes a device-resident volume");
    if (nverts) *nverts = 0;
    if (ntris) *ntris = 0;
    const Geom g = make_geom(vol);
    if (!has_cells(g)) return MISO_OK;
    RowState s;               // borrowed pointers only; nothing allocated, nothing released
    s.g = g;
    s.d_vol = vol->data;
    if (!s.lane_rows())
        return fail(MISO_ERR_ARG, "miso_fe_run_noalloc takes rows of at most 4096 voxels (nx = %lld)",
                    (long long)g.nx);
    const ScratchLayout lay(g);
    if (lay.end > kScratchBytes)
        return fail(MISO_ERR_ARG,
                    "volume %lld x %lld x %lld needs %zu bytes of row state, above the %zu-byte scratch of "
                    "miso_fe_run_noalloc",
                    (long long)g.nx, (long long)g.ny, (long long)g.nz, lay.end, kScratchBytes);
    void* base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    uint8_t* sp = (uint8_t*)base;
    s.stream = (hipStream_t)stream;
    s.d_mask = (uint64_t*)(sp + lay.mask);
    s.d_counts = (RowCounts*)(sp + lay.counts);
    s.d_trims = (RowTrim*)(sp + lay.trims);
    s.d_voff = (uint64_t*)(sp + lay.voff);
    s.d_toff = (uint64_t*)(sp + lay.toff);
    s.d_scan = (uint64_t*)(sp + lay.scan);
    s.d_totals = (uint64_t*)(sp + lay.totals);
    s.ntiles = (g.ny * g.nz + kScanTile - 1) / kScanTile;
    // passes 1-3 (pass 2 clears the scan state that pass 3 starts from); pass 4 checks the totals
    // against the caller's capacities on the device and writes nothing unless they fit
    if (int rc = s.launch_classify(isoval)) return rc;
    if (int rc = s.launch_count()) return rc;
    hipLaunchKernelGGL(k_fe_scan, dim3((unsigned)s.ntiles), dim3(256), 0, s.stream, (const RowCounts*)s.d_counts,
                       g.ny * g.nz, s.d_voff, s.d_toff, s.d_scan, s.ntiles, s.d_totals);
    HIPCALL(hipGetLastError());
    {
        const uint64_t* need = s.d_totals;
        if (s.unit_geometry())
            hipLaunchKernelGGL((k_fe_emit_rows<uint32_t, true>), dim3(s.lane_row_blocks()), dim3(256), 0, s.stream,
                               s.d_vol, (const uint64_t*)s.d_mask, g, isoval, (const RowCounts*)s.d_counts,
                               (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, points, normals, triangles,
                               (uint64_t)0, need, cap_vertices, cap_triangles);
        else
            hipLaunchKernelGGL((k_fe_emit_rows<uint32_t, false>), dim3(s.lane_row_blocks()), dim3(256), 0, s.stream,
                               s.d_vol, (const uint64_t*)s.d_mask, g, isoval, (const RowCounts*)s.d_counts,
                               (const uint64_t*)s.d_voff, (const uint64_t*)s.d_toff, points, normals, triangles,
                               (uint64_t)0, need, cap_vertices, cap_triangles);
        HIPCALL(hipGetLastError());
    }
    uint64_t tot[2] = {0, 0};
    HIPCALL(hipMemcpyAsync(tot, s.d_totals, sizeof(tot), hipMemcpyDeviceToHost, s.stream));
    HIPCALL(hipStreamSynchronize(s.stream));
    if (nverts) *nverts = tot[0];
    if (ntris) *ntris = tot[1];
    if (tot[0] > cap_vertices || tot[1] > cap_triangles)
        return fail(MISO_ERR_ARG, "the mesh has %llu vertices and %llu triangles, above the buffers' %llu and %llu "
                    "(nothing written; the counts are returned)", (unsigned long long)tot[0],
                    (unsigned long long)tot[1], (unsigned long long)cap_vertices, (unsigned long long)cap_triangles);
    return MISO_OK;
}

/* ====================================================================== library */

#include "batch_host.h"

extern "C" const char* miso_last_error(void) { return g_err.c_str(); }

extern "C" const char* miso_version(void) { return "miniiso 0.1 gfx950"; }
