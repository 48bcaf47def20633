/*
 * fe_oracle.c -- CPU restatement of the reference Flying Edges (TEST INFRASTRUCTURE ONLY).
 *
 * Follows flyingEdges/serial/FlyingEdgesAlgorithm.cpp pass by pass, and, for nthreads > 1,
 * the row-parallel loops and blocked two-level scans of flyingEdges/openmp/
 * FlyingEdgesAlgorithm.cpp. Geometry (positions, gradients, interpolation) follows
 * flyingEdges/util/Image3D.cpp:26-188 and FlyingEdgesAlgorithm.cpp:623-647 operation by
 * operation, in float, compiled with -ffp-contract=off (no fused multiply-add), which is how the
 * reference's x86-64 build evaluates them.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

typedef struct {            /* reference gridEdge, serial/FlyingEdgesAlgorithm.h:52-72 */
    size_t xl, xr, xstart, ystart, zstart;
} grid_edge;

struct ora_fe {
    const float* data;
    size_t nx, ny, nz;
    float sp[3], org[3];
    float iso;
    int nthreads;

    grid_edge* ge;          /* ny*nz */
    size_t* tri_counter;    /* (ny-1)*(nz-1) */
    unsigned char* edge_cases;  /* (nx-1)*ny*nz */
    unsigned char* cube_cases;  /* (nx-1)*(ny-1)*(nz-1) */

    /* snapshot of pass-2 counts (before the pass-3 scan overwrites them) */
    grid_edge* ge_counts;
    size_t* tri_counts;

    float* pts;
    float* nrm;
    uint64_t* tris;
    size_t npts, ntris;
    double times[6];
};

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---------------------------------------------------------------- geometry (Image3D.cpp) */

static inline float val(const ora_fe* h, size_t i, size_t j, size_t k)
{
    return h->data[k * h->nx * h->ny + j * h->nx + i];   /* Image3D.cpp:110-114 */
}

/* getValsCube, Image3D.cpp:26-43: v0=(i,j,k) v1=(i+1,j,k) v2=(i+1,j+1,k) v3=(i,j+1,k), v4..v7 at k+1 */
static void vals_cube(const ora_fe* h, size_t i, size_t j, size_t k, float v[8])
{
    v[0] = val(h, i, j, k);         v[1] = val(h, i + 1, j, k);
    v[2] = val(h, i + 1, j + 1, k); v[3] = val(h, i, j + 1, k);
    v[4] = val(h, i, j, k + 1);     v[5] = val(h, i + 1, j, k + 1);
    v[6] = val(h, i + 1, j + 1, k + 1); v[7] = val(h, i, j + 1, k + 1);
}

/* getPosCube, Image3D.cpp:45-87: corner 0 is origin + index*spacing, the others add spacing. */
static void pos_cube(const ora_fe* h, size_t i, size_t j, size_t k, float p[8][3])
{
    float x = h->org[0] + (float)i * h->sp[0];
    float y = h->org[1] + (float)j * h->sp[1];
    float z = h->org[2] + (float)k * h->sp[2];
    float x1 = x + h->sp[0], y1 = y + h->sp[1], z1 = z + h->sp[2];
    float c[8][3] = {{x, y, z}, {x1, y, z}, {x1, y1, z}, {x, y1, z},
                     {x, y, z1}, {x1, y, z1}, {x1, y1, z1}, {x, y1, z1}};
    memcpy(p, c, sizeof(c));
}

/* computeGradient, Image3D.cpp:116-188: one-sided at the faces, central inside; the result is
 * (lower - upper) / run, i.e. the negative gradient, not normalised. */
static void gradient(const ora_fe* h, size_t i, size_t j, size_t k, float g[3])
{
    size_t idx = i + j * h->nx + k * h->nx * h->ny;
    const float* d = h->data;
    size_t n[3] = {h->nx, h->ny, h->nz};
    size_t ix[3] = {i, j, k};
    size_t stride[3] = {1, h->nx, h->nx * h->ny};
    for (int a = 0; a < 3; ++a) {
        float hi, lo, run;
        if (ix[a] == 0) {
            hi = d[idx + stride[a]]; lo = d[idx]; run = h->sp[a];
        } else if (ix[a] == n[a] - 1) {
            hi = d[idx]; lo = d[idx - stride[a]]; run = h->sp[a];
        } else {
            hi = d[idx + stride[a]]; lo = d[idx - stride[a]]; run = 2 * h->sp[a];
        }
        g[a] = (lo - hi) / run;
    }
}

static void grad_cube(const ora_fe* h, size_t i, size_t j, size_t k, float g[8][3])
{
    gradient(h, i, j, k, g[0]);         gradient(h, i + 1, j, k, g[1]);
    gradient(h, i + 1, j + 1, k, g[2]); gradient(h, i, j + 1, k, g[3]);
    gradient(h, i, j, k + 1, g[4]);     gradient(h, i + 1, j, k + 1, g[5]);
    gradient(h, i + 1, j + 1, k + 1, g[6]); gradient(h, i, j + 1, k + 1, g[7]);
}

/* interpolateOnCube / interpolate, FlyingEdgesAlgorithm.cpp:623-647 */
static void interp_on_cube(float iso, const float c[8][3], const float v[8], int edge, float out[3])
{
    int i0 = ora_edge_vertices[edge][0], i1 = ora_edge_vertices[edge][1];
    float w = (iso - v[i0]) / (v[i1] - v[i0]);
    for (int a = 0; a < 3; ++a) out[a] = c[i0][a] + (w * (c[i1][a] - c[i0][a]));
}

/* ---------------------------------------------------------------- helpers (cpp:515-621) */

/* calcCaseEdge, cpp:556-574: bit 0 = left vertex below iso, bit 1 = right vertex below iso */
static inline unsigned char case_edge(int prev_ge, int cur_ge)
{
    if (prev_ge && cur_ge) return 0;
    if (!prev_ge && cur_ge) return 1;
    if (prev_ge && !cur_ge) return 2;
    return 3;
}

/* isCutEdge, cpp:515-554 */
static int is_cut_edge(const ora_fe* h, size_t i, size_t j, size_t k)
{
    size_t nx = h->nx, ny = h->ny, nz = h->nz;
    const unsigned char* ec = h->edge_cases;
    size_t idx = k * (nx - 1) * ny + j * (nx - 1) + i;
    if (ec[idx] == 1 || ec[idx] == 2) return 1;
    if (j != ny - 1) {
        size_t idy = k * (nx - 1) * ny + (j + 1) * (nx - 1) + i;
        if ((ec[idx] + ec[idy]) % 2 == 1) return 1;
    }
    if (k != nz - 1) {
        size_t idz = (k + 1) * (nx - 1) * ny + j * (nx - 1) + i;
        if ((ec[idx] + ec[idz]) % 2 == 1) return 1;
    }
    return 0;
}

/* calcCubeCase, cpp:576-604 */
static inline unsigned char cube_case(unsigned char e0, unsigned char e1, unsigned char e2,
                                      unsigned char e3)
{
    unsigned char c = 0;
    if (e0 == 0 || e0 == 2) c |= 1;
    if (e0 == 0 || e0 == 1) c |= 2;
    if (e1 == 0 || e1 == 1) c |= 4;
    if (e1 == 0 || e1 == 2) c |= 8;
    if (e2 == 0 || e2 == 2) c |= 16;
    if (e2 == 0 || e2 == 1) c |= 32;
    if (e3 == 0 || e3 == 1) c |= 64;
    if (e3 == 0 || e3 == 2) c |= 128;
    return c;
}

/* calcTrimValues, cpp:606-621 */
static void trim_values(const ora_fe* h, size_t j, size_t k, size_t* xl, size_t* xr)
{
    size_t ny = h->ny;
    const grid_edge* g0 = &h->ge[k * ny + j];
    const grid_edge* g1 = &h->ge[k * ny + j + 1];
    const grid_edge* g2 = &h->ge[(k + 1) * ny + j];
    const grid_edge* g3 = &h->ge[(k + 1) * ny + j + 1];
    size_t l = g0->xl, r = g0->xr;
    if (g1->xl < l) l = g1->xl;
    if (g2->xl < l) l = g2->xl;
    if (g3->xl < l) l = g3->xl;
    if (g1->xr > r) r = g1->xr;
    if (g2->xr > r) r = g2->xr;
    if (g3->xr > r) r = g3->xr;
    if (l > r) l = r;
    *xl = l;
    *xr = r;
}

/* ---------------------------------------------------------------- passes */

static void pass1_row(ora_fe* h, size_t j, size_t k)
{
    size_t nx = h->nx;
    unsigned char* ec = h->edge_cases + (nx - 1) * (k * h->ny + j);
    const float* row = h->data + nx * (k * h->ny + j);
    int prev = row[0] >= h->iso;
    for (size_t i = 1; i < nx; ++i) {
        int cur = row[i] >= h->iso;
        ec[i - 1] = case_edge(prev, cur);
        prev = cur;
    }
}

static void pass1_trim(ora_fe* h, size_t j, size_t k)
{
    grid_edge* g = &h->ge[k * h->ny + j];
    g->xl = h->nx;
    for (size_t i = 1; i < h->nx; ++i) {
        if (is_cut_edge(h, i - 1, j, k)) {
            if (g->xl == h->nx) g->xl = i - 1;
            g->xr = i;
        }
    }
}

/* pass 1, cpp:23-66 (OpenMP: openmp/...cpp:24-74) */
static void pass1(ora_fe* h)
{
    size_t total = h->ny * h->nz;
    #pragma omp parallel for num_threads(h->nthreads) if(h->nthreads > 1) schedule(static)
    for (size_t o = 0; o < total; ++o) pass1_row(h, o % h->ny, o / h->ny);
    #pragma omp parallel for num_threads(h->nthreads) if(h->nthreads > 1) schedule(static)
    for (size_t o = 0; o < total; ++o) pass1_trim(h, o % h->ny, o / h->ny);
}

/* pass 2, cpp:72-190 */
static void pass2_row(ora_fe* h, size_t j, size_t k)
{
    size_t nx = h->nx, ny = h->ny, nz = h->nz;
    size_t xl, xr;
    trim_values(h, j, k, &xl, &xr);
    grid_edge* g0 = &h->ge[k * ny + j];
    grid_edge* g1 = &h->ge[k * ny + j + 1];
    grid_edge* g2 = &h->ge[(k + 1) * ny + j];
    grid_edge* g3 = &h->ge[(k + 1) * ny + j + 1];
    const unsigned char* ec0 = h->edge_cases + (nx - 1) * (k * ny + j);
    const unsigned char* ec1 = h->edge_cases + (nx - 1) * (k * ny + j + 1);
    const unsigned char* ec2 = h->edge_cases + (nx - 1) * ((k + 1) * ny + j);
    const unsigned char* ec3 = h->edge_cases + (nx - 1) * ((k + 1) * ny + j + 1);
    size_t* tc = &h->tri_counter[k * (ny - 1) + j];
    unsigned char* cc = h->cube_cases + (nx - 1) * (k * (ny - 1) + j);
    int y_end = (j == ny - 2), z_end = (k == nz - 2);
    for (size_t i = xl; i != xr; ++i) {
        int x_end = (i == nx - 2);
        unsigned char c = cube_case(ec0[i], ec1[i], ec2[i], ec3[i]);
        cc[i] = c;
        if (c == 0 || c == 255) continue;
        *tc += ora_num_tris[c];
        const unsigned char* cut = ora_is_cut[c];
        g0->xstart += cut[0];
        g0->ystart += cut[3];
        g0->zstart += cut[8];
        if (x_end) { g0->ystart += cut[1]; g0->zstart += cut[9]; }
        if (y_end) { g1->xstart += cut[2]; g1->zstart += cut[10]; }
        if (z_end) { g2->xstart += cut[4]; g2->ystart += cut[7]; }
        if (x_end && y_end) g1->zstart += cut[11];
        if (x_end && z_end) g2->ystart += cut[5];
        if (y_end && z_end) g3->xstart += cut[6];
    }
}

static void pass2(ora_fe* h)
{
    size_t total = (h->ny - 1) * (h->nz - 1);
    #pragma omp parallel for num_threads(h->nthreads) if(h->nthreads > 1) schedule(static)
    for (size_t o = 0; o < total; ++o) pass2_row(h, o % (h->ny - 1), o / (h->ny - 1));
}

/* Exclusive scan of n entries of stride `stride` size_t-fields, `fields` consecutive fields per
 * entry, in place. Serial form: cpp:196-230. With threads: openmp/...cpp:209-332, each thread
 * scans a contiguous block, block totals are scanned by one thread, then added back. */
static size_t scan_fields(ora_fe* h, size_t* base, size_t n, size_t stride, int fields)
{
    if (h->nthreads <= 1) {
        size_t acc = 0;
        for (size_t r = 0; r < n; ++r)
            for (int f = 0; f < fields; ++f) {
                size_t t = base[r * stride + f];
                base[r * stride + f] = acc;
                acc += t;
            }
        return acc;
    }
    int nt = h->nthreads;
    size_t* parts = calloc((size_t)nt, sizeof(size_t));
    #pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        int cnt = omp_get_num_threads();
        size_t part = (n + (size_t)cnt - 1) / (size_t)cnt;
        size_t beg = part * (size_t)tid, end = beg + part;
        if (beg > n) beg = n;
        if (end > n) end = n;
        size_t acc = 0;
        for (size_t r = beg; r < end; ++r)
            for (int f = 0; f < fields; ++f) {
                size_t t = base[r * stride + f];
                base[r * stride + f] = acc;
                acc += t;
            }
        parts[tid] = acc;
        #pragma omp barrier
        #pragma omp single
        {
            for (int t = 1; t < cnt; ++t) parts[t] += parts[t - 1];
        }
        if (tid > 0)
            for (size_t r = beg; r < end; ++r)
                for (int f = 0; f < fields; ++f) base[r * stride + f] += parts[tid - 1];
        #pragma omp barrier
        #pragma omp single
        {
            parts[0] = parts[cnt - 1];
        }
    }
    size_t total = parts[0];
    free(parts);
    return total;
}

/* pass 3, cpp:196-269: triangle offsets per cell row, interleaved x|y|z point offsets per grid row,
 * then the zero-filled output allocation. */
static void pass3(ora_fe* h)
{
    h->ntris = scan_fields(h, h->tri_counter, (h->ny - 1) * (h->nz - 1), 1, 1);
    h->npts = scan_fields(h, &h->ge[0].xstart, h->ny * h->nz, sizeof(grid_edge) / sizeof(size_t), 3);
    h->pts = calloc(h->npts * 3 + 1, sizeof(float));
    h->nrm = calloc(h->npts * 3 + 1, sizeof(float));
    h->tris = calloc(h->ntris * 3 + 1, sizeof(uint64_t));
}

static inline void put3(float* dst, size_t idx, const float v[3])
{
    dst[3 * idx + 0] = v[0];
    dst[3 * idx + 1] = v[1];
    dst[3 * idx + 2] = v[2];
}

/* pass 4, cpp:275-497 */
static void pass4_row(ora_fe* h, size_t j, size_t k)
{
    size_t nx = h->nx, ny = h->ny, nz = h->nz;
    size_t xl, xr;
    trim_values(h, j, k, &xl, &xr);
    if (xl == xr) return;
    size_t tri = h->tri_counter[k * (ny - 1) + j];
    const unsigned char* cc = h->cube_cases + (nx - 1) * (k * (ny - 1) + j);
    const grid_edge* g0 = &h->ge[k * ny + j];
    const grid_edge* g1 = &h->ge[k * ny + j + 1];
    const grid_edge* g2 = &h->ge[(k + 1) * ny + j];
    const grid_edge* g3 = &h->ge[(k + 1) * ny + j + 1];
    size_t x0 = 0, y0 = 0, z0 = 0, x1 = 0, z1 = 0, x2 = 0, y2 = 0, x3 = 0;
    int y_end = (j == ny - 2), z_end = (k == nz - 2);
    for (size_t i = xl; i != xr; ++i) {
        int x_end = (i == nx - 2);
        unsigned char c = cc[i];
        if (c == 0 || c == 255) continue;
        const unsigned char* cut = ora_is_cut[c];
        float pc[8][3], vc[8], gc[8][3], p[3], n[3];
        pos_cube(h, i, j, k, pc);
        vals_cube(h, i, j, k, vc);
        grad_cube(h, i, j, k, gc);
        size_t gid[12] = {0};
#define EMIT(edge, idx_expr, cond)                                  \
        do {                                                        \
            size_t id_ = (idx_expr);                                \
            if (cond) {                                             \
                interp_on_cube(h->iso, pc, vc, edge, p);            \
                interp_on_cube(h->iso, gc, vc, edge, n);            \
                put3(h->pts, id_, p);                               \
                put3(h->nrm, id_, n);                               \
            }                                                       \
            gid[edge] = id_;                                        \
        } while (0)
        if (cut[0]) { EMIT(0, g0->xstart + x0, 1); ++x0; }
        if (cut[3]) { EMIT(3, g0->ystart + y0, 1); ++y0; }
        if (cut[8]) { EMIT(8, g0->zstart + z0, 1); ++z0; }
        if (cut[1]) EMIT(1, g0->ystart + y0, x_end);
        if (cut[9]) EMIT(9, g0->zstart + z0, x_end);
        if (cut[2]) { EMIT(2, g1->xstart + x1, y_end); ++x1; }
        if (cut[10]) { EMIT(10, g1->zstart + z1, y_end); ++z1; }
        if (cut[4]) { EMIT(4, g2->xstart + x2, z_end); ++x2; }
        if (cut[7]) { EMIT(7, g2->ystart + y2, z_end); ++y2; }
        if (cut[11]) EMIT(11, g1->zstart + z1, x_end && y_end);
        if (cut[5]) EMIT(5, g2->ystart + y2, x_end && z_end);
        if (cut[6]) { EMIT(6, g3->xstart + x3, y_end && z_end); ++x3; }
#undef EMIT
        const signed char* ct = ora_case_tris[c];
        for (int t = 0; ct[t] != -1; t += 3) {
            h->tris[3 * tri + 0] = gid[(int)ct[t]];
            h->tris[3 * tri + 1] = gid[(int)ct[t + 1]];
            h->tris[3 * tri + 2] = gid[(int)ct[t + 2]];
            ++tri;
        }
    }
}

static void pass4(ora_fe* h)
{
    size_t total = (h->ny - 1) * (h->nz - 1);
    #pragma omp parallel for num_threads(h->nthreads) if(h->nthreads > 1) schedule(static)
    for (size_t o = 0; o < total; ++o) pass4_row(h, o % (h->ny - 1), o / (h->ny - 1));
}

/* ---------------------------------------------------------------- API */

ora_fe* ora_fe_run(const float* data, int64_t nx, int64_t ny, int64_t nz,
                   const float spacing[3], const float origin[3], float iso, int nthreads)
{
    ora_tables_init();
    ora_fe* h = calloc(1, sizeof(ora_fe));
    h->data = data;
    h->nx = (size_t)nx; h->ny = (size_t)ny; h->nz = (size_t)nz;
    memcpy(h->sp, spacing, sizeof(h->sp));
    memcpy(h->org, origin, sizeof(h->org));
    h->iso = iso;
    h->nthreads = nthreads < 1 ? 1 : nthreads;
    if (nx < 2 || ny < 2 || nz < 2) {
        /* no cells: the reference's loops produce an empty mesh */
        h->pts = calloc(1, sizeof(float));
        h->nrm = calloc(1, sizeof(float));
        h->tris = calloc(1, sizeof(uint64_t));
        return h;
    }
    double t0 = now_s();
    /* the constructor allocates and zero-fills per-row and per-edge state (FlyingEdgesAlgorithm.h:29-39) */
    h->ge = calloc(h->ny * h->nz, sizeof(grid_edge));
    h->tri_counter = calloc((h->ny - 1) * (h->nz - 1), sizeof(size_t));
    h->edge_cases = calloc((h->nx - 1) * h->ny * h->nz, 1);
    h->cube_cases = calloc((h->nx - 1) * (h->ny - 1) * (h->nz - 1), 1);
    double t1 = now_s();
    pass1(h);
    double t2 = now_s();
    pass2(h);
    double t3 = now_s();
    /* keep the pre-scan counts for the row-count comparisons */
    h->ge_counts = malloc(h->ny * h->nz * sizeof(grid_edge));
    memcpy(h->ge_counts, h->ge, h->ny * h->nz * sizeof(grid_edge));
    h->tri_counts = malloc((h->ny - 1) * (h->nz - 1) * sizeof(size_t));
    memcpy(h->tri_counts, h->tri_counter, (h->ny - 1) * (h->nz - 1) * sizeof(size_t));
    double t3b = now_s();
    pass3(h);
    double t4 = now_s();
    pass4(h);
    double t5 = now_s();
    h->times[0] = t1 - t0;
    h->times[1] = t2 - t1;
    h->times[2] = t3 - t2;
    h->times[3] = t4 - t3b;
    h->times[4] = t5 - t4;
    h->times[5] = (t5 - t0) - (t3b - t3);
    return h;
}

void ora_fe_sizes(const ora_fe* h, uint64_t* npts, uint64_t* ntris)
{
    if (npts) *npts = h->npts;
    if (ntris) *ntris = h->ntris;
}

void ora_fe_times(const ora_fe* h, double out[6]) { memcpy(out, h->times, sizeof(h->times)); }

void ora_fe_mesh(const ora_fe* h, float* pts, float* nrm, uint64_t* tris)
{
    if (pts) memcpy(pts, h->pts, h->npts * 3 * sizeof(float));
    if (nrm) memcpy(nrm, h->nrm, h->npts * 3 * sizeof(float));
    if (tris) memcpy(tris, h->tris, h->ntris * 3 * sizeof(uint64_t));
}

void ora_fe_rows(const ora_fe* h, uint64_t* xl, uint64_t* xr, uint64_t* xcnt, uint64_t* ycnt,
                 uint64_t* zcnt, uint64_t* tcnt)
{
    if (!h->ge_counts) return;
    size_t rows = h->ny * h->nz;
    for (size_t r = 0; r < rows; ++r) {
        if (xl) xl[r] = h->ge_counts[r].xl;
        if (xr) xr[r] = h->ge_counts[r].xr;
        if (xcnt) xcnt[r] = h->ge_counts[r].xstart;
        if (ycnt) ycnt[r] = h->ge_counts[r].ystart;
        if (zcnt) zcnt[r] = h->ge_counts[r].zstart;
    }
    if (tcnt)
        for (size_t c = 0; c < (h->ny - 1) * (h->nz - 1); ++c) tcnt[c] = h->tri_counts[c];
}

void ora_fe_offsets(const ora_fe* h, uint64_t* xs, uint64_t* ys, uint64_t* zs, uint64_t* toff)
{
    if (!h->ge) return;
    size_t rows = h->ny * h->nz;
    for (size_t r = 0; r < rows; ++r) {
        if (xs) xs[r] = h->ge[r].xstart;
        if (ys) ys[r] = h->ge[r].ystart;
        if (zs) zs[r] = h->ge[r].zstart;
    }
    if (toff)
        for (size_t c = 0; c < (h->ny - 1) * (h->nz - 1); ++c) toff[c] = h->tri_counter[c];
}

void ora_fe_cases(const ora_fe* h, unsigned char* cube_cases, unsigned char* edge_cases)
{
    if (!h->cube_cases) return;
    if (cube_cases) memcpy(cube_cases, h->cube_cases, (h->nx - 1) * (h->ny - 1) * (h->nz - 1));
    if (edge_cases) memcpy(edge_cases, h->edge_cases, (h->nx - 1) * h->ny * h->nz);
}

void ora_fe_free(ora_fe* h)
{
    if (!h) return;
    free(h->ge); free(h->tri_counter); free(h->edge_cases); free(h->cube_cases);
    free(h->ge_counts); free(h->tri_counts);
    free(h->pts); free(h->nrm); free(h->tris);
    free(h);
}
