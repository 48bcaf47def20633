/*
 * mc_oracle.c -- CPU restatement of the reference serial Marching Cubes (TEST INFRASTRUCTURE ONLY).
 *
 * Follows marchingCubes/serial/main.cpp:42-192: cells are visited in z, y, x order over the whole
 * image; each cut edge gets the next point id the first time any cell's triangle list names it,
 * through a map keyed by the global edge id of marchingCubes/util/Image3D.cpp:168-256. Points and
 * normals use findCaseId / interpolate (marchingCubes/util/util.h:24-54) and getPosCube /
 * computeGradient (marchingCubes/util/Image3D.cpp:22-166), evaluated in float without FMA.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

struct ora_mc {
    float* pts;
    float* nrm;
    uint64_t* tris;
    uint64_t* gid;      /* global edge id of each point (Image3D.cpp:168-256) */
    size_t npts, ntris, cap_pts, cap_tris, cap_gid;
    double seconds;
};

/* open-addressing map global-edge-id -> point id (stands in for std::unordered_map) */
typedef struct {
    uint64_t* keys;
    uint64_t* vals;
    size_t mask, count;
} edge_map;

static void map_init(edge_map* m, size_t cap_pow2)
{
    m->mask = cap_pow2 - 1;
    m->count = 0;
    m->keys = malloc(cap_pow2 * sizeof(uint64_t));
    m->vals = malloc(cap_pow2 * sizeof(uint64_t));
    memset(m->keys, 0xff, cap_pow2 * sizeof(uint64_t));
}

static inline size_t map_hash(uint64_t k)
{
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33;
    return (size_t)k;
}

static void map_grow(edge_map* m);

/* returns 1 and sets *val if found; otherwise inserts (key, newval) and returns 0 */
static int map_find_or_insert(edge_map* m, uint64_t key, uint64_t newval, uint64_t* val)
{
    if ((m->count + 1) * 2 > m->mask + 1) map_grow(m);
    size_t p = map_hash(key) & m->mask;
    for (;;) {
        if (m->keys[p] == key) { *val = m->vals[p]; return 1; }
        if (m->keys[p] == UINT64_MAX) {
            m->keys[p] = key; m->vals[p] = newval; m->count++;
            *val = newval;
            return 0;
        }
        p = (p + 1) & m->mask;
    }
}

static void map_grow(edge_map* m)
{
    edge_map n;
    map_init(&n, (m->mask + 1) * 2);
    for (size_t p = 0; p <= m->mask; ++p)
        if (m->keys[p] != UINT64_MAX) {
            uint64_t dummy;
            map_find_or_insert(&n, m->keys[p], m->vals[p], &dummy);
        }
    free(m->keys); free(m->vals);
    *m = n;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void push3(float** arr, size_t n, size_t* cap, const float v[3])
{
    if (n + 1 > *cap) {
        *cap = *cap ? *cap * 2 : 1024;
        *arr = realloc(*arr, *cap * 3 * sizeof(float));
    }
    memcpy(*arr + 3 * n, v, 3 * sizeof(float));
}

ora_mc* ora_mc_run(const float* d, int64_t nx_, int64_t ny_, int64_t nz_,
                   const float sp[3], const float org[3], float iso)
{
    ora_tables_init();
    ora_mc* h = calloc(1, sizeof(ora_mc));
    if (nx_ < 2 || ny_ < 2 || nz_ < 2) return h;
    size_t nx = (size_t)nx_, ny = (size_t)ny_, nz = (size_t)nz_;
    double t0 = now_s();
    edge_map map;
    map_init(&map, 1 << 12);
    size_t cap_nrm = 0, cap_pts = 0;
    /* global edge ids, marchingCubes/util/Image3D.cpp:218-256 */
    size_t nxe = (nx - 1) * ny * nz, nye = nx * (ny - 1) * nz;
    for (size_t z = 0; z + 1 < nz; ++z)
    for (size_t y = 0; y + 1 < ny; ++y)
    for (size_t x = 0; x + 1 < nx; ++x) {
        size_t b = x + y * nx + z * nx * ny;
        float v[8] = {d[b], d[b + 1], d[b + nx + 1], d[b + nx],
                      d[b + nx * ny], d[b + nx * ny + 1], d[b + nx * ny + nx + 1], d[b + nx * ny + nx]};
        int c = 0;
        for (int q = 0; q < 8; ++q) if (v[q] >= iso) c |= 1 << q;
        if (c == 0 || c == 255) continue;
        /* getPosCube (Image3D.cpp:22-65) */
        float px = org[0] + (float)x * sp[0], py = org[1] + (float)y * sp[1], pz = org[2] + (float)z * sp[2];
        float px1 = px + sp[0], py1 = py + sp[1], pz1 = pz + sp[2];
        float pc[8][3] = {{px, py, pz}, {px1, py, pz}, {px1, py1, pz}, {px, py1, pz},
                          {px, py, pz1}, {px1, py, pz1}, {px1, py1, pz1}, {px, py1, pz1}};
        /* getGradCube / computeGradient (Image3D.cpp:67-166) */
        float gc[8][3];
        size_t corner[8][3] = {{x, y, z}, {x + 1, y, z}, {x + 1, y + 1, z}, {x, y + 1, z},
                               {x, y, z + 1}, {x + 1, y, z + 1}, {x + 1, y + 1, z + 1}, {x, y + 1, z + 1}};
        size_t dim[3] = {nx, ny, nz}, stride[3] = {1, nx, nx * ny};
        for (int q = 0; q < 8; ++q) {
            size_t idx = corner[q][0] + corner[q][1] * nx + corner[q][2] * nx * ny;
            for (int a = 0; a < 3; ++a) {
                float hi, lo, run;
                if (corner[q][a] == 0) { hi = d[idx + stride[a]]; lo = d[idx]; run = sp[a]; }
                else if (corner[q][a] == dim[a] - 1) { hi = d[idx]; lo = d[idx - stride[a]]; run = sp[a]; }
                else { hi = d[idx + stride[a]]; lo = d[idx - stride[a]]; run = 2 * sp[a]; }
                gc[q][a] = (lo - hi) / run;
            }
        }
        const signed char* te = ora_case_tris[c];
        for (; *te != -1; te += 3) {
            uint64_t tri[3];
            for (int s = 0; s < 3; ++s) {
                int e = te[s];
                size_t ge;
                switch (e) {
                case 0:  ge = x + (nx - 1) * y + ny * (nx - 1) * z; break;
                case 2:  ge = x + (nx - 1) * (y + 1) + ny * (nx - 1) * z; break;
                case 4:  ge = x + (nx - 1) * y + ny * (nx - 1) * (z + 1); break;
                case 6:  ge = x + (nx - 1) * (y + 1) + ny * (nx - 1) * (z + 1); break;
                case 3:  ge = nxe + x + nx * y + nx * (ny - 1) * z; break;
                case 1:  ge = nxe + (x + 1) + nx * y + nx * (ny - 1) * z; break;
                case 7:  ge = nxe + x + nx * y + nx * (ny - 1) * (z + 1); break;
                case 5:  ge = nxe + (x + 1) + nx * y + nx * (ny - 1) * (z + 1); break;
                case 8:  ge = nxe + nye + x + nx * y + nx * ny * z; break;
                case 9:  ge = nxe + nye + (x + 1) + nx * y + nx * ny * z; break;
                case 10: ge = nxe + nye + x + nx * (y + 1) + nx * ny * z; break;
                default: ge = nxe + nye + (x + 1) + nx * (y + 1) + nx * ny * z; break;
                }
                uint64_t id;
                if (!map_find_or_insert(&map, ge, h->npts, &id)) {
                    int v1 = ora_edge_vertices[e][0], v2 = ora_edge_vertices[e][1];
                    float w = (iso - v[v1]) / (v[v2] - v[v1]);
                    float p[3], n[3];
                    for (int a = 0; a < 3; ++a) {
                        p[a] = pc[v1][a] + (w * (pc[v2][a] - pc[v1][a]));
                        n[a] = gc[v1][a] + (w * (gc[v2][a] - gc[v1][a]));
                    }
                    push3(&h->pts, h->npts, &cap_pts, p);
                    push3(&h->nrm, h->npts, &cap_nrm, n);
                    if (h->npts + 1 > h->cap_gid) {
                        h->cap_gid = h->cap_gid ? h->cap_gid * 2 : 1024;
                        h->gid = realloc(h->gid, h->cap_gid * sizeof(uint64_t));
                    }
                    h->gid[h->npts] = ge;
                    h->npts++;
                }
                tri[s] = id;
            }
            /* the degenerate-triangle filter of main.cpp:181-184 */
            if (tri[0] != tri[1] && tri[1] != tri[2] && tri[2] != tri[0]) {
                if (h->ntris + 1 > h->cap_tris) {
                    h->cap_tris = h->cap_tris ? h->cap_tris * 2 : 1024;
                    h->tris = realloc(h->tris, h->cap_tris * 3 * sizeof(uint64_t));
                }
                memcpy(h->tris + 3 * h->ntris, tri, sizeof(tri));
                h->ntris++;
            }
        }
    }
    free(map.keys); free(map.vals);
    h->seconds = now_s() - t0;
    return h;
}

void ora_mc_sizes(const ora_mc* h, uint64_t* npts, uint64_t* ntris, double* seconds)
{
    if (npts) *npts = h->npts;
    if (ntris) *ntris = h->ntris;
    if (seconds) *seconds = h->seconds;
}

void ora_mc_mesh(const ora_mc* h, float* pts, float* nrm, uint64_t* tris)
{
    if (pts && h->npts) memcpy(pts, h->pts, h->npts * 3 * sizeof(float));
    if (nrm && h->npts) memcpy(nrm, h->nrm, h->npts * 3 * sizeof(float));
    if (tris && h->ntris) memcpy(tris, h->tris, h->ntris * 3 * sizeof(uint64_t));
}

void ora_mc_edge_ids(const ora_mc* h, uint64_t* gid)
{
    if (gid && h->npts) memcpy(gid, h->gid, h->npts * sizeof(uint64_t));
}

void ora_mc_free(ora_mc* h)
{
    if (!h) return;
    free(h->pts); free(h->nrm); free(h->tris); free(h->gid);
    free(h);
}
