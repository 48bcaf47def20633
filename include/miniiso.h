/*
 * miniiso.h -- C ABI of the MI355X isosurface engine (libminiiso.so).
 *
 * Plain C types only: pointers, sizes, status codes. No exception crosses this boundary; every
 * call returns MISO_OK (0) or a negative status, and miso_last_error() describes the last failure
 * on the calling thread. One handle is used from one host thread at a time; each handle owns one
 * HIP stream (or borrows the caller's); a slab handle owns two, plus one for host-to-device copies.
 *
 * Reference interfaces these entry points replace (paths relative to the miniIsosurface root):
 *   FlyingEdgesAlgorithm(Image3D const&, scalar_t const&)   flyingEdges/serial/FlyingEdgesAlgorithm.h:29-39
 *   void pass1(); pass2(); pass3(); pass4();                 flyingEdges/serial/FlyingEdgesAlgorithm.h:41-47
 *   util::TriangleMesh moveOutput();                         flyingEdges/serial/FlyingEdgesAlgorithm.h:49
 *   (the OpenMP class has the same interface: flyingEdges/openmp/FlyingEdgesAlgorithm.h:29-51)
 *   template<T> TriangleMesh<T> MarchingCubes(Image3D<T> const&, T const&)
 *                                                            marchingCubes/serial/main.cpp:42-44
 *   MarchingCubes(image, isoval, nSectionsX, nSectionsY, nSectionsZ)
 *                                                            marchingCubes/openmp/main.cpp:159-162
 *   the duplicate-free variant's MarchingCubes and util::duplicateRemover
 *                                                            marchingCubes/openmpDupFree/main.cpp:159-331
 *   loadImageSections (sections with ghost cells)            marchingCubes/mpi/mpiutil.h:87-221
 */
#ifndef MINIISO_H
#define MINIISO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MISO_OK 0
#define MISO_ERR_HIP (-1)       /* a HIP runtime call failed */
#define MISO_ERR_ALLOC (-2)     /* device or host allocation failed */
#define MISO_ERR_ARG (-3)       /* invalid argument */
#define MISO_ERR_STATE (-4)     /* call out of order (e.g. pass 3 before pass 2) */
#define MISO_ERR_BUDGET (-5)    /* request exceeds the admission budget */

/* An image volume: data[i + nx*(j + ny*k)], x fastest, as util::Image3D stores it
 * (flyingEdges/util/Image3D.h:31-37, Image3D.cpp:110-114). */
/* A device-resident volume is read by kernels on the handle's stream: the caller passes the
 * stream that produced the data (or one ordered after it). With stream == NULL the library
 * creates a blocking stream, which is ordered after work on the null stream but not after other
 * non-blocking streams. */
typedef struct miso_volume {
    const float* data;       /* host pointer (on_device == 0) or device pointer (on_device == 1) */
    int64_t nx, ny, nz;
    float spacing[3];
    float origin[3];         /* the reference's zeroPos */
    int32_t on_device;
} miso_volume;

/* ---------------------------------------------------------------- Flying Edges */
typedef struct miso_fe miso_fe;

/* FlyingEdgesAlgorithm(image, isoval) (serial/FlyingEdgesAlgorithm.h:29-39). A host volume is
 * copied to the device here; a device volume is borrowed and must outlive the handle.
 * `stream` is a hipStream_t to run on, or NULL for a stream owned by the handle. */
int miso_fe_create(const miso_volume* vol, float isoval, void* stream, miso_fe** out);

/* Re-targets an existing handle to another isovalue without reallocating (passes restart at 1). */
int miso_fe_set_isoval(miso_fe* h, float isoval);

/* pass1()..pass4() (serial/FlyingEdgesAlgorithm.h:41-47); must be called in order 1, 2, 3, 4.
 * Each call waits for its kernels to finish, so host timers around it measure the pass. */
int miso_fe_pass(miso_fe* h, int pass);

/* All four passes with a single host synchronisation (after pass 3, to size the output). */
int miso_fe_run(miso_fe* h);

/* Vertex and triangle counts; valid after pass 3 (the reference allocates its output there). */
int miso_fe_counts(const miso_fe* h, uint64_t* nverts, uint64_t* ntris);

/* moveOutput() (serial/FlyingEdgesAlgorithm.h:49): copies the mesh to host arrays of
 * nverts*3 floats (points, normals) and ntris*3 indices of 8 bytes (the reference's size_t).
 * Any pointer may be NULL to skip that array. */
int miso_fe_copy_out(miso_fe* h, float* points, float* normals, uint64_t* triangles);

/* Device-resident output (owned by the handle until the next run or destroy). Triangle indices
 * are 32-bit when nverts < 2^32 (index_bits = 32), else 64-bit. */
int miso_fe_device_output(miso_fe* h, const float** points, const float** normals,
                          const void** triangles, int* index_bits);

/* Forces 64-bit triangle indices (index_bits 64) or the default (32 when they fit). */
int miso_fe_set_index_bits(miso_fe* h, int index_bits);

/* Device time of the last run's passes in milliseconds: [pass1, pass2, pass3, pass4]; 0 for a
 * pass that miso_fe_set_timing leaves untimed. */
int miso_fe_pass_times(const miso_fe* h, float ms[4]);

/* Which passes the next runs time: bit p - 1 for pass p (default 0, none). Each timed pass
 * records a pair of HIP events on the handle's stream, about 4-10 us of device time per pass
 * between the kernels. No counterpart in the reference, whose drivers time on the host. */
int miso_fe_set_timing(miso_fe* h, unsigned passes);


/* Per grid row (ny*nz entries, r = j + ny*k): xl, xr (pass 1 trims), the cut x/y/z edge counts
 * (pass 2, before the scan) and the vertex offset (after pass 3); per cell row, stored at the
 * grid-row index of its (j,k) with j < ny-1 and k < nz-1: triangle count and offset. Arrays may
 * be NULL. Valid after pass 3. */
int miso_fe_rows(const miso_fe* h, uint32_t* xl, uint32_t* xr, uint32_t* xcount,
                 uint32_t* ycount, uint32_t* zcount, uint32_t* tcount, uint64_t* voffset,
                 uint64_t* toffset);

/* The reference's cubeCases after pass 2 ((nx-1)*(ny-1)*(nz-1) bytes, x fastest): the
 * marching-cubes case inside each cell row's trimmed range [xl, xr) and 0 elsewhere
 * (FlyingEdgesAlgorithm.cpp:106-113). With `all_cells` nonzero, the true case of every cell. */
int miso_fe_cube_cases(miso_fe* h, uint8_t* cases, int all_cells);

void miso_fe_destroy(miso_fe* h);

/* ---------------------------------------------------------------- Flying Edges over z-slabs */
/* The section decomposition with ghost cells of the MPI variant (marchingCubes/mpi/mpiutil.h:
 * 87-221: each section reads one ghost plane below and two above), along z and run slab after
 * slab on one GPU, for volumes larger than the device memory or to bound it. Slab s owns the
 * grid planes [k0, k1), at most `planes` of them: the vertices of those planes' rows and the
 * triangles of their cells, numbered as the whole-volume run numbers them. Concatenating the parts
 * in order gives miso_fe_run's mesh, bit for bit. A host volume is borrowed until the last
 * miso_fe_slabs_next and streamed to the device one slab ahead of the one being computed (from
 * pinned memory the copy overlaps the computation); a device volume is read in place. While the
 * caller takes slab s's part, slab s + 1 is already being counted. Device memory: (planes + 3)
 * planes of volume twice for a host volume, two sets of row state for (planes + 3) planes (about
 * 40 bytes per grid row each, plus the sign mask), and the largest part. */
typedef struct miso_fe_slabs miso_fe_slabs;
int miso_fe_slabs_create(const miso_volume* vol, float isoval, int64_t planes, void* stream,
                         miso_fe_slabs** out);
/* Extracts the next slab; *done = 1 (and nothing extracted) once all planes are done. */
int miso_fe_slabs_next(miso_fe_slabs* h, int* done);
/* Starts over from the first slab at a new isovalue, keeping the device buffers (the same
 * volume, still borrowed). */
int miso_fe_slabs_restart(miso_fe_slabs* h, float isoval);
/* The current slab: its planes [k0, k1), and its part of the mesh: vertices vbase .. vbase +
 * nverts - 1 and triangles tbase .. tbase + ntris - 1 of the whole mesh. */
int miso_fe_slabs_part(const miso_fe_slabs* h, int64_t* k0, int64_t* k1, uint64_t* vbase,
                       uint64_t* nverts, uint64_t* tbase, uint64_t* ntris);
/* The current part into host buffers (nverts x 3 floats twice, ntris x 3 ids that number the
 * whole mesh's vertices), or its device buffers (valid until the next call on the handle). */
int miso_fe_slabs_copy_part(miso_fe_slabs* h, float* points, float* normals, uint64_t* triangles);
int miso_fe_slabs_device_part(miso_fe_slabs* h, const float** points, const float** normals,
                              const uint64_t** triangles);
void miso_fe_slabs_destroy(miso_fe_slabs* h);

/* ---------------------------------------------------------------- Marching Cubes */
typedef struct miso_mc miso_mc;

/* MarchingCubes(image, isoval) (marchingCubes/serial/main.cpp:42-192): same triangle order as
 * Flying Edges; vertices numbered in the serial reference's first-visit order. */
int miso_mc_create(const miso_volume* vol, float isoval, void* stream, miso_mc** out);
int miso_mc_run(miso_mc* h);
int miso_mc_counts(const miso_mc* h, uint64_t* nverts, uint64_t* ntris);
int miso_mc_copy_out(miso_mc* h, float* points, float* normals, uint64_t* triangles);
int miso_mc_device_output(miso_mc* h, const float** points, const float** normals,
                          const void** triangles, int* index_bits);
int miso_mc_time_ms(const miso_mc* h, float* ms);
void miso_mc_destroy(miso_mc* h);
/* Vertex order of the next runs. FIRST_VISIT (the default): the serial reference's
 * (marchingCubes/serial/main.cpp:145-177). EDGE: sorted by global edge id, the order of the
 * duplicate-free OpenMP variant (util::duplicateRemover, marchingCubes/util/DuplicateRemover.h:
 * 34-70, called at marchingCubes/openmpDupFree/main.cpp:324-331); deterministic, with triangles
 * in cell order. It equals that variant's output for one section, and for any sectioning when
 * spacing is 1 and origin 0; otherwise a seam point there may come from a later section's cell
 * and differ in the last place. */
#define MISO_MC_ORDER_FIRST_VISIT 0
#define MISO_MC_ORDER_EDGE 1
int miso_mc_set_order(miso_mc* h, int order);

/* ---------------------------------------------------------------- batched Flying Edges */
/* No counterpart in the reference, which extracts one volume per process
 * (flyingEdges/serial/main.cpp:110-158). A batch is a fixed list of jobs -- device-resident
 * volumes of any shape (dimensions >= 2, rows of at most 4096 voxels), each with its own
 * isovalue; several jobs may share one volume -- extracted by one persistent launch per pass,
 * replayed from captured HIP graphs. Each job's result is bit-identical to miso_fe_run on that
 * volume and isovalue; its triangles index its own vertices from 0 (32-bit). */
typedef struct miso_batch miso_batch;

int miso_batch_create(const miso_volume* volumes, const float* isovals, int64_t njobs, void* stream,
                      miso_batch** out);
/* passes 1-4 for every job; synchronous */
int miso_batch_run(miso_batch* b);
int miso_batch_totals(const miso_batch* b, uint64_t* nverts, uint64_t* ntris);
int miso_batch_job_counts(const miso_batch* b, int64_t job, uint64_t* nverts, uint64_t* ntris);
/* device pointers into the batch's shared output, valid until the next run or destroy */
int miso_batch_job_output(const miso_batch* b, int64_t job, const float** points, const float** normals,
                          const uint32_t** triangles);
/* host copy of one job's mesh, triangles widened to 64-bit (util::TriangleMesh's size_t) */
int miso_batch_copy_out(miso_batch* b, int64_t job, float* points, float* normals, uint64_t* triangles);
/* launch shape of the persistent kernels: resident blocks per CU and work items per ticket */
int miso_batch_set_tuning(miso_batch* b, int blocks_per_cu, int chunk);
int miso_batch_get_tuning(const miso_batch* b, int* blocks_per_cu, int* chunk);
/* Pass 4's work queue. DISPATCH (the default): one queue, row blocks in job order. XCD: one
 * queue per XCD of the MI355X (a persistent block serves the queue of the XCD it runs on, block
 * b on XCD b mod 8); each XCD owns every eighth row block of a job, in the tile order of the
 * single-volume pass 4, so that neighbouring rows meet in one XCD's L2. Same output. */
#define MISO_BATCH_QUEUE_DISPATCH 0
#define MISO_BATCH_QUEUE_XCD 1
int miso_batch_set_queue_order(miso_batch* b, int order);
int miso_batch_get_queue_order(const miso_batch* b, int* order);
/* How miso_batch_run launches once the output capacity is known from an earlier run. GRAPHS (the
 * default): one captured graph of six kernels (clear, passes 1, 2, 3, per-job counts, 4). FUSED:
 * one cooperative kernel whose resident blocks run the same phases with a grid-wide barrier
 * between them (at most the blocks per CU that fit at once; MISO_ERR_ARG where the device has no
 * cooperative launch). Same output; a run whose output outgrows the buffers ends with pass 4
 * from the graph either way. */
#define MISO_BATCH_LAUNCH_GRAPHS 0
#define MISO_BATCH_LAUNCH_FUSED 1
int miso_batch_set_launch(miso_batch* b, int mode);
int miso_batch_get_launch(const miso_batch* b, int* mode);
/* times every setting of a small grid on this batch (reps runs each) -- resident blocks per CU,
 * items per ticket, pass 4's queue order, launch mode -- and keeps the fastest */
int miso_batch_autotune(miso_batch* b, int reps, float* best_ms);
/* device time of the last run, passes 1-4 */
int miso_batch_time_ms(const miso_batch* b, float* ms);
void miso_batch_destroy(miso_batch* b);

/* Split-phase use for a serving loop (no counterpart in the reference either): a batch without
 * jobs whose job list is replaced per step, passes 1-3 enqueued without waiting, the counts then
 * used to admit a prefix of the jobs by the memory their output needs, and pass 4 of that prefix
 * written into caller-owned device buffers (job j's output starts at its first vertex / first
 * triangle as given by miso_batch_job_offsets; jobs after the prefix get no output). */
int miso_batch_create_empty(void* stream, miso_batch** out);
int miso_batch_set_jobs(miso_batch* b, const miso_volume* volumes, const float* isovals, int64_t njobs);
int miso_batch_count_async(miso_batch* b);
/* 1 when the batch's stream is idle, 0 while work is in flight */
int miso_batch_query(miso_batch* b);
int miso_batch_count_wait(miso_batch* b);
int miso_batch_job_offsets(const miso_batch* b, int64_t job, uint64_t* first_vertex, uint64_t* first_triangle);
int miso_batch_emit_async(miso_batch* b, int64_t njobs_emit, float* points, float* normals, uint32_t* triangles,
                          uint64_t cap_vertices, uint64_t cap_triangles);
int miso_batch_wait(miso_batch* b);

/* ---------------------------------------------------------------- Flying Edges without allocation */
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): miso_fe_run in one
 * call on a device-resident volume, with no device allocation and no free. The row state of the
 * four passes (sign mask, row counts and trims, offsets, scan state, totals) lives in a static
 * device array of the library, 64 MiB per GPU, held from the library's load until the process
 * ends; volumes whose row state exceeds it (one bit per voxel plus about 40 bytes per grid row:
 * up to about 700^3 voxels) are refused with MISO_ERR_ARG, as are rows of more than 4096 voxels. The caller passes the output buffers: points and normals of cap_vertices * 3 floats,
 * triangles of cap_triangles * 3 uint32 indices (each triangle indexes this mesh's vertices).
 * Kernels run on `stream` (NULL: the null stream); the call returns when they are done, with the
 * mesh's counts in *nverts and *ntris. If the mesh does not fit the buffers, nothing is written
 * to them, the counts are still returned, and the status is MISO_ERR_ARG. Every call on one GPU
 * shares that GPU's array: calls that overlap in time on one GPU (from several threads or
 * streams) race and must be serialised by the caller. Same output as miso_fe_run. */
int miso_fe_run_noalloc(const miso_volume* vol, float isoval, void* stream, float* points, float* normals,
                        uint32_t* triangles, uint64_t cap_vertices, uint64_t cap_triangles, uint64_t* nverts,
                        uint64_t* ntris);

/* ---------------------------------------------------------------- library */
const char* miso_last_error(void);
/* "miniiso <version> gfx950" */
const char* miso_version(void);

#ifdef __cplusplus
}
#endif
#endif
