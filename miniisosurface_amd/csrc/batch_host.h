/*
 * batch_host.h -- host side of the batched Flying Edges entry points (miso_batch_* in
 * include/miniiso.h). Included once, by capi.hip, after its helpers.
 *
 * A batch is a fixed list of jobs (device volume, isovalue). Passes 1-3 and the per-job count
 * gather run as one captured hipGraph; pass 4 as a second one, or both as a single graph once
 * the output capacity is known from an earlier run. The output of all jobs shares one buffer of
 * points, one of normals and one of 32-bit triangle indices.
 */

struct miso_batch {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int njobs = 0;
    std::vector<BatchJob> jobs;
    std::vector<int64_t> cls_prefix, lane_prefix;    // work items per classify group / per job, prefix sums
    std::vector<int> group_job;                      // classify group q: jobs [group_job[q], group_job[q + 1])
    int ngroups = 0;
    int* d_group_job = nullptr;
    BatchJob* d_jobs = nullptr;
    int64_t* d_cls_prefix = nullptr;
    int64_t* d_lane_prefix = nullptr;
    int64_t R = 0, mask_words = 0, ntiles = 0;
    uint64_t* d_mask = nullptr;
    RowCounts* d_counts = nullptr;
    RowTrim* d_trims = nullptr;
    uint64_t* d_voff = nullptr;
    uint64_t* d_toff = nullptr;
    uint64_t* d_state = nullptr;      // [0..2] queue tickets, [3] totals V, [4] totals T, [5..12] the XCD
                                      // queues' tickets of pass 4, [13..] scan state
    int64_t state_words = 0;
    uint64_t* d_jobout = nullptr;     // 4 per job: first vertex, first triangle, vertices, triangles
    uint64_t* h_jobout = nullptr;     // pinned copy, then the two totals
    float* d_pts = nullptr;
    float* d_nrm = nullptr;
    uint32_t* d_tris = nullptr;
    uint64_t cap_v = 0, cap_t = 0;
    int blocks_per_cu = 4, chunk = 4, num_cus = 256;
    int queue_order = MISO_BATCH_QUEUE_DISPATCH;     // pass 4's work queue: one, or one per XCD
    int launch = MISO_BATCH_LAUNCH_GRAPHS;           // graph of six kernels, or k_batch_fused
    int fused_per_cu = 0;                            // k_batch_fused blocks resident per CU (0: none)
    std::vector<int64_t> xcd_prefix;                 // 8 x (njobs + 1): XCD x's pass 4 items before job j
    int64_t* d_xcd_prefix = nullptr;
    int64_t cap_xcd = 0;
    int64_t cap_jobs = 0, cap_prefix_c = 0, cap_prefix_l = 0, cap_mask = 0, cap_counts = 0, cap_trims = 0;
    int64_t cap_voff = 0, cap_toff = 0, cap_state = 0, cap_jobout = 0, cap_host = 0, cap_groups = 0;
    bool counted = false;             // split-phase use: counts of the current job set are on the host
    hipGraphExec_t g_scan = nullptr, g_emit = nullptr, g_all = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    bool ran = false;

    uint64_t* tickets() { return d_state; }
    uint64_t* totals() { return d_state + 3; }
    uint64_t* xcd_tickets() { return d_state + 5; }
    uint64_t* scan_state() { return d_state + 13; }
};

namespace {

bool geom_unit(const Geom& g)
{
    return g.sp[0] == 1.0f && g.sp[1] == 1.0f && g.sp[2] == 1.0f && g.org[0] == 0.0f && g.org[1] == 0.0f &&
           g.org[2] == 0.0f && g.nx < (1 << 24) && g.ny < (1 << 24) && g.nz < (1 << 24);
}

bool geom_lane_rows(const Geom& g)
{
    const int64_t planes = (63 + g.ny) / g.ny + 3;
    return g.W <= 64 && (g.ny + 130) * 3 * g.nx < (int64_t(1) << 32) && planes * g.nx * g.ny < (int64_t(1) << 31);
}

int64_t geom_rows_per_wave(const Geom& g)
{
    int64_t wp = 1;
    while (wp < g.W) wp <<= 1;
    return 64 / wp;
}

void destroy_graphs(miso_batch* b)
{
    if (b->g_scan) (void)hipGraphExecDestroy(b->g_scan);
    if (b->g_emit) (void)hipGraphExecDestroy(b->g_emit);
    if (b->g_all) (void)hipGraphExecDestroy(b->g_all);
    b->g_scan = b->g_emit = b->g_all = nullptr;
}

unsigned batch_grid(const miso_batch* b, int64_t nitems)
{
    const int64_t want = (int64_t)b->num_cus * b->blocks_per_cu;
    const int64_t need = (nitems + b->chunk - 1) / b->chunk;
    return (unsigned)std::max<int64_t>(1, std::min(want, need));
}

int enqueue_scan(miso_batch* b)
{
    const int64_t ncls = b->cls_prefix.back(), nlane = b->lane_prefix.back();
    hipLaunchKernelGGL(k_batch_zero, dim3((unsigned)std::min<int64_t>(1024, (b->state_words + 255) / 256)), dim3(256),
                       0, b->stream, b->d_state, b->state_words);
    hipLaunchKernelGGL(k_batch_classify, dim3(batch_grid(b, ncls)), dim3(256), 0, b->stream,
                       (const BatchJob*)b->d_jobs, (const int*)b->d_group_job, (const int64_t*)b->d_cls_prefix,
                       b->ngroups, ncls, b->chunk, b->d_mask, b->tickets() + 0);
    hipLaunchKernelGGL(k_batch_count, dim3(batch_grid(b, nlane)), dim3(256), 0, b->stream, (const BatchJob*)b->d_jobs,
                       (const int64_t*)b->d_lane_prefix, b->njobs, nlane, b->chunk, (const uint64_t*)b->d_mask,
                       b->d_counts, b->d_trims, b->tickets() + 1);
    hipLaunchKernelGGL(k_fe_scan, dim3((unsigned)b->ntiles), dim3(256), 0, b->stream, (const RowCounts*)b->d_counts,
                       b->R, b->d_voff, b->d_toff, b->scan_state(), b->ntiles, b->totals());
    hipLaunchKernelGGL(k_batch_job_counts, dim3((unsigned)((b->njobs + 255) / 256)), dim3(256), 0, b->stream,
                       (const BatchJob*)b->d_jobs, b->njobs, b->R, (const uint64_t*)b->d_voff,
                       (const uint64_t*)b->d_toff, (const uint64_t*)b->totals(), b->d_jobout);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(b->h_jobout, b->d_jobout, 4 * (size_t)b->njobs * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           b->stream));
    HIPCALL(hipMemcpyAsync(b->h_jobout + 4 * b->njobs, b->totals(), 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           b->stream));
    return MISO_OK;
}

/* pass 4 for jobs [0, nemit) into the given buffers */
int enqueue_emit_into(miso_batch* b, int64_t nemit, float* pts, float* nrm, uint32_t* tris, uint64_t cap_v,
                      uint64_t cap_t)
{
    const int64_t nlane = b->lane_prefix[nemit];
    if (nlane == 0) return MISO_OK;
    const uint64_t* need = nemit < b->njobs ? b->d_jobout + 4 * nemit : b->totals();
    // the emit queue's tickets start at zero on every emit, also a second emit after one count
    // (a fill kernel, not a memset node, so that the enqueue can be captured)
    const bool xcd = b->queue_order == MISO_BATCH_QUEUE_XCD;
    uint64_t* tk = xcd ? b->xcd_tickets() : b->tickets() + 2;
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, b->stream, tk, (int64_t)(xcd ? 8 : 1));
    unsigned grid = batch_grid(b, nlane);
    if (xcd) grid = std::max(8u, (grid + 7u) / 8u * 8u);     // every XCD queue has blocks
    hipLaunchKernelGGL(k_batch_emit, dim3(grid), dim3(256), 0, b->stream, (const BatchJob*)b->d_jobs,
                       (const int64_t*)b->d_lane_prefix, b->njobs, nlane, b->chunk, (const uint64_t*)b->d_mask,
                       (const RowCounts*)b->d_counts, (const uint64_t*)b->d_voff, (const uint64_t*)b->d_toff,
                       need, cap_v, cap_t, pts, nrm, tris, tk,
                       xcd ? (const int64_t*)b->d_xcd_prefix : (const int64_t*)nullptr, (int)nemit);
    HIPCALL(hipGetLastError());
    return MISO_OK;
}

int enqueue_emit(miso_batch* b)
{
    return enqueue_emit_into(b, b->njobs, b->d_pts, b->d_nrm, b->d_tris, b->cap_v, b->cap_t);
}

/* captures what `fn` enqueues on the batch stream into an executable graph */
template <class Fn>
int capture(miso_batch* b, hipGraphExec_t* out, Fn fn)
{
    hipGraph_t graph = nullptr;
    HIPCALL(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = fn();
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(MISO_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(MISO_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return MISO_OK;
}

int batch_capture_emit(miso_batch* b)
{
    if (b->g_emit) (void)hipGraphExecDestroy(b->g_emit);
    if (b->g_all) (void)hipGraphExecDestroy(b->g_all);
    b->g_emit = b->g_all = nullptr;
    if (int rc = capture(b, &b->g_emit, [&] { return enqueue_emit(b); })) return rc;
    return capture(b, &b->g_all, [&] {
        if (int rc = enqueue_scan(b)) return rc;
        return enqueue_emit(b);
    });
}

int batch_reserve(miso_batch* b, uint64_t v, uint64_t t)
{
    if (v <= b->cap_v && t <= b->cap_t && b->d_pts) return MISO_OK;
    HIPCALL(hipStreamSynchronize(b->stream));
    const uint64_t nv = std::max(v, b->cap_v), nt = std::max(t, b->cap_t);
    dev_free(b->d_pts);
    dev_free(b->d_nrm);
    dev_free(b->d_tris);
    b->cap_v = b->cap_t = 0;
    if (int rc = dev_alloc(&b->d_pts, (size_t)(3 * nv))) return rc;
    if (int rc = dev_alloc(&b->d_nrm, (size_t)(3 * nv))) return rc;
    if (int rc = dev_alloc(&b->d_tris, (size_t)(3 * nt))) return rc;
    b->cap_v = nv;
    b->cap_t = nt;
    return batch_capture_emit(b);
}

void batch_release(miso_batch* b)
{
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_graphs(b);
    dev_free(b->d_jobs);
    dev_free(b->d_cls_prefix);
    dev_free(b->d_group_job);
    dev_free(b->d_lane_prefix);
    dev_free(b->d_xcd_prefix);
    dev_free(b->d_mask);
    dev_free(b->d_counts);
    dev_free(b->d_trims);
    dev_free(b->d_voff);
    dev_free(b->d_toff);
    dev_free(b->d_state);
    dev_free(b->d_jobout);
    dev_free(b->d_pts);
    dev_free(b->d_nrm);
    dev_free(b->d_tris);
    if (b->h_jobout) (void)hipHostFree(b->h_jobout);
    b->h_jobout = nullptr;
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    b->ev0 = b->ev1 = nullptr;
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    b->stream = nullptr;
}

int batch_setup(miso_batch* b, void* user_stream)
{
    if (user_stream) {
        b->stream = (hipStream_t)user_stream;
    } else {
        HIPCALL(hipStreamCreateWithFlags(&b->stream, hipStreamDefault));   // ordered after null-stream work
        b->own_stream = true;
    }
    int dev = 0;
    HIPCALL(hipGetDevice(&dev));
    HIPCALL(hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    int coop = 0;
    HIPCALL(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (coop)
        HIPCALL(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->fused_per_cu, k_batch_fused, 256, 0));
    HIPCALL(hipEventCreate(&b->ev0));
    HIPCALL(hipEventCreate(&b->ev1));
    return MISO_OK;
}

template <typename T>
int grow(T** p, int64_t* cap, int64_t need)
{
    if (need <= *cap && *p) return MISO_OK;
    dev_free(*p);
    *cap = 0;
    const int64_t n = std::max<int64_t>(need, 1);
    if (int rc = dev_alloc(p, (size_t)n)) return rc;
    *cap = n;
    return MISO_OK;
}

/* Validates the jobs and lays out their rows; grows the device arrays as needed. */
int batch_set_jobs(miso_batch* b, const miso_volume* vols, const float* isovals, int64_t njobs)
{
    if (njobs <= 0 || njobs > (1 << 24)) return fail(MISO_ERR_ARG, "job count %lld out of range", (long long)njobs);
    if (!vols || !isovals) return fail(MISO_ERR_ARG, "volumes or isovalues NULL");
    std::vector<BatchJob> jobs(njobs);
    std::vector<int64_t> lane(njobs + 1, 0);
    int64_t R = 0, mw = 0;
    for (int64_t j = 0; j < njobs; ++j) {
        const miso_volume* v = &vols[j];
        if (int rc = check_volume(v)) return rc;
        if (!v->on_device) return fail(MISO_ERR_ARG, "job %lld: batch volumes must be device-resident", (long long)j);
        BatchJob& J = jobs[j];
        J.g = make_geom(v);
        if (!has_cells(J.g)) return fail(MISO_ERR_ARG, "job %lld: dimensions below 2", (long long)j);
        if (!geom_lane_rows(J.g)) return fail(MISO_ERR_ARG, "job %lld: volume too large for a batch", (long long)j);
        if (3 * J.g.nx * J.g.ny * J.g.nz >= (int64_t(1) << 32))
            return fail(MISO_ERR_ARG, "job %lld: vertex ids may exceed 32 bits", (long long)j);
        J.vol = v->data;
        J.iso = isovals[j];
        J.unit = geom_unit(J.g) ? 1 : 0;
        J.row_base = R;
        J.mask_base = mw;
        const int64_t rows = J.g.ny * J.g.nz;
        R += rows;
        mw += rows * J.g.W;
        const int64_t per_block = 4 * geom_rows_per_wave(J.g);
        lane[j + 1] = lane[j] + (rows + per_block - 1) / per_block;
    }
    // classify groups: runs of consecutive jobs on one volume with the same dimensions (one
    // volume at several isovalues), whose rows pass 1 then loads once for all of them
    std::vector<int> grp{0};
    std::vector<int64_t> cls{0};
    for (int64_t j = 0; j < njobs;) {
        const Geom& g = jobs[j].g;
        int64_t e = j + 1;
        while (e < njobs && e - j < kClassifyGroup && jobs[e].vol == jobs[j].vol && jobs[e].g.nx == g.nx &&
               jobs[e].g.ny == g.ny && jobs[e].g.nz == g.nz)
            ++e;
        const int64_t per_item = batch_classify_rows(g);
        cls.push_back(cls.back() + (g.ny * g.nz + per_item - 1) / per_item);
        grp.push_back((int)e);
        j = e;
    }
    HIPCALL(hipStreamSynchronize(b->stream));   // the previous job set is no longer in use
    destroy_graphs(b);
    b->ran = false;
    b->njobs = (int)njobs;
    b->jobs.swap(jobs);
    b->cls_prefix.swap(cls);
    b->group_job.swap(grp);
    b->ngroups = (int)b->group_job.size() - 1;
    b->lane_prefix.swap(lane);
    b->R = R;
    b->mask_words = mw;
    b->ntiles = (R + kScanTile - 1) / kScanTile;
    b->state_words = 13 + 1 + 2 * b->ntiles;
    // pass 4's XCD queues: XCD x owns the dispatch indices 8 t + x of each job's row blocks
    b->xcd_prefix.assign(8 * (size_t)(njobs + 1), 0);
    for (int x = 0; x < 8; ++x)
        for (int64_t j = 0; j < njobs; ++j) {
            const int64_t n = b->lane_prefix[j + 1] - b->lane_prefix[j];
            b->xcd_prefix[x * (njobs + 1) + j + 1] = b->xcd_prefix[x * (njobs + 1) + j] + (n > x ? (n - x + 7) / 8 : 0);
        }
    int rc = 0;
    if ((rc = grow(&b->d_jobs, &b->cap_jobs, njobs))) return rc;
    if ((rc = grow(&b->d_cls_prefix, &b->cap_prefix_c, njobs + 1))) return rc;
    if ((rc = grow(&b->d_group_job, &b->cap_groups, njobs + 1))) return rc;
    if ((rc = grow(&b->d_lane_prefix, &b->cap_prefix_l, njobs + 1))) return rc;
    if ((rc = grow(&b->d_xcd_prefix, &b->cap_xcd, 8 * (njobs + 1)))) return rc;
    if ((rc = grow(&b->d_mask, &b->cap_mask, mw))) return rc;
    if ((rc = grow(&b->d_counts, &b->cap_counts, R))) return rc;
    if ((rc = grow(&b->d_trims, &b->cap_trims, R))) return rc;
    if ((rc = grow(&b->d_voff, &b->cap_voff, R))) return rc;
    if ((rc = grow(&b->d_toff, &b->cap_toff, R))) return rc;
    if ((rc = grow(&b->d_state, &b->cap_state, b->state_words))) return rc;
    if ((rc = grow(&b->d_jobout, &b->cap_jobout, 4 * njobs))) return rc;
    if (njobs + 1 > b->cap_host) {
        if (b->h_jobout) (void)hipHostFree(b->h_jobout);
        b->h_jobout = nullptr;
        b->cap_host = 0;
        HIPCALL(hipHostMalloc((void**)&b->h_jobout, (4 * (size_t)njobs + 2) * sizeof(uint64_t), hipHostMallocDefault));
        b->cap_host = njobs + 1;
    }
    HIPCALL(hipMemcpyAsync(b->d_jobs, b->jobs.data(), njobs * sizeof(BatchJob), hipMemcpyHostToDevice, b->stream));
    HIPCALL(hipMemcpyAsync(b->d_cls_prefix, b->cls_prefix.data(), (b->ngroups + 1) * sizeof(int64_t),
                           hipMemcpyHostToDevice, b->stream));
    HIPCALL(hipMemcpyAsync(b->d_group_job, b->group_job.data(), (b->ngroups + 1) * sizeof(int), hipMemcpyHostToDevice,
                           b->stream));
    HIPCALL(hipMemcpyAsync(b->d_lane_prefix, b->lane_prefix.data(), (njobs + 1) * sizeof(int64_t),
                           hipMemcpyHostToDevice, b->stream));
    HIPCALL(hipMemcpyAsync(b->d_xcd_prefix, b->xcd_prefix.data(), 8 * (njobs + 1) * sizeof(int64_t),
                           hipMemcpyHostToDevice, b->stream));
    HIPCALL(hipStreamSynchronize(b->stream));
    return MISO_OK;
}

int batch_init(miso_batch* b, const miso_volume* vols, const float* isovals, int64_t njobs, void* user_stream)
{
    if (int rc = batch_setup(b, user_stream)) return rc;
    if (int rc = batch_set_jobs(b, vols, isovals, njobs)) return rc;
    return capture(b, &b->g_scan, [&] { return enqueue_scan(b); });
}

/* Passes 1-4 as one cooperative launch of k_batch_fused (all blocks resident at once) after a
 * small kernel clears its tickets and scan state (cheaper there than before a grid barrier that
 * every block reaches at once), then the per-job counts and totals to the host. */
int enqueue_fused(miso_batch* b)
{
    const bool xcd = b->queue_order == MISO_BATCH_QUEUE_XCD;
    int64_t grid = (int64_t)b->num_cus * std::min(b->blocks_per_cu, b->fused_per_cu);
    if (xcd) grid = grid / 8 * 8;   // every XCD queue has blocks
    if (grid < 8) return fail(MISO_ERR_ARG, "fused launch: %lld resident blocks", (long long)grid);
    const BatchJob* jobs = b->d_jobs;
    const int* group_job = b->d_group_job;
    const int64_t* cls_prefix = b->d_cls_prefix;
    const int64_t* lane_prefix = b->d_lane_prefix;
    const int64_t* xprefix = xcd ? b->d_xcd_prefix : nullptr;
    int ngroups = b->ngroups, njobs = b->njobs, chunk = b->chunk;
    int64_t ncls = b->cls_prefix.back(), nlane = b->lane_prefix.back();
    int64_t R = b->R, ntiles = b->ntiles;
    uint64_t cap_v = b->cap_v, cap_t = b->cap_t;
    void* args[] = {&jobs,      &group_job, &cls_prefix, &ngroups,     &ncls,  &lane_prefix, &njobs,
                    &nlane,     &chunk,     &b->d_mask,  &b->d_counts, &b->d_trims, &b->d_voff, &b->d_toff,
                    &b->d_state, &R,        &ntiles,     &b->d_jobout, &cap_v, &cap_t,     &b->d_pts,
                    &b->d_nrm,  &b->d_tris, &xprefix};
    hipLaunchKernelGGL(k_batch_zero, dim3((unsigned)std::min<int64_t>(1024, (b->state_words + 255) / 256)), dim3(256),
                       0, b->stream, b->d_state, b->state_words);
    HIPCALL(hipLaunchCooperativeKernel((const void*)k_batch_fused, dim3((unsigned)grid), dim3(256), args, 0,
                                       b->stream));
    HIPCALL(hipMemcpyAsync(b->h_jobout, b->d_jobout, 4 * (size_t)b->njobs * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           b->stream));
    HIPCALL(hipMemcpyAsync(b->h_jobout + 4 * b->njobs, b->totals(), 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           b->stream));
    return MISO_OK;
}

int batch_run(miso_batch* b)
{
    if (b->njobs == 0) return fail(MISO_ERR_STATE, "batch has no jobs");
    if (!b->g_scan)
        if (int rc = capture(b, &b->g_scan, [&] { return enqueue_scan(b); })) return rc;
    HIPCALL(hipEventRecord(b->ev0, b->stream));
    const uint64_t* tot = b->h_jobout + 4 * b->njobs;
    if (b->ran && b->d_pts && b->g_emit && b->launch == MISO_BATCH_LAUNCH_FUSED) {
        // as the graph below: pass 4 writes nothing if the totals outgrew the buffers
        if (int rc = enqueue_fused(b)) return rc;
        HIPCALL(hipStreamSynchronize(b->stream));
        if (tot[0] > b->cap_v || tot[1] > b->cap_t) {
            if (int rc = batch_reserve(b, tot[0], tot[1])) return rc;
            HIPCALL(hipGraphLaunch(b->g_emit, b->stream));
        }
    } else if (b->ran && b->g_all) {
        // capacity known from the last run: all four passes in one graph; the emit pass writes
        // nothing if the totals outgrew the buffers, and is then run again below
        HIPCALL(hipGraphLaunch(b->g_all, b->stream));
        HIPCALL(hipStreamSynchronize(b->stream));
        if (tot[0] > b->cap_v || tot[1] > b->cap_t) {
            if (int rc = batch_reserve(b, tot[0], tot[1])) return rc;
            HIPCALL(hipGraphLaunch(b->g_emit, b->stream));
        }
    } else {
        HIPCALL(hipGraphLaunch(b->g_scan, b->stream));
        HIPCALL(hipStreamSynchronize(b->stream));
        if (int rc = batch_reserve(b, tot[0], tot[1])) return rc;
        if (!b->g_emit)
            if (int rc = batch_capture_emit(b)) return rc;
        HIPCALL(hipGraphLaunch(b->g_emit, b->stream));
    }
    HIPCALL(hipEventRecord(b->ev1, b->stream));
    HIPCALL(hipStreamSynchronize(b->stream));
    HIPCALL(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
    b->ran = true;
    b->counted = true;
    return MISO_OK;
}

}  // namespace

extern "C" int miso_batch_create(const miso_volume* volumes, const float* isovals, int64_t njobs, void* stream,
                                 miso_batch** out)
{
    if (!out) return fail(MISO_ERR_ARG, "out is NULL");
    *out = nullptr;
    miso_batch* b = new (std::nothrow) miso_batch();
    if (!b) return fail(MISO_ERR_ALLOC, "out of host memory");
    if (int rc = batch_init(b, volumes, isovals, njobs, stream)) {
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return MISO_OK;
}

extern "C" int miso_batch_run(miso_batch* b)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    return batch_run(b);
}

extern "C" int miso_batch_totals(const miso_batch* b, uint64_t* nverts, uint64_t* ntris)
{
    if (!b || !b->ran) return fail(MISO_ERR_STATE, "batch has not run");
    if (nverts) *nverts = b->h_jobout[4 * b->njobs];
    if (ntris) *ntris = b->h_jobout[4 * b->njobs + 1];
    return MISO_OK;
}

extern "C" int miso_batch_job_counts(const miso_batch* b, int64_t job, uint64_t* nverts, uint64_t* ntris)
{
    if (!b || !b->ran) return fail(MISO_ERR_STATE, "batch has not run");
    if (job < 0 || job >= b->njobs) return fail(MISO_ERR_ARG, "job %lld out of range", (long long)job);
    if (nverts) *nverts = b->h_jobout[4 * job + 2];
    if (ntris) *ntris = b->h_jobout[4 * job + 3];
    return MISO_OK;
}

extern "C" int miso_batch_job_output(const miso_batch* b, int64_t job, const float** points, const float** normals,
                                     const uint32_t** triangles)
{
    if (!b || !b->ran) return fail(MISO_ERR_STATE, "batch has not run");
    if (job < 0 || job >= b->njobs) return fail(MISO_ERR_ARG, "job %lld out of range", (long long)job);
    const uint64_t v0 = b->h_jobout[4 * job], t0 = b->h_jobout[4 * job + 1];
    if (points) *points = b->d_pts + 3 * v0;
    if (normals) *normals = b->d_nrm + 3 * v0;
    if (triangles) *triangles = b->d_tris + 3 * t0;
    return MISO_OK;
}

extern "C" int miso_batch_copy_out(miso_batch* b, int64_t job, float* points, float* normals, uint64_t* triangles)
{
    if (!b || !b->ran) return fail(MISO_ERR_STATE, "batch has not run");
    if (job < 0 || job >= b->njobs) return fail(MISO_ERR_ARG, "job %lld out of range", (long long)job);
    const uint64_t v0 = b->h_jobout[4 * job], t0 = b->h_jobout[4 * job + 1];
    const uint64_t nv = b->h_jobout[4 * job + 2], nt = b->h_jobout[4 * job + 3];
    if (points && nv)
        HIPCALL(hipMemcpyAsync(points, b->d_pts + 3 * v0, 3 * nv * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (normals && nv)
        HIPCALL(hipMemcpyAsync(normals, b->d_nrm + 3 * v0, 3 * nv * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (triangles && nt) {
        uint32_t* tmp = reinterpret_cast<uint32_t*>(triangles) + 3 * nt;
        HIPCALL(hipMemcpyAsync(tmp, b->d_tris + 3 * t0, 3 * nt * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
        HIPCALL(hipStreamSynchronize(b->stream));
        for (uint64_t q = 0; q < 3 * nt; ++q) triangles[q] = tmp[q];
    }
    HIPCALL(hipStreamSynchronize(b->stream));
    return MISO_OK;
}

extern "C" int miso_batch_set_tuning(miso_batch* b, int blocks_per_cu, int chunk)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    if (blocks_per_cu < 1 || blocks_per_cu > 64 || chunk < 1 || chunk > 1024)
        return fail(MISO_ERR_ARG, "tuning out of range (blocks_per_cu 1..64, chunk 1..1024)");
    if (blocks_per_cu == b->blocks_per_cu && chunk == b->chunk) return MISO_OK;
    HIPCALL(hipStreamSynchronize(b->stream));
    b->blocks_per_cu = blocks_per_cu;
    b->chunk = chunk;
    destroy_graphs(b);
    if (b->njobs == 0) return MISO_OK;
    if (int rc = capture(b, &b->g_scan, [&] { return enqueue_scan(b); })) return rc;
    if (b->d_pts) return batch_capture_emit(b);
    return MISO_OK;
}

extern "C" int miso_batch_set_queue_order(miso_batch* b, int order)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    if (order != MISO_BATCH_QUEUE_DISPATCH && order != MISO_BATCH_QUEUE_XCD)
        return fail(MISO_ERR_ARG, "queue order is MISO_BATCH_QUEUE_DISPATCH or MISO_BATCH_QUEUE_XCD");
    if (order == b->queue_order) return MISO_OK;
    HIPCALL(hipStreamSynchronize(b->stream));
    b->queue_order = order;
    if (b->g_emit || b->g_all) {
        if (b->g_emit) (void)hipGraphExecDestroy(b->g_emit);
        if (b->g_all) (void)hipGraphExecDestroy(b->g_all);
        b->g_emit = b->g_all = nullptr;
        if (b->d_pts) return batch_capture_emit(b);
    }
    return MISO_OK;
}

extern "C" int miso_batch_get_queue_order(const miso_batch* b, int* order)
{
    if (!b || !order) return fail(MISO_ERR_ARG, "NULL argument");
    *order = b->queue_order;
    return MISO_OK;
}

extern "C" int miso_batch_set_launch(miso_batch* b, int mode)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    if (mode != MISO_BATCH_LAUNCH_GRAPHS && mode != MISO_BATCH_LAUNCH_FUSED)
        return fail(MISO_ERR_ARG, "launch mode is MISO_BATCH_LAUNCH_GRAPHS or MISO_BATCH_LAUNCH_FUSED");
    if (mode == MISO_BATCH_LAUNCH_FUSED && b->fused_per_cu < 1)
        return fail(MISO_ERR_ARG, "the device has no cooperative launch for the fused kernel");
    b->launch = mode;
    return MISO_OK;
}

extern "C" int miso_batch_get_launch(const miso_batch* b, int* mode)
{
    if (!b || !mode) return fail(MISO_ERR_ARG, "NULL argument");
    *mode = b->launch;
    return MISO_OK;
}

extern "C" int miso_batch_get_tuning(const miso_batch* b, int* blocks_per_cu, int* chunk)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
    if (chunk) *chunk = b->chunk;
    return MISO_OK;
}

extern "C" int miso_batch_autotune(miso_batch* b, int reps, float* best_ms)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    if (reps < 1) reps = 1;
    static const int kBlocks[] = {2, 4, 8};
    static const int kChunk[] = {1, 4, 16};
    static const int kOrder[] = {MISO_BATCH_QUEUE_DISPATCH, MISO_BATCH_QUEUE_XCD};
    static const int kLaunch[] = {MISO_BATCH_LAUNCH_GRAPHS, MISO_BATCH_LAUNCH_FUSED};
    float best = -1.0f;
    int bb = b->blocks_per_cu, bc = b->chunk, bo = b->queue_order, bl = b->launch;
    for (int l : kLaunch)
        for (int o : kOrder)
            for (int x : kBlocks)
                for (int c : kChunk) {
                    if (l == MISO_BATCH_LAUNCH_FUSED && (b->fused_per_cu < 1 || x > b->fused_per_cu)) continue;
                    if (int rc = miso_batch_set_tuning(b, x, c)) return rc;
                    if (int rc = miso_batch_set_queue_order(b, o)) return rc;
                    if (int rc = miso_batch_set_launch(b, l)) return rc;
                    if (int rc = batch_run(b)) return rc;   // warm: sizes the output, captures the graphs
                    float sum = 0.0f;
                    for (int r = 0; r < reps; ++r) {
                        if (int rc = batch_run(b)) return rc;
                        sum += b->last_ms;
                    }
                    const float ms = sum / reps;
                    if (best < 0.0f || ms < best) {
                        best = ms;
                        bb = x;
                        bc = c;
                        bo = o;
                        bl = l;
                    }
                }
    if (int rc = miso_batch_set_tuning(b, bb, bc)) return rc;
    if (int rc = miso_batch_set_queue_order(b, bo)) return rc;
    if (int rc = miso_batch_set_launch(b, bl)) return rc;
    if (best_ms) *best_ms = best;
    return MISO_OK;
}

extern "C" int miso_batch_create_empty(void* stream, miso_batch** out)
{
    if (!out) return fail(MISO_ERR_ARG, "out is NULL");
    *out = nullptr;
    miso_batch* b = new (std::nothrow) miso_batch();
    if (!b) return fail(MISO_ERR_ALLOC, "out of host memory");
    if (int rc = batch_setup(b, stream)) {
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return MISO_OK;
}

extern "C" int miso_batch_set_jobs(miso_batch* b, const miso_volume* volumes, const float* isovals, int64_t njobs)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    b->counted = false;
    return batch_set_jobs(b, volumes, isovals, njobs);
}

extern "C" int miso_batch_count_async(miso_batch* b)
{
    if (!b || b->njobs == 0) return fail(MISO_ERR_STATE, "batch has no jobs");
    b->counted = false;
    HIPCALL(hipEventRecord(b->ev0, b->stream));
    return enqueue_scan(b);
}

extern "C" int miso_batch_query(miso_batch* b)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(MISO_ERR_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

extern "C" int miso_batch_count_wait(miso_batch* b)
{
    if (!b || b->njobs == 0) return fail(MISO_ERR_STATE, "batch has no jobs");
    HIPCALL(hipStreamSynchronize(b->stream));
    b->counted = true;
    b->ran = true;   // per-job counts and totals are valid
    return MISO_OK;
}

extern "C" int miso_batch_job_offsets(const miso_batch* b, int64_t job, uint64_t* first_vertex,
                                      uint64_t* first_triangle)
{
    if (!b || !b->counted) return fail(MISO_ERR_STATE, "batch has not been counted");
    if (job < 0 || job >= b->njobs) return fail(MISO_ERR_ARG, "job %lld out of range", (long long)job);
    if (first_vertex) *first_vertex = b->h_jobout[4 * job];
    if (first_triangle) *first_triangle = b->h_jobout[4 * job + 1];
    return MISO_OK;
}

extern "C" int miso_batch_emit_async(miso_batch* b, int64_t njobs_emit, float* points, float* normals,
                                     uint32_t* triangles, uint64_t cap_vertices, uint64_t cap_triangles)
{
    if (!b || !b->counted) return fail(MISO_ERR_STATE, "batch has not been counted");
    if (njobs_emit < 0 || njobs_emit > b->njobs) return fail(MISO_ERR_ARG, "njobs_emit out of range");
    const uint64_t nv = njobs_emit < b->njobs ? b->h_jobout[4 * njobs_emit] : b->h_jobout[4 * b->njobs];
    const uint64_t nt = njobs_emit < b->njobs ? b->h_jobout[4 * njobs_emit + 1] : b->h_jobout[4 * b->njobs + 1];
    if (nv > cap_vertices || nt > cap_triangles)
        return fail(MISO_ERR_ARG, "output of %llu vertices / %llu triangles exceeds the buffers", (unsigned long long)nv,
                    (unsigned long long)nt);
    if ((nv && (!points || !normals)) || (nt && !triangles)) return fail(MISO_ERR_ARG, "output buffer NULL");
    if (int rc = enqueue_emit_into(b, njobs_emit, points, normals, triangles, cap_vertices, cap_triangles)) return rc;
    HIPCALL(hipEventRecord(b->ev1, b->stream));
    return MISO_OK;
}

extern "C" int miso_batch_wait(miso_batch* b)
{
    if (!b) return fail(MISO_ERR_ARG, "batch is NULL");
    HIPCALL(hipStreamSynchronize(b->stream));
    return MISO_OK;
}

extern "C" int miso_batch_time_ms(const miso_batch* b, float* ms)
{
    if (!b || !b->ran) return fail(MISO_ERR_STATE, "batch has not run");
    if (ms) *ms = b->last_ms;
    return MISO_OK;
}

extern "C" void miso_batch_destroy(miso_batch* b)
{
    if (!b) return;
    batch_release(b);
    delete b;
}
