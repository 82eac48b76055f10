/*
 * tgx_frame.hip — the batched multi-camera frame as one captured hipGraph,
 * and the autotuner that picks the tuning (launch shapes and queue policies)
 * the graph is captured with.
 *
 * A frame is: visibility of every camera in one batched call, then per
 * camera either the plain one-bounce GI (tgx_gi_device's passes) or the
 * neural-cache frame (tgx_nrc_frame_device's passes after visibility).
 * Creating a graph runs the frame once as ordinary launches, which sizes
 * every context buffer, then records the same launches with stream capture
 * and instantiates them. Replays advance the cache's training through its
 * device counters. While a graph exists the context neither grows nor frees
 * buffers, since the graph holds their addresses.
 */
#include <algorithm>

#include "tgx_gi_dev.h"

#include <chrono>

struct tgx_frame_graph {
    hipGraph_t      graph = nullptr;
    hipGraphExec_t  exec = nullptr;
    void*           cams_dev = nullptr;   /* visibility camera records */
    uint32_t*       acc = nullptr;        /* frame totals, see gi_acc_k */
    std::vector<tgx_camera> cams;
    uint32_t        w = 0, h = 0, flags = 0;
    tgx_gi_params   gi = {};
    uint64_t*       vis = nullptr;
    float*          rad = nullptr;
    float*          irr = nullptr;
};

/* totals: 0 visible pixels, 1 bounce rays, 2 shadow rays, 3 bounce hits,
 * 4 training rays, 5 fallback pixels */
#define ACC_N 8u

__global__ void frame_clear_k(uint32_t* acc)
{
    if (threadIdx.x < ACC_N) acc[threadIdx.x] = 0u;
}

/* Adds one camera's GI counters to the frame totals. mode: 1 the plain
 * block ran, 2 training ran, 4 the cache's fallback count is valid. */
__global__ void gi_acc_k(const uint32_t* __restrict__ gcnt, const uint32_t* __restrict__ nrc_q, uint32_t mode,
                         uint32_t* __restrict__ acc)
{
    acc[0] += gcnt[GI_CNT_VISIBLE];
    if (mode & 1u) {
        acc[1] += gcnt[GI_CNT_PLAIN + 1];
        acc[2] += gcnt[GI_CNT_PLAIN + 3];
        acc[3] += gcnt[GI_CNT_PLAIN + 2];
    }
    if (mode & 2u) acc[4] += gcnt[GI_CNT_TRAIN + 1] + gcnt[GI_CNT_TRAIN + 3];
    if ((mode & 4u) && nrc_q) acc[5] += nrc_q[1];
}

/* Frame totals of a multi-camera plain GI frame traced in one launch: the
 * cameras' visible counts and the shared counters. */
__global__ void gi_acc_all_k(const uint32_t* __restrict__ counts, uint32_t n, const uint32_t* __restrict__ gcnt,
                             uint32_t* __restrict__ acc)
{
    if (threadIdx.x != 0) return;
    uint32_t v = 0;
    for (uint32_t c = 0; c < n; c++) v += counts[c];
    acc[0] += v;
    acc[1] += gcnt[GI_CNT_PLAIN + 1];
    acc[2] += gcnt[GI_CNT_PLAIN + 3];
    acc[3] += gcnt[GI_CNT_PLAIN + 2];
}

/* One frame as stream-ordered launches; `record` selects the capture-safe
 * visibility path. */
static int frame_body(tgx_context* ctx, tgx_frame_graph* g, bool record)
{
    const uint32_t n = (uint32_t)g->cams.size();
    const size_t px = (size_t)g->w * g->h;
    int rc;
    hipLaunchKernelGGL(frame_clear_k, dim3(1), dim3(64), 0, ctx->stream, g->acc);
    TGX_LAUNCH_CHECK(ctx, "frame_clear_k");
    rc = record ? tgx_visibility_record(ctx, g->cams_dev, g->cams.data(), n, g->w, g->h, g->vis)
                : tgx_visibility_device(ctx, g->cams.data(), n, g->w, g->h, g->vis);
    if (rc) return rc;
    const bool neural = (g->flags & TGX_FRAME_NEURAL) != 0;
    uint32_t mode = 1u;
    if (neural) {
        const uint32_t nf = ctx->nrc_cfg.flags;
        mode = ((nf & TGX_NRC_FLAG_NO_FALLBACK) ? 0u : 5u) | ((nf & TGX_NRC_FLAG_NO_TRAINING) ? 0u : 2u);
    }
    /* Cameras go through the merged paths in groups whose per-camera state
     * stays bounded: at most 2^26 pixels (32 cameras at 1080p) and 2^28
     * bounce samples per group. */
    const uint64_t by_px = (1ull << 26) / px, by_samples = (1ull << 28) / (px * g->gi.spp);
    const uint32_t group = (uint32_t)std::max<uint64_t>(1u, std::min<uint64_t>({(uint64_t)n, by_px, by_samples}));
    if (!neural && n > 1 && group > 1) {
        /* Plain GI of every camera of a group through one ray queue and one
         * trace launch: each camera's surfaces and pixel list in its own
         * slice. */
        if ((rc = tgx_buffer_reserve(ctx, &ctx->gi_counts, 4ull * GI_CNT_TOTAL, "gi counters")) ||
            (rc = tgx_buffer_reserve(ctx, &ctx->gi_bsurf, sizeof(gi_surf_rec) * group * px, "gi surfaces per camera")) ||
            (rc = tgx_buffer_reserve(ctx, &ctx->gi_blist, 12ull * group * px, "gi pixel lists per camera")) ||
            (rc = tgx_buffer_reserve(ctx, &ctx->gi_bcount, 4ull * group, "gi visible counts per camera")))
            return rc;
        gi_surf_rec* surf = (gi_surf_rec*)ctx->gi_bsurf.ptr;
        uint32_t* lists = (uint32_t*)ctx->gi_blist.ptr;
        uint32_t* counts = (uint32_t*)ctx->gi_bcount.ptr;
        uint32_t* cnt = (uint32_t*)ctx->gi_counts.ptr;
        for (uint32_t c0 = 0; c0 < n; c0 += group) {
            const uint32_t m = std::min(group, n - c0);
            std::vector<gi_trace_set> sets(m);
            for (uint32_t k = 0; k < m; k++) {
                const uint32_t c = c0 + k;
                float* rad = g->rad + 3 * c * px;
                float* irr = g->irr ? g->irr + 3 * c * px : nullptr;
                if ((rc = tgx_gi_surfaces_at(ctx, &g->cams[c], g->w, g->h, g->vis + c * px, rad, irr, surf + k * px,
                                             lists + 3 * k * px, counts + k)))
                    return rc;
                sets[k] = {g->gi, lists + 3 * k * px, counts + k, (uint32_t)px, cnt + GI_CNT_PLAIN, nullptr, nullptr,
                           rad, irr, nullptr, surf + k * px, 0u};
            }
            if ((rc = tgx_gi_trace_sets(ctx, sets.data(), m))) return rc;
            hipLaunchKernelGGL(gi_acc_all_k, dim3(1), dim3(64), 0, ctx->stream, (const uint32_t*)counts, m,
                               (const uint32_t*)cnt, g->acc);
            TGX_LAUNCH_CHECK(ctx, "gi_acc_all_k");
        }
        return TGX_OK;
    }
    if (neural && n > 1 && !(ctx->nrc_cfg.flags & TGX_NRC_FLAG_NO_TRAINING)) {
        /* groups run in camera order, so the result is the camera-by-camera
         * sequence's either way */
        for (uint32_t c0 = 0; c0 < n; c0 += group) {
            const uint32_t m = std::min(group, n - c0);
            if ((rc = tgx_nrc_frames_from_vis(ctx, g->cams.data() + c0, m, g->w, g->h, &g->gi, g->vis + c0 * px,
                                              g->rad + 3 * c0 * px, g->irr ? g->irr + 3 * c0 * px : nullptr, g->acc)))
                return rc;
        }
        return TGX_OK;
    }
    for (uint32_t c = 0; c < n; c++) {
        const uint64_t* vis = g->vis + c * px;
        float* rad = g->rad + 3 * c * px;
        float* irr = g->irr ? g->irr + 3 * c * px : nullptr;
        if (neural) {
            if ((rc = tgx_nrc_frame_from_vis(ctx, &g->cams[c], g->w, g->h, &g->gi, vis, rad, irr))) return rc;
        } else {
            if ((rc = tgx_gi_surfaces(ctx, &g->cams[c], g->w, g->h, vis, rad, irr))) return rc;
            uint32_t* cnt = (uint32_t*)ctx->gi_counts.ptr;
            if ((rc = tgx_gi_trace_list(ctx, &g->gi, (const uint32_t*)ctx->gi_list.ptr, cnt + GI_CNT_VISIBLE,
                                        g->w * g->h, cnt + GI_CNT_PLAIN, nullptr, nullptr, rad, irr, nullptr)))
                return rc;
        }
        hipLaunchKernelGGL(gi_acc_k, dim3(1), dim3(1), 0, ctx->stream, (const uint32_t*)ctx->gi_counts.ptr,
                           neural ? (const uint32_t*)ctx->nrc_queue.ptr : nullptr, mode, g->acc);
        TGX_LAUNCH_CHECK(ctx, "gi_acc_k");
    }
    return TGX_OK;
}

static int check_frame_args(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                            const tgx_gi_params* gi, uint32_t flags, uint64_t* vis, float* rad)
{
    if (!cams || !gi || !vis || !rad || n_cams == 0 || n_cams > 1024)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "frame needs 1..1024 cameras and output buffers");
    if (flags & ~TGX_FRAME_NEURAL) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "unknown frame flags");
    if ((flags & TGX_FRAME_NEURAL) && !ctx->nrc_ready)
        return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "a neural frame needs tgx_nrc_create first");
    return tgx_gi_validate(ctx, w, h, gi);
}

static void fill_graph(tgx_frame_graph* g, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                       const tgx_gi_params* gi, uint32_t flags, uint64_t* vis, float* rad, float* irr)
{
    g->cams.assign(cams, cams + n_cams);
    g->w = w; g->h = h; g->flags = flags; g->gi = *gi;
    g->vis = vis; g->rad = rad; g->irr = irr;
}

static int read_totals(tgx_context* ctx, const uint32_t* acc, uint64_t primary, tgx_frame_stats* out)
{
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t a[ACC_N];
    TGX_HIP_CHECK(ctx, hipMemcpy(a, acc, sizeof(a), hipMemcpyDeviceToHost));
    out->primary_rays = primary;
    out->visible_pixels = a[0];
    out->bounce_rays = a[1];
    out->shadow_rays = a[2];
    out->bounce_hits = a[3];
    out->train_rays = a[4];
    out->fallback_pixels = a[5];
    return TGX_OK;
}

extern "C" {

int tgx_frame_batch(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                    const tgx_gi_params* gi, uint32_t flags, uint64_t* vis_dev, float* radiance_dev,
                    float* irradiance_dev, tgx_frame_stats* stats_out)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    int rc;
    if ((rc = check_frame_args(ctx, cams, n_cams, w, h, gi, flags, vis_dev, radiance_dev))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->frame_acc, 4 * ACC_N, "frame totals"))) return rc;
    tgx_frame_graph g;
    fill_graph(&g, cams, n_cams, w, h, gi, flags, vis_dev, radiance_dev, irradiance_dev);
    g.acc = (uint32_t*)ctx->frame_acc.ptr;
    if ((rc = frame_body(ctx, &g, false))) return rc;
    if (stats_out) return read_totals(ctx, g.acc, (uint64_t)n_cams * w * h, stats_out);
    return TGX_OK;
}

int tgx_frame_graph_create(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                           const tgx_gi_params* gi, uint32_t flags, uint64_t* vis_dev, float* radiance_dev,
                           float* irradiance_dev, tgx_frame_graph** out)
{
    if (!ctx || !out) return TGX_ERR_INVALID_ARG;
    *out = nullptr;
    ctx->last_error.clear();
    int rc;
    if ((rc = check_frame_args(ctx, cams, n_cams, w, h, gi, flags, vis_dev, radiance_dev))) return rc;
    tgx_frame_graph* g = new tgx_frame_graph();
    fill_graph(g, cams, n_cams, w, h, gi, flags, vis_dev, radiance_dev, irradiance_dev);
    auto fail = [&](int code) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        if (g->cams_dev) (void)hipFree(g->cams_dev);
        if (g->acc) (void)hipFree(g->acc);
        delete g;
        return code;
    };
    hipError_t e = hipMalloc((void**)&g->acc, 4 * ACC_N);
    if (e != hipSuccess) return fail(tgx_fail(ctx, TGX_ERR_OUT_OF_MEMORY, "frame totals: %s", hipGetErrorString(e)));
    if ((rc = tgx_vis_cams_upload(ctx, cams, n_cams, w, h, &g->cams_dev))) return fail(rc);
    /* ordinary run: sizes every buffer the captured launches will use */
    if ((rc = frame_body(ctx, g, false))) return fail(rc);
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return fail(tgx_fail(ctx, TGX_ERR_HIP, "warm-up frame: %s", hipGetErrorString(e)));
    const bool prof = ctx->profiling;
    const tgx_tuning tuning = ctx->tuning;
    ctx->profiling = false;
    /* rocPRIM's radix sort clears its histograms with memsets, which a graph
     * must not hold: captured frames trace unsorted (same results) */
    ctx->tuning.trace_sort = 0u;
    ctx->n_graphs++; /* from here on no buffer may move */
    if ((e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed)) != hipSuccess) {
        ctx->n_graphs--;
        ctx->profiling = prof;
        ctx->tuning = tuning;
        return fail(tgx_fail(ctx, TGX_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e)));
    }
    rc = frame_body(ctx, g, true);
    e = hipStreamEndCapture(ctx->stream, &g->graph);
    ctx->profiling = prof;
    ctx->tuning = tuning;
    if (rc || e != hipSuccess) {
        ctx->n_graphs--;
        if (!rc) rc = tgx_fail(ctx, TGX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        return fail(rc);
    }
    /* refuse memset nodes: cleared memory is written by fill kernels only */
    size_t n_nodes = 0;
    if (hipGraphGetNodes(g->graph, nullptr, &n_nodes) == hipSuccess && n_nodes) {
        std::vector<hipGraphNode_t> nodes(n_nodes);
        if (hipGraphGetNodes(g->graph, nodes.data(), &n_nodes) == hipSuccess)
            for (size_t i = 0; i < n_nodes; i++) {
                hipGraphNodeType t;
                if (hipGraphNodeGetType(nodes[i], &t) == hipSuccess && t == hipGraphNodeTypeMemset) {
                    ctx->n_graphs--;
                    return fail(tgx_fail(ctx, TGX_ERR_HIP, "captured frame contains a memset node"));
                }
            }
    }
    if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess) {
        ctx->n_graphs--;
        return fail(tgx_fail(ctx, TGX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e)));
    }
    *out = g;
    return TGX_OK;
}

int tgx_frame_graph_launch(tgx_context* ctx, tgx_frame_graph* g)
{
    if (!ctx || !g || !g->exec) return TGX_ERR_INVALID_ARG;
    TGX_HIP_CHECK(ctx, hipGraphLaunch(g->exec, ctx->stream));
    return TGX_OK;
}

int tgx_frame_graph_stats(tgx_context* ctx, tgx_frame_graph* g, tgx_frame_stats* out)
{
    if (!ctx || !g || !out) return TGX_ERR_INVALID_ARG;
    return read_totals(ctx, g->acc, (uint64_t)g->cams.size() * g->w * g->h, out);
}

int tgx_frame_graph_destroy(tgx_context* ctx, tgx_frame_graph* g)
{
    if (!ctx || !g) return TGX_ERR_INVALID_ARG;
    (void)hipStreamSynchronize(ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->cams_dev) (void)hipFree(g->cams_dev);
    if (g->acc) (void)hipFree(g->acc);
    if (ctx->n_graphs) ctx->n_graphs--;
    delete g;
    return TGX_OK;
}

int tgx_autotune(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                 const tgx_gi_params* gi, uint32_t flags, uint64_t* vis_dev, float* radiance_dev,
                 float* irradiance_dev, uint32_t frames, tgx_tuning* best_out)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    int rc;
    if ((rc = check_frame_args(ctx, cams, n_cams, w, h, gi, flags, vis_dev, radiance_dev))) return rc;
    if (frames == 0) frames = 2;
    tgx_frame_graph g;
    fill_graph(&g, cams, n_cams, w, h, gi, flags, vis_dev, radiance_dev, irradiance_dev);
    if ((rc = tgx_buffer_reserve(ctx, &ctx->frame_acc, 4 * ACC_N, "frame totals"))) return rc;
    g.acc = (uint32_t*)ctx->frame_acc.ptr;
    hipError_t e;
    const tgx_tuning start = ctx->tuning;
    tgx_tuning best = start;
    double best_ms = 1e300;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    /* Coordinate descent over the launch shapes and queue policies, with
     * unsorted rays as captured frames trace them: each knob in turn takes
     * each of its values with the others at their best so far; a candidate
     * runs one untimed frame, then `frames` timed ones. No knob changes a
     * result. */
    struct knob {
        uint32_t tgx_tuning::*field;
        uint32_t n;
        uint32_t values[6];
    };
    const knob knobs[] = {
        {&tgx_tuning::trace_regroup, 4u, {0u, 1u, 2u, 4u}},          /* trace launch form, node steps per round */
        {&tgx_tuning::trace_skip, 2u, {0u, 1u}},                      /* exact empty-brick skip (regrouped form) */
        {&tgx_tuning::trace_blocks_per_cu, 2u, {1u, 2u}},             /* lane-machine workgroups per CU */
        {&tgx_tuning::trace_dda, 2u, {0u, 1u}},                       /* lane-machine DDA without iteration count */
        {&tgx_tuning::trace_node, 4u, {1u, 2u, 3u, 4u}},              /* node iterations per lane step */
        {&tgx_tuning::trace_cross, 5u, {0u, 24u, 32u, 40u, 48u}},     /* empty-brick crossing threshold */
        {&tgx_tuning::trace_refill, 6u, {4u, 5u, 6u, 8u, 12u, 16u}}, /* idle lanes that trigger a queue refill */
        {&tgx_tuning::trace_chunk, 3u, {0u, 32u, 64u}},               /* rays per queue cursor atomic */
        {&tgx_tuning::vis_setup_blocks, 4u, {256u, 1024u, 2048u, 4096u}}, /* cull / scatter workgroups per camera */
        {&tgx_tuning::vis_tile, 3u, {8u, 16u, 32u}},                  /* visibility tile edge */
    };
    best.trace_sort = 0u;
    best.trace_vote = 0u;
    auto time_candidate = [&](const tgx_tuning& t, double& ms_out) -> int {
        ctx->tuning = t;
        int r = frame_body(ctx, &g, false);
        if (r) return r;
        (void)hipEventRecord(t0, ctx->stream);
        for (uint32_t f = 0; f < frames && !r; f++) r = frame_body(ctx, &g, false);
        (void)hipEventRecord(t1, ctx->stream);
        if (r) return r;
        if ((e = hipEventSynchronize(t1)) != hipSuccess)
            return tgx_fail(ctx, TGX_ERR_HIP, "autotune timing: %s", hipGetErrorString(e));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        ms_out = ms;
        return TGX_OK;
    };
    rc = time_candidate(best, best_ms);
    for (const knob& kb : knobs) {
        if (rc) break;
        /* the skip exists in the regrouped form only, the workgroups per CU
         * in the lane-machine form only */
        if (kb.field == &tgx_tuning::trace_skip && best.trace_regroup == 0u) continue;
        if ((kb.field == &tgx_tuning::trace_blocks_per_cu || kb.field == &tgx_tuning::trace_dda ||
             kb.field == &tgx_tuning::trace_node || kb.field == &tgx_tuning::trace_cross ||
             kb.field == &tgx_tuning::trace_chunk) &&
            best.trace_regroup != 0u)
            continue;
        /* the crossing exists in the uncounted DDA only */
        if (kb.field == &tgx_tuning::trace_cross && best.trace_dda == 0u) continue;
        const uint32_t cur = best.*kb.field;
        for (uint32_t i = 0; i < kb.n && !rc; i++) {
            if (kb.values[i] == cur) continue;
            tgx_tuning t = best;
            t.*kb.field = kb.values[i];
            double ms = 0.0;
            if ((rc = time_candidate(t, ms))) {
                /* a graph holds the buffers and this candidate needs larger ones */
                if (rc == TGX_ERR_CAPACITY && ctx->n_graphs > 0) {
                    rc = TGX_OK;
                    ctx->last_error.clear();
                    continue;
                }
                break;
            }
            if (ms < best_ms) { best_ms = ms; best = t; }
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    ctx->tuning = rc ? start : best;
    if (!rc && best_out) *best_out = best;
    return rc;
}

} /* extern "C" */
