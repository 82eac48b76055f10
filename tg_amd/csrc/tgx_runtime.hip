/*
 * tgx_runtime.hip — context, scene upload, device memory and profiling
 * entry points of the C ABI (declared in include/tgx.h).
 */
#include "tgx_internal.h"

#include <algorithm>

int tgx_fail(tgx_context* ctx, int code, const char* fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

int tgx_buffer_reserve(tgx_context* ctx, tgx_buffer* b, size_t bytes, const char* what)
{
    if (bytes <= b->bytes && b->ptr) return TGX_OK;
    if (ctx->n_graphs > 0)
        return tgx_fail(ctx, TGX_ERR_CAPACITY,
                        "%s needs %zu bytes, but a frame graph holds the context's buffers: destroy it first", what, bytes);
    if (b->ptr) {
        /* The old contents are dead; wait for in-flight users first. */
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess)
            return tgx_fail(ctx, TGX_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
        (void)hipFree(b->ptr);
        b->ptr = nullptr;
        b->bytes = 0;
    }
    size_t alloc = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b->ptr, alloc);
    if (e != hipSuccess) {
        b->ptr = nullptr;
        return tgx_fail(ctx, TGX_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) for %s: %s", alloc, what,
                        hipGetErrorString(e));
    }
    b->bytes = alloc;
    return TGX_OK;
}

void tgx_buffer_release(tgx_buffer* b)
{
    if (b->ptr) (void)hipFree(b->ptr);
    b->ptr = nullptr;
    b->bytes = 0;
}

tgx_launch_scope::tgx_launch_scope(tgx_context* c, const char* name, bool on) : ctx(c)
{
    if (!on || !ctx->profiling) return;
    for (size_t i = 0; i < ctx->stats.size(); i++)
        if (strncmp(ctx->stats[i].name, name, sizeof(ctx->stats[i].name)) == 0) stat_idx = (int)i;
    if (stat_idx < 0) {
        tgx_kernel_stat s = {};
        strncpy(s.name, name, sizeof(s.name) - 1);
        ctx->stats.push_back(s);
        stat_idx = (int)ctx->stats.size() - 1;
    }
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
        stat_idx = -1;
        return;
    }
    (void)hipEventRecord(start, ctx->stream);
}

tgx_launch_scope::~tgx_launch_scope()
{
    if (stat_idx < 0) return;
    (void)hipEventRecord(stop, ctx->stream);
    ctx->pending.push_back({stat_idx, start, stop});
}

int tgx_scene_replace_begin(tgx_context* ctx, const char* what)
{
    if (ctx->n_graphs > 0)
        return tgx_fail(ctx, TGX_ERR_CAPACITY, "%s: a frame graph holds the scene's buffers: destroy it first", what);
    ctx->svo_stale = true;
    return TGX_OK;
}

int tgx_svo_check_current(tgx_context* ctx)
{
    if (!ctx->svo_built) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no SVO built");
    if (ctx->svo_stale)
        return tgx_fail(ctx, TGX_ERR_NOT_BUILT,
                        "the scene changed since the SVO was built: call tgx_svo_update or tgx_svo_build");
    return TGX_OK;
}

/* Grid-stride 16-byte copy. */
__global__ __launch_bounds__(256) void tgx_copy_k(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" {

int tgx_create(int device, tgx_context** out_ctx)
{
    if (!out_ctx) return TGX_ERR_INVALID_ARG;
    *out_ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TGX_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return TGX_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return TGX_ERR_HIP;
    tgx_context* ctx = new tgx_context();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return TGX_ERR_HIP;
    }
    if (hipHostMalloc(&ctx->host_pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return TGX_ERR_HIP;
    }
    if (tgx_vis_scratch_init(ctx) != TGX_OK) {
        (void)hipHostFree(ctx->host_pinned);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return TGX_ERR_HIP;
    }
    *out_ctx = ctx;
    return TGX_OK;
}

int tgx_destroy(tgx_context* ctx)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    tgx_buffer* bufs[] = {&ctx->objects, &ctx->cluster_pointers, &ctx->cluster_idx_to_object,
                          &ctx->bits, &ctx->materials, &ctx->luts, &ctx->vis_clusters, &ctx->vis_solid, &ctx->vis_objcnt,
                          &ctx->vis_occ, &ctx->vis_obj_occ, &ctx->vis_scan_sums,
                          &ctx->vis_rec_flags, &ctx->vis_global_keys,
                          &ctx->vis_counters, &ctx->vis_tile_count, &ctx->vis_tile_off,
                          &ctx->vis_tile_list, &ctx->vis_cams, &ctx->vis_out, &ctx->vis_stats_dev,
                          &ctx->frame_acc, &ctx->pick_buf, &ctx->dyn_stage, &ctx->dyn_tmp, &ctx->cl_dirty};
    for (tgx_buffer* b : bufs) tgx_buffer_release(b);
    tgx_svo_release(ctx);
    tgx_nrc_release(ctx);
    if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
    if (ctx->vis_stats_event) (void)hipEventDestroy(ctx->vis_stats_event);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TGX_OK;
}

const char* tgx_last_error(const tgx_context* ctx)
{
    return ctx ? ctx->last_error.c_str() : "null context";
}

int tgx_synchronize(tgx_context* ctx)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

static int upload(tgx_context* ctx, tgx_buffer* b, const void* src, size_t bytes, const char* what)
{
    int rc = tgx_buffer_reserve(ctx, b, bytes, what);
    if (rc) return rc;
    if (bytes && src) TGX_HIP_CHECK(ctx, hipMemcpyAsync(b->ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return TGX_OK;
}

int tgx_scene_upload(tgx_context* ctx, const tgx_scene_desc* s)
{
    if (!ctx || !s) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    ctx->scene_version++;
    if (!s->objects || !s->cluster_pointers || !s->cluster_idx_to_object || !s->bits)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "scene arrays must not be NULL");
    if (s->n_objects == 0 || s->n_cluster_pointers == 0 || s->n_clusters == 0)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "empty scene");
    if ((uint64_t)s->n_cluster_pointers >= (1ull << TGX_CLUSTER_POINTER_BITS))
        return tgx_fail(ctx, TGX_ERR_CAPACITY, "%u cluster pointers exceed the 31-bit field",
                        s->n_cluster_pointers);
    /* Validate the maps the kernels index with, so no kernel reads out of
     * bounds (the reference asserts in debug builds only). */
    for (uint32_t c = 0; c < s->n_cluster_pointers; c++)
        if (s->cluster_pointers[c] >= s->n_clusters)
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "cluster pointer %u -> idx %u >= %u", c,
                            s->cluster_pointers[c], s->n_clusters);
    for (uint32_t c = 0; c < s->n_clusters; c++)
        if (s->cluster_idx_to_object[c] >= s->n_objects)
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "cluster %u -> object %u >= %u", c,
                            s->cluster_idx_to_object[c], s->n_objects);
    uint64_t covered = 0;
    for (uint32_t o = 0; o < s->n_objects; o++) {
        const tgx_object* ob = &s->objects[o];
        const uint64_t n = (uint64_t)ob->n_cluster_pointers_per_dim[0] * ob->n_cluster_pointers_per_dim[1] *
                           ob->n_cluster_pointers_per_dim[2];
        if (n == 0 || (uint64_t)ob->first_cluster_pointer + n > s->n_cluster_pointers)
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "object %u cluster range out of bounds", o);
        if (s->luts && ob->lut_idx >= s->n_luts)
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "object %u lut_idx %u >= %u", o, ob->lut_idx,
                            s->n_luts);
        /* tg's model: an object's pointers map to clusters of that object */
        for (uint64_t c = ob->first_cluster_pointer; c < ob->first_cluster_pointer + n; c++)
            if (s->cluster_idx_to_object[s->cluster_pointers[c]] != o)
                return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "cluster pointer %llu of object %u maps to object %u",
                                (unsigned long long)c, o, s->cluster_idx_to_object[s->cluster_pointers[c]]);
        covered += n;
    }
    /* ... and the objects' ranges (disjoint by the check above) cover every pointer */
    if (covered != s->n_cluster_pointers)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "objects cover %llu of %u cluster pointers",
                        (unsigned long long)covered, s->n_cluster_pointers);
    int rc;
    if ((rc = tgx_scene_replace_begin(ctx, "tgx_scene_upload"))) return rc;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    ctx->has_scene = false;
    ctx->has_materials = false;
    ctx->dynamic = false;
    ctx->svo_prev_valid = false;
    ctx->dyn_objects.clear();
    ctx->dyn_free_objects.clear();
    ctx->dyn_free_clusters.clear();
    if ((rc = upload(ctx, &ctx->objects, s->objects, sizeof(tgx_object) * s->n_objects, "objects"))) return rc;
    if ((rc = upload(ctx, &ctx->cluster_pointers, s->cluster_pointers, 4ull * s->n_cluster_pointers, "cluster pointers"))) return rc;
    if ((rc = upload(ctx, &ctx->cluster_idx_to_object, s->cluster_idx_to_object, 4ull * s->n_clusters, "cluster->object"))) return rc;
    if ((rc = upload(ctx, &ctx->bits, s->bits, 64ull * s->n_clusters, "solid bits"))) return rc;
    if (s->materials)
        if ((rc = upload(ctx, &ctx->materials, s->materials, 512ull * s->n_clusters, "materials"))) return rc;
    if (s->luts)
        if ((rc = upload(ctx, &ctx->luts, s->luts, sizeof(tgx_lut_entry) * TGX_LUT_ENTRIES * s->n_luts, "luts"))) return rc;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_objects = s->n_objects;
    ctx->n_cluster_pointers = s->n_cluster_pointers;
    ctx->n_clusters = s->n_clusters;
    ctx->n_luts = s->luts ? s->n_luts : 0;
    ctx->has_materials = s->materials != nullptr;
    ctx->has_scene = true;
    return tgx_vis_solid_refresh(ctx);
}

int tgx_device_alloc(tgx_context* ctx, size_t bytes, void** out_dev)
{
    if (!ctx || !out_dev) return TGX_ERR_INVALID_ARG;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(out_dev, bytes ? bytes : 1);
    if (e != hipSuccess) return tgx_fail(ctx, TGX_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return TGX_OK;
}

int tgx_device_free(tgx_context* ctx, void* dev)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    TGX_HIP_CHECK(ctx, hipFree(dev));
    return TGX_OK;
}

int tgx_memcpy_d2h(tgx_context* ctx, void* dst_host, const void* src_dev, size_t bytes)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_memcpy_h2d(tgx_context* ctx, void* dst_dev, const void* src_host, size_t bytes)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_memcpy_d2d(tgx_context* ctx, void* dst_dev, const void* src_dev, size_t bytes)
{
    if (!ctx || (bytes && (!dst_dev || !src_dev)) || bytes % 16) return TGX_ERR_INVALID_ARG;
    if (bytes == 0) return TGX_OK;
    const uint64_t n = bytes / 16;
    const uint64_t blocks = (n + 255) / 256;
    const uint32_t grid = blocks < 65536 ? (uint32_t)blocks : 65536u;
    {
        tgx_launch_scope s(ctx, "tgx_copy_k");
        hipLaunchKernelGGL(tgx_copy_k, dim3(grid), dim3(256), 0, ctx->stream, (uint4*)dst_dev, (const uint4*)src_dev, n);
    }
    TGX_LAUNCH_CHECK(ctx, "tgx_copy_k");
    return TGX_OK;
}

int tgx_tuning_get(tgx_context* ctx, tgx_tuning* out)
{
    if (!ctx || !out) return TGX_ERR_INVALID_ARG;
    *out = ctx->tuning;
    return TGX_OK;
}

int tgx_tuning_set(tgx_context* ctx, const tgx_tuning* t)
{
    if (!ctx || !t) return TGX_ERR_INVALID_ARG;
    if (t->trace_sort > 3u || t->trace_refill < 1u || t->trace_refill > 64u || t->trace_blocks_per_cu < 1u ||
        t->trace_blocks_per_cu > 2u || t->trace_vote > 64u || t->trace_regroup > 8u || t->trace_skip > 1u ||
        t->vis_setup_blocks < 64u || t->vis_setup_blocks > 16384u || t->trace_dda > 1u ||
        (t->vis_tile != 8u && t->vis_tile != 16u && t->vis_tile != 32u) || t->trace_cross > 64u ||
        t->trace_node < 1u || t->trace_node > 4u || t->trace_burst < 2u || t->trace_burst > 32u ||
        (t->trace_burst & 1u) || t->trace_chunk > 256u || t->trace_small > 64u)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "tuning out of range");
    ctx->tuning = *t;
    return TGX_OK;
}

int tgx_profile_enable(tgx_context* ctx, int enable)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->profiling = enable != 0;
    return TGX_OK;
}

static void resolve_pending(tgx_context* ctx)
{
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            ctx->stats[p.stat_idx].launches += 1;
            ctx->stats[p.stat_idx].total_ms += ms;
        }
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    ctx->pending.clear();
}

int tgx_profile_reset(tgx_context* ctx)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    resolve_pending(ctx);
    ctx->stats.clear();
    return TGX_OK;
}

int tgx_profile_get(tgx_context* ctx, tgx_kernel_stat* out, uint32_t capacity, uint32_t* out_n)
{
    if (!ctx || !out_n) return TGX_ERR_INVALID_ARG;
    resolve_pending(ctx);
    const uint32_t n = (uint32_t)ctx->stats.size();
    for (uint32_t i = 0; i < n && i < capacity && out; i++) out[i] = ctx->stats[i];
    *out_n = n;
    return TGX_OK;
}

int tgx_visibility_decode(uint64_t word, float* out_depth, uint32_t* out_cluster_pointer,
                          uint32_t* out_voxel_idx)
{
    const float depth = (float)(word >> TGX_VIS_DEPTH_SHIFT) / 16777215.0f;
    if (out_depth) *out_depth = depth;
    if (depth < 1.0f) {
        if (out_cluster_pointer) *out_cluster_pointer = (uint32_t)(word >> TGX_VIS_CPTR_SHIFT) & TGX_VIS_CPTR_MASK;
        if (out_voxel_idx) *out_voxel_idx = (uint32_t)word & TGX_VIS_VOXEL_MASK;
        return 1;
    }
    if (out_cluster_pointer) *out_cluster_pointer = 0xFFFFFFFFu;
    if (out_voxel_idx) *out_voxel_idx = 0xFFFFFFFFu;
    return 0;
}

} /* extern "C" */
