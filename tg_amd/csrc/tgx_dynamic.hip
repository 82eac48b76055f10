/*
 * tgx_dynamic.hip — the dynamic scene path of tg's ray tracer: a scene with
 * fixed object and cluster capacities whose objects are created and
 * destroyed one at a time (tg_raytracer_create / create_object /
 * destroy_object, tgvk_raytracer.c:662-712, 805-868, 994-1067).
 *
 * As in tg, object and cluster indices come from LIFO free stacks that
 * start as [capacity - 1, ..., 0] (so the first pop is 0); a new object's
 * cluster pointers are appended after the live ones, and destroying an
 * object returns its cluster indices in pointer order, shifts every later
 * cluster pointer down over the gap and lowers the first pointer of every
 * later object. The scene arrays stay in HBM: only the new object's record
 * and cluster data cross PCIe, and the compaction runs on the device. The
 * host keeps the free stacks and a copy of the object records.
 */
#include "tgx_internal.h"

/* dst[i] = src[i] for i < n (the compaction's two halves go through the
 * context's scratch buffer, so source and destination never overlap) */
__global__ void dyn_copy_k(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

/* New clusters: pointer first + k -> index idx[k], index -> object, and the
 * k-th staged bit words and materials into the index's slots. */
__global__ void dyn_place_k(const uint32_t* __restrict__ idx, uint32_t n, uint32_t first, uint32_t object,
                            const uint32_t* __restrict__ stage_bits, const uint8_t* __restrict__ stage_mats,
                            uint32_t* __restrict__ cptrs, uint32_t* __restrict__ c2o, uint32_t* __restrict__ bits,
                            uint8_t* __restrict__ mats, uint8_t* __restrict__ dirty)
{
    const uint32_t k = blockIdx.x;
    if (k >= n) return;
    const uint32_t c = idx[k];
    if (threadIdx.x == 0) {
        cptrs[first + k] = c;
        c2o[c] = object;
        dirty[c] = 1u;
    }
    if (threadIdx.x < 16u) bits[(size_t)c * 16u + threadIdx.x] = stage_bits ? stage_bits[(size_t)k * 16u + threadIdx.x] : 0u;
    for (uint32_t v = threadIdx.x; v < 512u; v += blockDim.x)
        mats[(size_t)c * 512u + v] = stage_mats ? stage_mats[(size_t)k * 512u + v] : (uint8_t)0;
}

/* Lowers first_cluster_pointer of live objects at or past `end` by n. */
__global__ void dyn_shift_objects_k(tgx_object* __restrict__ objects, uint32_t n_objects, uint32_t end, uint32_t n)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_objects) return;
    tgx_object& ob = objects[o];
    const bool live = ob.n_cluster_pointers_per_dim[0] && ob.n_cluster_pointers_per_dim[1] && ob.n_cluster_pointers_per_dim[2];
    if (live && ob.first_cluster_pointer >= end) ob.first_cluster_pointer -= n;
}

static int sync_objects(tgx_context* ctx, uint32_t o)
{
    TGX_HIP_CHECK(ctx, hipMemcpyAsync((tgx_object*)ctx->objects.ptr + o, &ctx->dyn_objects[o], sizeof(tgx_object),
                                      hipMemcpyHostToDevice, ctx->stream));
    return TGX_OK;
}

extern "C" {

int tgx_scene_reserve(tgx_context* ctx, uint32_t object_capacity, uint32_t cluster_capacity, const tgx_lut_entry* luts,
                      uint32_t n_luts)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    ctx->scene_version++;
    if (object_capacity == 0 || cluster_capacity == 0 || !luts || n_luts == 0)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "capacities, LUTs and n_luts must be non-zero");
    if ((uint64_t)cluster_capacity >= (1ull << TGX_CLUSTER_POINTER_BITS))
        return tgx_fail(ctx, TGX_ERR_CAPACITY, "%u clusters exceed the 31-bit pointer field", cluster_capacity);
    int rc;
    if ((rc = tgx_scene_replace_begin(ctx, "tgx_scene_reserve"))) return rc;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    ctx->has_scene = false;
    ctx->has_materials = false;
    ctx->dynamic = false;
    ctx->svo_prev_valid = false;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->objects, sizeof(tgx_object) * (size_t)object_capacity, "objects")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->cluster_pointers, 4ull * cluster_capacity, "cluster pointers")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->cluster_idx_to_object, 4ull * cluster_capacity, "cluster->object")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->bits, 64ull * cluster_capacity, "solid bits")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->materials, 512ull * cluster_capacity, "materials")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->luts, sizeof(tgx_lut_entry) * TGX_LUT_ENTRIES * (size_t)n_luts, "luts")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->dyn_tmp, 4ull * cluster_capacity, "compaction scratch")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->cl_dirty, 4ull * ((cluster_capacity + 3ull) / 4ull), "cluster dirty flags")))
        return rc;
    ctx->dyn_objects.assign(object_capacity, tgx_object{});
    ctx->dyn_free_objects.resize(object_capacity);
    ctx->dyn_free_clusters.resize(cluster_capacity);
    for (uint32_t i = 0; i < object_capacity; i++) ctx->dyn_free_objects[i] = object_capacity - i - 1u;
    for (uint32_t i = 0; i < cluster_capacity; i++) ctx->dyn_free_clusters[i] = cluster_capacity - i - 1u;
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->objects.ptr, 0, sizeof(tgx_object) * (size_t)object_capacity, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->cluster_idx_to_object.ptr, 0, 4ull * cluster_capacity, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->bits.ptr, 0, 64ull * cluster_capacity, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->materials.ptr, 0, 512ull * cluster_capacity, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->cl_dirty.ptr, 0, 4ull * ((cluster_capacity + 3ull) / 4ull), ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->luts.ptr, luts, sizeof(tgx_lut_entry) * TGX_LUT_ENTRIES * (size_t)n_luts,
                                      hipMemcpyHostToDevice, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_objects = object_capacity;
    ctx->n_clusters = cluster_capacity;
    ctx->n_cluster_pointers = 0;
    ctx->n_luts = n_luts;
    ctx->has_materials = true;
    ctx->has_scene = true;
    ctx->dynamic = true;
    return TGX_OK;
}

int tgx_object_create(tgx_context* ctx, const uint32_t n_clusters_per_dim[3], const float translation[3],
                      const float rotation[16], uint32_t lut_idx, const uint32_t* bits, const uint8_t* materials,
                      uint32_t* out_object_idx)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!ctx->dynamic) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no dynamic scene: call tgx_scene_reserve first");
    if (ctx->n_graphs > 0) return tgx_fail(ctx, TGX_ERR_CAPACITY, "a frame graph holds the scene: destroy it first");
    ctx->scene_version++;
    if (!n_clusters_per_dim || !translation || !rotation || !out_object_idx)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "null argument");
    const uint64_t n64 = (uint64_t)n_clusters_per_dim[0] * n_clusters_per_dim[1] * n_clusters_per_dim[2];
    if (n64 == 0) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "an object needs at least one cluster per axis");
    if (lut_idx >= ctx->n_luts) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "lut_idx %u >= %u", lut_idx, ctx->n_luts);
    if (ctx->dyn_free_objects.empty()) return tgx_fail(ctx, TGX_ERR_CAPACITY, "no free object index");
    if (n64 > ctx->dyn_free_clusters.size() || ctx->n_cluster_pointers + n64 > ctx->n_clusters)
        return tgx_fail(ctx, TGX_ERR_CAPACITY, "%llu clusters requested, %zu free", (unsigned long long)n64,
                        ctx->dyn_free_clusters.size());
    const uint32_t n = (uint32_t)n64;
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->dyn_stage, (4ull + 64ull + 512ull) * n, "object staging"))) return rc;
    /* pops, as tg_raytracer_create_object (:816, :855) */
    const uint32_t o = ctx->dyn_free_objects.back();
    ctx->dyn_free_objects.pop_back();
    std::vector<uint32_t> idx(n);
    for (uint32_t k = 0; k < n; k++) {
        idx[k] = ctx->dyn_free_clusters.back();
        ctx->dyn_free_clusters.pop_back();
    }
    tgx_object& ob = ctx->dyn_objects[o];
    for (int a = 0; a < 3; a++) {
        ob.n_cluster_pointers_per_dim[a] = n_clusters_per_dim[a];
        ob.translation[a] = translation[a];
    }
    ob.first_cluster_pointer = ctx->n_cluster_pointers;
    ob.lut_idx = lut_idx;
    for (int i = 0; i < 16; i++) ob.rotation[i] = rotation[i];
    uint32_t* d_idx = (uint32_t*)ctx->dyn_stage.ptr;
    uint32_t* d_bits = d_idx + n;
    uint8_t* d_mats = (uint8_t*)(d_bits + 16ull * n);
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(d_idx, idx.data(), 4ull * n, hipMemcpyHostToDevice, ctx->stream));
    if (bits) TGX_HIP_CHECK(ctx, hipMemcpyAsync(d_bits, bits, 64ull * n, hipMemcpyHostToDevice, ctx->stream));
    if (materials) TGX_HIP_CHECK(ctx, hipMemcpyAsync(d_mats, materials, 512ull * n, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = sync_objects(ctx, o))) return rc;
    hipLaunchKernelGGL(dyn_place_k, dim3(n), dim3(128), 0, ctx->stream, d_idx, n, ob.first_cluster_pointer, o,
                       bits ? d_bits : nullptr, materials ? d_mats : nullptr, (uint32_t*)ctx->cluster_pointers.ptr,
                       (uint32_t*)ctx->cluster_idx_to_object.ptr, (uint32_t*)ctx->bits.ptr, (uint8_t*)ctx->materials.ptr,
                       (uint8_t*)ctx->cl_dirty.ptr);
    TGX_LAUNCH_CHECK(ctx, "dyn_place_k");
    /* the host arrays and the staged copies must outlive the copies */
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_cluster_pointers += n;
    ctx->svo_stale = true;
    *out_object_idx = o;
    return TGX_OK;
}

int tgx_object_destroy(tgx_context* ctx, uint32_t o)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!ctx->dynamic) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no dynamic scene: call tgx_scene_reserve first");
    if (ctx->n_graphs > 0) return tgx_fail(ctx, TGX_ERR_CAPACITY, "a frame graph holds the scene: destroy it first");
    if (o >= ctx->dyn_objects.size()) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "object %u out of range", o);
    /* cluster pointers and objects move: the visibility pass's
     * occupied-pointer lists are rebuilt */
    ctx->scene_version++;
    tgx_object& ob = ctx->dyn_objects[o];
    const uint32_t n = ob.n_cluster_pointers_per_dim[0] * ob.n_cluster_pointers_per_dim[1] * ob.n_cluster_pointers_per_dim[2];
    if (n == 0) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "object %u is not alive", o);
    const uint32_t first = ob.first_cluster_pointer, end = first + n;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    /* return the cluster indices in pointer order (:1010-1015) */
    std::vector<uint32_t> freed(n);
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(freed.data(), (const uint32_t*)ctx->cluster_pointers.ptr + first, 4ull * n,
                                      hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t k = 0; k < n; k++) ctx->dyn_free_clusters.push_back(freed[k]);
    /* compact the pointers and lower later objects' first pointers (:1019-1055) */
    const uint32_t tail = ctx->n_cluster_pointers - end;
    if (tail) {
        uint32_t* cp = (uint32_t*)ctx->cluster_pointers.ptr;
        uint32_t* tmp = (uint32_t*)ctx->dyn_tmp.ptr;
        hipLaunchKernelGGL(dyn_copy_k, dim3((tail + 255) / 256), dim3(256), 0, ctx->stream, cp + end, tmp, tail);
        TGX_LAUNCH_CHECK(ctx, "dyn_copy_k");
        hipLaunchKernelGGL(dyn_copy_k, dim3((tail + 255) / 256), dim3(256), 0, ctx->stream, tmp, cp + first, tail);
        TGX_LAUNCH_CHECK(ctx, "dyn_copy_k");
        const uint32_t n_obj = (uint32_t)ctx->dyn_objects.size();
        hipLaunchKernelGGL(dyn_shift_objects_k, dim3((n_obj + 255) / 256), dim3(256), 0, ctx->stream,
                           (tgx_object*)ctx->objects.ptr, n_obj, end, n);
        TGX_LAUNCH_CHECK(ctx, "dyn_shift_objects_k");
        for (tgx_object& other : ctx->dyn_objects) {
            const bool live = other.n_cluster_pointers_per_dim[0] && other.n_cluster_pointers_per_dim[1] &&
                              other.n_cluster_pointers_per_dim[2];
            if (live && other.first_cluster_pointer >= end) other.first_cluster_pointer -= n;
        }
    }
    ctx->n_cluster_pointers -= n;
    /* return the object index; its record is cleared (not initialised) */
    ob = tgx_object{};
    int rc;
    if ((rc = sync_objects(ctx, o))) return rc;
    ctx->dyn_free_objects.push_back(o);
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->svo_stale = true;
    return TGX_OK;
}

int tgx_lut_set(tgx_context* ctx, uint32_t lut_idx, uint32_t entry, const tgx_lut_entry* value)
{
    if (!ctx || !value) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene");
    if (lut_idx >= ctx->n_luts || entry >= TGX_LUT_ENTRIES)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "LUT %u entry %u outside %u LUTs of %u entries", lut_idx, entry,
                        ctx->n_luts, TGX_LUT_ENTRIES);
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    /* stream-ordered: frames queued before see the old entry, later ones
     * the new; the copy waits for its host source */
    TGX_HIP_CHECK(ctx, hipMemcpyAsync((tgx_lut_entry*)ctx->luts.ptr + (size_t)lut_idx * TGX_LUT_ENTRIES + entry, value,
                                      sizeof(tgx_lut_entry), hipMemcpyHostToDevice, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_scene_get_counts(tgx_context* ctx, tgx_scene_counts* out)
{
    if (!ctx || !out) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene");
    out->n_objects = ctx->n_objects;
    out->n_cluster_pointers = ctx->n_cluster_pointers;
    out->n_clusters = ctx->n_clusters;
    out->n_luts = ctx->n_luts;
    out->has_materials = ctx->has_materials ? 1u : 0u;
    return TGX_OK;
}

} /* extern "C" */
