/*
 * tgx_gen_dev.hip — synthetic scenes generated in the context's device
 * buffers (the GPU voxel generation tg leaves as a TODO at
 * tgvk_raytracer.c:868). Object records, LUTs and each object's noise
 * offset come from the host generator; the cluster maps and every voxel's
 * solid bit and material are computed on the GPU with the shared value
 * noise (tgx_noise.h), so the bytes equal tgx_gen_scene_fill's.
 */
#include "tgx_internal.h"
#include "tgx_gen_int.h"
#include "tgx_noise.h"

struct gen_obj {
    uint32_t first, nx, ny, nz;
    double   ox, oy, oz, pad;
};
static_assert(sizeof(gen_obj) == 48, "gen_obj layout");

/* cluster pointer c -> index c; index -> the object whose range holds it */
__global__ void gen_maps_k(const gen_obj* __restrict__ objs, uint32_t n_objects, uint32_t n_clusters,
                           uint32_t* __restrict__ cptrs, uint32_t* __restrict__ c2o)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_clusters) return;
    cptrs[c] = c;
    uint32_t lo = 0, hi = n_objects;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (objs[mid].first <= c) lo = mid; else hi = mid;
    }
    c2o[c] = lo;
}

/* One lane per voxel; a wave covers 64 consecutive voxels of one cluster,
 * i.e. two solid-bit words, written from the wave's ballot. */
__global__ __launch_bounds__(256) void gen_voxels_k(const gen_obj* __restrict__ objs, const uint32_t* __restrict__ c2o,
                                                    uint64_t n_voxels, uint64_t key_solid, uint64_t key_mat,
                                                    uint32_t* __restrict__ bits, uint8_t* __restrict__ materials)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_voxels; g += stride) {
        /* n_voxels is a multiple of 512 and g - lane a multiple of 64: the
         * whole wave takes this iteration together */
        const uint32_t c = (uint32_t)(g >> 9), v = (uint32_t)(g & 511u);
        const gen_obj o = objs[c2o[c]];
        const uint32_t rc = c - o.first;
        const uint32_t cx = rc % o.nx, cy = (rc / o.nx) % o.ny, cz = rc / (o.nx * o.ny);
        int solid;
        uint8_t m;
        tgx_noise_voxel(key_solid, key_mat, o.ox, o.oy, o.oz, cx, cy, cz, v, &solid, &m);
        materials[g] = m;
        const unsigned long long ball = __ballot(solid);
        if ((threadIdx.x & 63u) == 0u) {
            bits[g >> 5] = (uint32_t)ball;
            bits[(g >> 5) + 1] = (uint32_t)(ball >> 32);
        }
    }
}

extern "C" {

int tgx_gen_scene_device(tgx_context* ctx, const tgx_gen_params* p)
{
    if (!ctx || !p) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    ctx->scene_version++;
    tgx_gen_sizes sz;
    int rc = tgx_gen_scene_sizes(p, &sz);
    if (rc) return tgx_fail(ctx, rc, "bad generator parameters or more than 2^31 clusters");
    if (p->kind != TGX_GEN_NOISE_OBJECTS) {
        /* the 64-cluster Bernoulli scene is drawn from one sequential stream: host */
        std::vector<tgx_object> objs(sz.n_objects);
        std::vector<uint32_t> cptrs(sz.n_cluster_pointers), c2o(sz.n_clusters), bits(16ull * sz.n_clusters);
        std::vector<uint8_t> mats(512ull * sz.n_clusters);
        std::vector<tgx_lut_entry> luts(256ull * sz.n_luts);
        if ((rc = tgx_gen_scene_fill(p, objs.data(), cptrs.data(), c2o.data(), bits.data(), mats.data(), luts.data())))
            return tgx_fail(ctx, rc, "host generator failed");
        tgx_scene_desc d = {objs.data(), sz.n_objects, cptrs.data(), sz.n_cluster_pointers, c2o.data(),
                            sz.n_clusters, bits.data(), mats.data(), luts.data(), sz.n_luts};
        return tgx_scene_upload(ctx, &d);
    }
    std::vector<tgx_object> objs(sz.n_objects);
    std::vector<double> off(3ull * sz.n_objects);
    std::vector<tgx_lut_entry> luts(256ull * sz.n_luts);
    if ((rc = tgx_gen_noise_objects(p, objs.data(), off.data()))) return tgx_fail(ctx, rc, "object generation failed");
    tgx_gen_luts(p, luts.data());
    std::vector<gen_obj> go(sz.n_objects);
    for (uint32_t o = 0; o < sz.n_objects; o++)
        go[o] = {objs[o].first_cluster_pointer, objs[o].n_cluster_pointers_per_dim[0],
                 objs[o].n_cluster_pointers_per_dim[1], objs[o].n_cluster_pointers_per_dim[2], off[3ull * o],
                 off[3ull * o + 1], off[3ull * o + 2], 0.0};
    if ((rc = tgx_scene_replace_begin(ctx, "tgx_gen_scene_device"))) return rc;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    ctx->has_scene = false;
    ctx->has_materials = false;
    ctx->dynamic = false;
    ctx->svo_prev_valid = false;
    ctx->dyn_objects.clear();
    ctx->dyn_free_objects.clear();
    ctx->dyn_free_clusters.clear();
    if ((rc = tgx_buffer_reserve(ctx, &ctx->objects, sizeof(tgx_object) * sz.n_objects, "objects")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->cluster_pointers, 4ull * sz.n_cluster_pointers, "cluster pointers")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->cluster_idx_to_object, 4ull * sz.n_clusters, "cluster->object")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->bits, 64ull * sz.n_clusters, "solid bits")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->materials, 512ull * sz.n_clusters, "materials")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->luts, sizeof(tgx_lut_entry) * luts.size(), "luts")))
        return rc;
    gen_obj* d_obj = nullptr;
    TGX_HIP_CHECK(ctx, hipMalloc(&d_obj, sizeof(gen_obj) * go.size()));
    hipError_t e = hipMemcpyAsync(d_obj, go.data(), sizeof(gen_obj) * go.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->objects.ptr, objs.data(), sizeof(tgx_object) * objs.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->luts.ptr, luts.data(), sizeof(tgx_lut_entry) * luts.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gen_maps_k, dim3((sz.n_clusters + 255) / 256), dim3(256), 0, ctx->stream, d_obj,
                           sz.n_objects, sz.n_clusters, (uint32_t*)ctx->cluster_pointers.ptr,
                           (uint32_t*)ctx->cluster_idx_to_object.ptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        tgx_launch_scope s(ctx, "gen_voxels_k");
        const uint64_t n_voxels = 512ull * sz.n_clusters;
        const uint64_t want = (n_voxels + 255) / 256;
        hipLaunchKernelGGL(gen_voxels_k, dim3((uint32_t)(want < 1048576u ? want : 1048576u)), dim3(256), 0, ctx->stream,
                           d_obj, (const uint32_t*)ctx->cluster_idx_to_object.ptr, n_voxels,
                           tgx_gen_noise_key(p->seed, 0), tgx_gen_noise_key(p->seed, 1),
                           (uint32_t*)ctx->bits.ptr, (uint8_t*)ctx->materials.ptr);
        e = hipGetLastError();
    }
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_obj);
    if (e != hipSuccess || es != hipSuccess)
        return tgx_fail(ctx, TGX_ERR_HIP, "scene generation failed: %s", hipGetErrorString(e != hipSuccess ? e : es));
    ctx->n_objects = sz.n_objects;
    ctx->n_cluster_pointers = sz.n_cluster_pointers;
    ctx->n_clusters = sz.n_clusters;
    ctx->n_luts = sz.n_luts;
    ctx->has_materials = true;
    ctx->has_scene = true;
    return tgx_vis_solid_refresh(ctx);
}

int tgx_scene_export(tgx_context* ctx, tgx_object* objects, uint32_t* cluster_pointers, uint32_t* cluster_idx_to_object,
                     uint32_t* bits, uint8_t* materials, tgx_lut_entry* luts)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene");
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (objects) TGX_HIP_CHECK(ctx, hipMemcpy(objects, ctx->objects.ptr, sizeof(tgx_object) * ctx->n_objects, hipMemcpyDeviceToHost));
    if (cluster_pointers)
        TGX_HIP_CHECK(ctx, hipMemcpy(cluster_pointers, ctx->cluster_pointers.ptr, 4ull * ctx->n_cluster_pointers, hipMemcpyDeviceToHost));
    if (cluster_idx_to_object)
        TGX_HIP_CHECK(ctx, hipMemcpy(cluster_idx_to_object, ctx->cluster_idx_to_object.ptr, 4ull * ctx->n_clusters, hipMemcpyDeviceToHost));
    if (bits) TGX_HIP_CHECK(ctx, hipMemcpy(bits, ctx->bits.ptr, 64ull * ctx->n_clusters, hipMemcpyDeviceToHost));
    if (materials && ctx->has_materials)
        TGX_HIP_CHECK(ctx, hipMemcpy(materials, ctx->materials.ptr, 512ull * ctx->n_clusters, hipMemcpyDeviceToHost));
    else if (materials)
        memset(materials, 0, 512ull * ctx->n_clusters);
    if (luts && ctx->luts.ptr)
        TGX_HIP_CHECK(ctx, hipMemcpy(luts, ctx->luts.ptr, sizeof(tgx_lut_entry) * TGX_LUT_ENTRIES * ctx->n_luts, hipMemcpyDeviceToHost));
    return TGX_OK;
}

} /* extern "C" */
