/*
 * tgx_shade.hip — the shading pass over the visibility buffer
 * (shading.frag:114-337): decode of the 64-bit word, material and LUT
 * lookup, the hit voxel's face normal, the reference's GGX lighting with an
 * ambient term, and its debug views. One lane per pixel writes one RGBA
 * float4; the pass is a streaming read of the visibility buffer plus
 * gathers of the per-cluster maps, object records and materials.
 *
 * The pick readback (tgx_pick) decodes single pixels the same way.
 *
 * Arithmetic follows shading.frag's operation order in fp32 (no contraction),
 * so the oracle's transcription gives identical bits. GLSL leaves the
 * precision of pow() to the implementation; pow(x, 5.0) is evaluated here
 * (and in the oracle) as ((x x)(x x)) x.
 */
#include "tgx_internal.h"
#include "tgx_m4.h"

struct shade_cam {
    float pos[3];
    float bl[3], br[3], tr[3], tl[3];
    float far_plane;
};

struct shade_consts {
    uint32_t view;
    float    lx, ly, lz;  /* normalised light direction */
    float    radiance, ambient, metallic, roughness;
    float    miss[4];
};

#define SHADE_PI 3.14159265358979323846f

__device__ static inline float sh_min(float a, float b) { return a < b ? a : b; }
__device__ static inline float sh_max(float a, float b) { return a > b ? a : b; }
__device__ static inline float sh_clamp01(float a) { return sh_min(sh_max(a, 0.0f), 1.0f); }
__device__ static inline float sh_sign(float a) { return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f); }

/* tg_hash_u32 (util.inc:47-56) */
__device__ static inline uint32_t sh_hash(uint32_t v)
{
    v ^= v >> 16; v *= 0x85ebca6bu; v ^= v >> 13; v *= 0xc2b2ae35u; v ^= v >> 16;
    return v;
}

/* Three chained hashes as RGB in [0, 1] (shading.frag:235-242). */
__device__ static inline float4 sh_hash_color(uint32_t v)
{
    const uint32_t h0 = sh_hash(v), h1 = sh_hash(h0), h2 = sh_hash(h1);
    return make_float4((float)h0 / 4294967295.0f, (float)h1 / 4294967295.0f, (float)h2 / 4294967295.0f, 1.0f);
}

/* tg_shade (shading.frag:82-109) with white light of `radiance` per channel. */
__device__ static inline float3 sh_ggx(float3 n, float3 v, float3 l, float3 alb, float3 spec, float metallic,
                                       float roughness, float radiance)
{
    float hx = v.x + l.x, hy = v.y + l.y, hz = v.z + l.z;
    const float hm = sqrtf(hx * hx + hy * hy + hz * hz);
    hx = hx / hm; hy = hy / hm; hz = hz / hm;
    const float h_dot_n = hx * n.x + hy * n.y + hz * n.z;
    const float l_dot_n = n.x * l.x + n.y * l.y + n.z * l.z;
    const float n_dot_v = n.x * v.x + n.y * v.y + n.z * v.z;
    const float chn = sh_clamp01(h_dot_n), cln = sh_clamp01(l_dot_n), cnv = sh_clamp01(n_dot_v);
    /* tg_distribution_ggx */
    const float a = roughness * roughness;
    const float a_sqr = a * a;
    float denom = chn * chn * (a_sqr - 1.0f) + 1.0f;
    denom = SHADE_PI * denom * denom;
    const float d = a_sqr / denom;
    /* tg_fresnel_schlick */
    const float f_lambda = 1.0f - roughness;
    const float u = 1.0f - cnv;
    const float u2 = u * u;
    const float f = f_lambda + (1.0f - f_lambda) * (u2 * u2 * u);
    /* tg_geometry_smith */
    const float r = roughness + 1.0f;
    const float k = r * r / 8.0f;
    const float g2 = cnv / (cnv * (1.0f - k) + k);
    const float g1 = cln / (cln * (1.0f - k) + k);
    const float g = g1 * g2;
    const float dfg = d * f * g;
    const float den = 4.0f * cnv * cln;
    const float sf = dfg / sh_max(den, 0.001f);
    const float df = (1.0f - f) * (1.0f - metallic);
    float3 o;
    o.x = (df * alb.x / SHADE_PI + spec.x * sf) * radiance * cln;
    o.y = (df * alb.y / SHADE_PI + spec.y * sf) * radiance * cln;
    o.z = (df * alb.z / SHADE_PI + spec.z * sf) * radiance * cln;
    return o;
}

__global__ __launch_bounds__(256) void shade_k(const tgx_object* __restrict__ objects,
                                               const uint32_t* __restrict__ cptrs, const uint32_t* __restrict__ c2o,
                                               const uint8_t* __restrict__ materials,
                                               const tgx_lut_entry* __restrict__ luts, shade_cam cam, uint32_t W,
                                               uint32_t H, const uint64_t* __restrict__ vis, shade_consts k,
                                               float4* __restrict__ out)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= W * H) return;
    const uint32_t px = pix % W, py = pix / W;
    const uint64_t word = vis[pix];
    const float depth = (float)(word >> 40) / 16777215.0f;
    if (!(depth < 1.0f)) {
        out[pix] = make_float4(k.miss[0], k.miss[1], k.miss[2], k.miss[3]);
        return;
    }
    const uint32_t cptr = (uint32_t)(word >> 9) & 0x7FFFFFFFu;
    const uint32_t v = (uint32_t)word & 511u;
    const uint32_t cidx = cptrs[cptr];
    const uint32_t oidx = c2o[cidx];
    const tgx_object ob = objects[oidx];
    const uint32_t mat = materials ? (uint32_t)materials[(size_t)cidx * 512u + v] : 0u;
    const tgx_lut_entry e = luts[(size_t)ob.lut_idx * TGX_LUT_ENTRIES + mat];

    /* ray in cluster space and the hit voxel's face normal (shading.frag:141-229) */
    const float fx = ((float)px + 0.5f) / (float)W;
    const float fy = 1.0f - ((float)py + 0.5f) / (float)H;
    const uint32_t rel = cptr - ob.first_cluster_pointer;
    const uint32_t nx = ob.n_cluster_pointers_per_dim[0], ny = ob.n_cluster_pointers_per_dim[1];
    const uint32_t rx = rel % nx, ry = (rel / nx) % ny, rz = rel / (nx * ny);
    const tgx_m4 w2c = tgx_world_to_cluster(ob.rotation, ob.translation, ob.n_cluster_pointers_per_dim, rx, ry, rz);
    const tgx_v4 o4 = tgx_m4_mulv4(w2c, cam.pos[0], cam.pos[1], cam.pos[2], 1.0f);
    const float omy = 1.0f - fy, omx = 1.0f - fx;
    const float lx = cam.bl[0] * omy + cam.tl[0] * fy, ly = cam.bl[1] * omy + cam.tl[1] * fy, lz = cam.bl[2] * omy + cam.tl[2] * fy;
    const float qx = cam.br[0] * omy + cam.tr[0] * fy, qy = cam.br[1] * omy + cam.tr[1] * fy, qz = cam.br[2] * omy + cam.tr[2] * fy;
    const float dwx = lx * omx + qx * fx, dwy = ly * omx + qy * fx, dwz = lz * omx + qz * fx;
    const tgx_v4 d4 = tgx_m4_mulv4(w2c, dwx, dwy, dwz, 0.0f);
    const float dm = sqrtf(d4.x * d4.x + d4.y * d4.y + d4.z * d4.z);
    const float ddx = d4.x / dm, ddy = d4.y / dm, ddz = d4.z / dm;
    const float vmx = (float)(v % 8u), vmy = (float)((v / 8u) % 8u), vmz = (float)(v / 64u);
    /* tg_intersect_ray_aabb (collide.inc:3-24) with the voxel box */
    const float ax0 = ddx == 0.0f ? -3.402823466e+38f : (vmx - o4.x) / ddx;
    const float ay0 = ddy == 0.0f ? -3.402823466e+38f : (vmy - o4.y) / ddy;
    const float az0 = ddz == 0.0f ? -3.402823466e+38f : (vmz - o4.z) / ddz;
    const float bx0 = ddx == 0.0f ? 3.402823466e+38f : ((vmx + 1.0f) - o4.x) / ddx;
    const float by0 = ddy == 0.0f ? 3.402823466e+38f : ((vmy + 1.0f) - o4.y) / ddy;
    const float bz0 = ddz == 0.0f ? 3.402823466e+38f : ((vmz + 1.0f) - o4.z) / ddz;
    const float enter = sh_max(sh_max(sh_min(ax0, bx0), sh_min(ay0, by0)), sh_min(az0, bz0));
    const float exit_ = sh_min(sh_min(sh_max(ax0, bx0), sh_max(ay0, by0)), sh_max(az0, bz0));
    float3 n = make_float3(0.0f, 0.0f, 0.0f);
    if (exit_ > 0.0f && enter <= exit_) {
        float hx = o4.x, hy = o4.y, hz = o4.z;
        if (enter > 0.0f) { hx = o4.x + enter * ddx; hy = o4.y + enter * ddy; hz = o4.z + enter * ddz; }
        float mx = hx - (vmx + 0.5f), my = hy - (vmy + 0.5f), mz = hz - (vmz + 0.5f);
        if (fabsf(mx) > fabsf(my)) {
            my = 0.0f;
            if (fabsf(mx) > fabsf(mz)) { mx = sh_sign(mx); mz = 0.0f; } else { mz = sh_sign(mz); mx = 0.0f; }
        } else {
            mx = 0.0f;
            if (fabsf(my) > fabsf(mz)) { my = sh_sign(my); mz = 0.0f; } else { mz = sh_sign(mz); my = 0.0f; }
        }
        tgx_m4 rot;
        for (int i = 0; i < 16; i++) rot.m[i] = ob.rotation[i];
        const tgx_v4 n4 = tgx_m4_mulv4(rot, mx, my, mz, 0.0f);
        const float nm = sqrtf(n4.x * n4.x + n4.y * n4.y + n4.z * n4.z);
        n = make_float3(n4.x / nm, n4.y / nm, n4.z / nm);
    }

    float4 c;
    switch (k.view) {
    case TGX_SHADE_OBJECT_INDEX: c = sh_hash_color(oidx); break;
    case TGX_SHADE_DEPTH: {
        const float t = sh_min(1.0f, 8.0f * depth);
        c = make_float4(t, t, t, 1.0f);
        break;
    }
    case TGX_SHADE_CLUSTER_INDEX:
    case TGX_SHADE_BLOCKS: c = sh_hash_color(cidx); break;
    case TGX_SHADE_VOXEL_INDEX: c = sh_hash_color(v); break;
    case TGX_SHADE_COLOR_LUT_INDEX: c = sh_hash_color(mat); break;
    case TGX_SHADE_COLOR: c = make_float4(e.albedo[0], e.albedo[1], e.albedo[2], 1.0f); break;
    case TGX_SHADE_NORMAL: c = make_float4(n.x * 0.5f + 0.5f, n.y * 0.5f + 0.5f, n.z * 0.5f + 0.5f, 1.0f); break;
    default: {
        /* TGX_SHADE_LIT (LUT albedo) and TGX_SHADE_SHADING (white) */
        const float3 alb = k.view == TGX_SHADE_SHADING ? make_float3(1.0f, 1.0f, 1.0f)
                                                         : make_float3(e.albedo[0], e.albedo[1], e.albedo[2]);
        /* hit_position_ws = camera + depth far dir (shading.frag:231) */
        const float t = depth * cam.far_plane;
        const float hpx = cam.pos[0] + t * dwx, hpy = cam.pos[1] + t * dwy, hpz = cam.pos[2] + t * dwz;
        float vx = cam.pos[0] - hpx, vy = cam.pos[1] - hpy, vz = cam.pos[2] - hpz;
        const float vm = sqrtf(vx * vx + vy * vy + vz * vz);
        vx = vx / vm; vy = vy / vm; vz = vz / vm;
        const float om = 1.0f - k.metallic;
        const float3 spec = make_float3(0.04f * om + alb.x * k.metallic, 0.04f * om + alb.y * k.metallic,
                                        0.04f * om + alb.z * k.metallic);
        const float rough = k.roughness < 0.0f ? e.roughness : k.roughness;
        const float3 lo = sh_ggx(n, make_float3(vx, vy, vz), make_float3(k.lx, k.ly, k.lz), alb, spec, k.metallic,
                                 rough, k.radiance);
        c = make_float4(k.ambient * alb.x + lo.x, k.ambient * alb.y + lo.y, k.ambient * alb.z + lo.z, 1.0f);
        break;
    }
    }
    out[pix] = c;
}

/* Pick readback: one lane per requested pixel. */
__global__ void pick_k(const uint32_t* __restrict__ cptrs, const uint32_t* __restrict__ c2o, shade_cam cam,
                       uint32_t has_cam, uint32_t W, uint32_t H, const uint64_t* __restrict__ vis,
                       const uint32_t* __restrict__ xy, uint32_t n, tgx_pick_result* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t px = xy[2 * i], py = xy[2 * i + 1];
    const uint64_t word = vis[(size_t)py * W + px];
    tgx_pick_result r;
    r.depth = (float)(word >> 40) / 16777215.0f;
    r.position[0] = r.position[1] = r.position[2] = 0.0f;
    if (r.depth < 1.0f) {
        r.hit = 1u;
        r.cluster_pointer = (uint32_t)(word >> 9) & 0x7FFFFFFFu;
        r.voxel = (uint32_t)word & 511u;
        r.cluster_idx = cptrs[r.cluster_pointer];
        r.object_idx = c2o[r.cluster_idx];
        if (has_cam) {
            const float fx = ((float)px + 0.5f) / (float)W;
            const float fy = 1.0f - ((float)py + 0.5f) / (float)H;
            const float omy = 1.0f - fy, omx = 1.0f - fx;
            const float lx = cam.bl[0] * omy + cam.tl[0] * fy, ly = cam.bl[1] * omy + cam.tl[1] * fy,
                        lz = cam.bl[2] * omy + cam.tl[2] * fy;
            const float qx = cam.br[0] * omy + cam.tr[0] * fy, qy = cam.br[1] * omy + cam.tr[1] * fy,
                        qz = cam.br[2] * omy + cam.tr[2] * fy;
            const float t = r.depth * cam.far_plane;
            r.position[0] = cam.pos[0] + t * (lx * omx + qx * fx);
            r.position[1] = cam.pos[1] + t * (ly * omx + qy * fx);
            r.position[2] = cam.pos[2] + t * (lz * omx + qz * fx);
        }
    } else {
        r.hit = 0u;
        r.cluster_pointer = r.cluster_idx = r.object_idx = r.voxel = 0xFFFFFFFFu;
    }
    out[i] = r;
}

static void cam_of(const tgx_camera* cam, shade_cam& sc)
{
    for (int a = 0; a < 3; a++) {
        sc.pos[a] = cam->camera[a];
        sc.bl[a] = cam->ray_bl[a]; sc.br[a] = cam->ray_br[a];
        sc.tr[a] = cam->ray_tr[a]; sc.tl[a] = cam->ray_tl[a];
    }
    sc.far_plane = cam->far_plane;
}

extern "C" {

int tgx_pick(tgx_context* ctx, const tgx_camera* cam, const uint64_t* vis_dev, uint32_t w, uint32_t h,
             const uint32_t* xy, uint32_t n, tgx_pick_result* out_host)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (n == 0) return TGX_OK;
    if (!vis_dev || !xy || !out_host) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene uploaded");
    if (n > (1u << 20)) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "at most 2^20 picks per call");
    for (uint32_t i = 0; i < n; i++)
        if (xy[2 * i] >= w || xy[2 * i + 1] >= h)
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "pick %u at (%u, %u) is outside %ux%u", i, xy[2 * i],
                            xy[2 * i + 1], w, h);
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->pick_buf, (8ull + sizeof(tgx_pick_result)) * n, "picks"))) return rc;
    uint32_t* d_xy = (uint32_t*)ctx->pick_buf.ptr;
    tgx_pick_result* d_out = (tgx_pick_result*)(d_xy + 2ull * n);
    shade_cam sc = {};
    if (cam) cam_of(cam, sc);
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(d_xy, xy, 8ull * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(pick_k, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       (const uint32_t*)ctx->cluster_pointers.ptr, (const uint32_t*)ctx->cluster_idx_to_object.ptr, sc,
                       cam ? 1u : 0u, w, h, vis_dev, d_xy, n, d_out);
    TGX_LAUNCH_CHECK(ctx, "pick_k");
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(out_host, d_out, sizeof(tgx_pick_result) * n, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_shade_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis_dev,
                     const tgx_shade_params* prm, float* rgba_dev)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!cam || !vis_dev || !prm || !rgba_dev) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene uploaded");
    if (w == 0 || h == 0 || w > 65535 || h > 65535 || (uint64_t)w * h >= (1ull << 31))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "invalid resolution %ux%u", w, h);
    if (prm->view > TGX_SHADE_SHADING) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "unknown view %u", prm->view);
    /* every hit pixel reads its LUT entry (shading.frag:132-135) */
    if (!ctx->luts.ptr || ctx->n_luts == 0) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "the scene has no LUTs");
    shade_cam sc;
    cam_of(cam, sc);
    shade_consts k;
    k.view = prm->view;
    /* normalize(light_dir) as GLSL does: v / sqrt(dot(v, v)) */
    const float l0 = prm->light_dir[0], l1 = prm->light_dir[1], l2 = prm->light_dir[2];
    const float lm = sqrtf(l0 * l0 + l1 * l1 + l2 * l2);
    if (!(lm > 0.0f)) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "light direction must be non-zero");
    k.lx = l0 / lm; k.ly = l1 / lm; k.lz = l2 / lm;
    k.radiance = prm->light_radiance;
    k.ambient = prm->ambient;
    k.metallic = prm->metallic;
    k.roughness = prm->roughness;
    for (int i = 0; i < 4; i++) k.miss[i] = prm->miss[i];
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const uint32_t n = w * h;
    {
        tgx_launch_scope s(ctx, "shade_k");
        hipLaunchKernelGGL(shade_k, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                           (const tgx_object*)ctx->objects.ptr, (const uint32_t*)ctx->cluster_pointers.ptr,
                           (const uint32_t*)ctx->cluster_idx_to_object.ptr, tgx_scene_materials(ctx),
                           (const tgx_lut_entry*)ctx->luts.ptr, sc, w, h, vis_dev, k, (float4*)rgba_dev);
    }
    TGX_LAUNCH_CHECK(ctx, "shade_k");
    return TGX_OK;
}

} /* extern "C" */
