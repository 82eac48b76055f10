/*
 * tgx_oracle_shade.c — TEST INFRASTRUCTURE ONLY: scalar C restatement of
 * tg's shading pass (shading.frag:114-337), the checker of tgx_shade.hip.
 *
 * Per pixel: decode of the visibility word (:116-120), cluster pointer ->
 * cluster -> object (:124-126), material byte and LUT entry (:128-135;
 * fp32 albedo in place of the reference's RGBA8), the hit voxel's face
 * normal from a cluster-space slab test (:141-229), the world hit from the
 * quantised depth (:231), then one of the debug views (:233-290) or the
 * GGX lighting of tg_shade (:82-109, :291-322). fp32 in GLSL's operation
 * order (GLSL mix(x, y, a) = x (1 - a) + y a, dot left to right,
 * normalize(v) = v / sqrt(dot(v, v))); pow(x, 5.0), whose precision GLSL
 * leaves open, is ((x x)(x x)) x.
 */
#include "tgx_oracle.h"

#include <math.h>
#include <string.h>

#define PI_F 3.14159265358979323846f

typedef struct { float x, y, z; } sv3;

static float smin(float a, float b) { return a < b ? a : b; }   /* tg_math.h:17-23 */
static float smax(float a, float b) { return a > b ? a : b; }
static float sclamp01(float a) { return smin(smax(a, 0.0f), 1.0f); }
static float ssign(float a) { return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f); }
static float sdot(sv3 a, sv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static sv3 snormalize(sv3 a)
{
    const float m = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    sv3 r = {a.x / m, a.y / m, a.z / m};
    return r;
}
/* column-major m4 (tg_math.h:240-275) times (x, y, z, w), xyz of the result */
static sv3 smulv4(const float m[16], float x, float y, float z, float w)
{
    sv3 r;
    r.x = x * m[0] + y * m[4] + z * m[8] + w * m[12];
    r.y = x * m[1] + y * m[5] + z * m[9] + w * m[13];
    r.z = x * m[2] + y * m[6] + z * m[10] + w * m[14];
    return r;
}

/* tg_hash_u32 (util.inc:47-56) */
static uint32_t shash(uint32_t v)
{
    uint32_t r = v;
    r ^= r >> 16;
    r *= 0x85ebca6bu;
    r ^= r >> 13;
    r *= 0xc2b2ae35u;
    r ^= r >> 16;
    return r;
}

static void hash_color(uint32_t v, float out[4])
{
    const uint32_t h0 = shash(v), h1 = shash(h0), h2 = shash(h1);
    out[0] = (float)h0 / 4294967295.0f;
    out[1] = (float)h1 / 4294967295.0f;
    out[2] = (float)h2 / 4294967295.0f;
    out[3] = 1.0f;
}

/* tg_distribution_ggx (shading.frag:50-58) */
static float distribution_ggx(float clamped_n_dot_h, float roughness)
{
    const float a = roughness * roughness;
    const float a_sqr = a * a;
    float denom = (clamped_n_dot_h * clamped_n_dot_h * (a_sqr - 1.0f) + 1.0f);
    denom = PI_F * denom * denom;
    return a_sqr / denom;
}

/* tg_geometry_schlick_ggf (:60-67) */
static float geometry_schlick_ggf(float n_dot_v, float roughness)
{
    const float r = (roughness + 1.0f);
    const float k = (r * r) / 8.0f;
    const float denom = n_dot_v * (1.0f - k) + k;
    return n_dot_v / denom;
}

/* tg_geometry_smith (:69-75) */
static float geometry_smith(float clamped_n_dot_v, float clamped_n_dot_l, float roughness)
{
    const float ggx2 = geometry_schlick_ggf(clamped_n_dot_v, roughness);
    const float ggx1 = geometry_schlick_ggf(clamped_n_dot_l, roughness);
    return ggx1 * ggx2;
}

/* tg_fresnel_schlick (:77-81) */
static float fresnel_schlick(float u, float roughness)
{
    const float f_lambda = 1.0f - roughness;
    const float x = 1.0f - u;
    const float x2 = x * x;
    return f_lambda + (1.0f - f_lambda) * (x2 * x2 * x);
}

/* tg_shade (:83-109) with radiance vec3(radiance) */
static sv3 shade(sv3 n, sv3 v, sv3 l, sv3 diffuse_albedo, sv3 specular_albedo, float metallic, float roughness,
                 float radiance)
{
    sv3 h = {v.x + l.x, v.y + l.y, v.z + l.z};
    h = snormalize(h);
    const float h_dot_n = sdot(h, n);
    const float l_dot_n = sdot(n, l);
    const float n_dot_v = sdot(n, v);
    const float clamped_h_dot_n = sclamp01(h_dot_n);
    const float clamped_l_dot_n = sclamp01(l_dot_n);
    const float clamped_n_dot_v = sclamp01(n_dot_v);
    const float d = distribution_ggx(clamped_h_dot_n, roughness);
    const float f = fresnel_schlick(clamped_n_dot_v, roughness);
    const float g = geometry_smith(clamped_n_dot_v, clamped_l_dot_n, roughness);
    const float dfg = d * f * g;
    const float denominator = 4.0f * clamped_n_dot_v * clamped_l_dot_n;
    const float s = dfg / smax(denominator, 0.001f);
    const sv3 specular = {specular_albedo.x * s, specular_albedo.y * s, specular_albedo.z * s};
    const float kd = (1.0f - f) * (1.0f - metallic);
    const sv3 diffuse = {kd * diffuse_albedo.x / PI_F, kd * diffuse_albedo.y / PI_F, kd * diffuse_albedo.z / PI_F};
    sv3 r;
    r.x = (diffuse.x + specular.x) * radiance * clamped_l_dot_n;
    r.y = (diffuse.y + specular.y) * radiance * clamped_l_dot_n;
    r.z = (diffuse.z + specular.z) * radiance * clamped_l_dot_n;
    return r;
}

void tgxo_shade_pixel(const tgxo_scene* sc, const tgx_camera* cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                      uint64_t word, const tgx_shade_params* p, float out[4])
{
    const float depth_24b = (float)(word >> 40) / 16777215.0f;
    if (!(depth_24b < 1.0f)) {
        memcpy(out, p->miss, 4 * sizeof(float));
        return;
    }
    const uint32_t cluster_pointer = (uint32_t)(word >> 9) & 2147483647u;
    const uint32_t voxel_idx = (uint32_t)word & 511u;
    const uint32_t cluster_idx = sc->cluster_pointers[cluster_pointer];
    const uint32_t object_idx = sc->cluster_idx_to_object[cluster_idx];
    const tgx_object* o = &sc->objects[object_idx];
    const uint32_t color_lut_idx = sc->materials ? sc->materials[(size_t)cluster_idx * 512u + voxel_idx] : 0u;
    const tgx_lut_entry* e = &sc->luts[(size_t)o->lut_idx * TGX_LUT_ENTRIES + color_lut_idx];

    /* pixel ray (gl_FragCoord is the pixel centre) */
    const float fx = ((float)px + 0.5f) / (float)w;
    const float fy = 1.0f - ((float)py + 0.5f) / (float)h;
    float ws2ms[16];
    tgxo_world_to_cluster(o, cluster_pointer - o->first_cluster_pointer, ws2ms);
    const sv3 origin_ws = {cam->camera[0], cam->camera[1], cam->camera[2]};
    const sv3 origin_ms = smulv4(ws2ms, origin_ws.x, origin_ws.y, origin_ws.z, 1.0f);
    /* tg_lerp_corner_ray_directions_nn (common.inc:40-46) */
    sv3 dir_ws;
    {
        const float omy = 1.0f - fy, omx = 1.0f - fx;
        const sv3 left = {cam->ray_bl[0] * omy + cam->ray_tl[0] * fy, cam->ray_bl[1] * omy + cam->ray_tl[1] * fy,
                          cam->ray_bl[2] * omy + cam->ray_tl[2] * fy};
        const sv3 right = {cam->ray_br[0] * omy + cam->ray_tr[0] * fy, cam->ray_br[1] * omy + cam->ray_tr[1] * fy,
                           cam->ray_br[2] * omy + cam->ray_tr[2] * fy};
        dir_ws.x = left.x * omx + right.x * fx;
        dir_ws.y = left.y * omx + right.y * fx;
        dir_ws.z = left.z * omx + right.z * fx;
    }
    const sv3 dir_ms = snormalize(smulv4(ws2ms, dir_ws.x, dir_ws.y, dir_ws.z, 0.0f));

    /* face normal of the hit voxel (:183-229); zero when the slab test fails */
    const float vmin[3] = {(float)(voxel_idx % 8u), (float)((voxel_idx / 8u) % 8u), (float)(voxel_idx / 64u)};
    const float vmax[3] = {vmin[0] + 1.0f, vmin[1] + 1.0f, vmin[2] + 1.0f};
    const float o3[3] = {origin_ms.x, origin_ms.y, origin_ms.z}, d3[3] = {dir_ms.x, dir_ms.y, dir_ms.z};
    sv3 normal_ws = {0.0f, 0.0f, 0.0f};
    float enter, exit_;
    if (tgxo_intersect_ray_aabb(o3, d3, vmin, vmax, &enter, &exit_)) {
        sv3 hit = origin_ms;
        if (enter > 0.0f) {
            hit.x = origin_ms.x + enter * dir_ms.x;
            hit.y = origin_ms.y + enter * dir_ms.y;
            hit.z = origin_ms.z + enter * dir_ms.z;
        }
        sv3 n = {hit.x - (vmin[0] + 0.5f), hit.y - (vmin[1] + 0.5f), hit.z - (vmin[2] + 0.5f)};
        if (fabsf(n.x) > fabsf(n.y)) {
            n.y = 0.0f;
            if (fabsf(n.x) > fabsf(n.z)) { n.x = ssign(n.x); n.z = 0.0f; }
            else { n.z = ssign(n.z); n.x = 0.0f; }
        } else {
            n.x = 0.0f;
            if (fabsf(n.y) > fabsf(n.z)) { n.y = ssign(n.y); n.z = 0.0f; }
            else { n.z = ssign(n.z); n.y = 0.0f; }
        }
        normal_ws = snormalize(smulv4(o->rotation, n.x, n.y, n.z, 0.0f));
    }
    const float t = depth_24b * cam->far_plane;
    const sv3 hit_ws = {origin_ws.x + t * dir_ws.x, origin_ws.y + t * dir_ws.y, origin_ws.z + t * dir_ws.z};

    switch (p->view) {
    case TGX_SHADE_OBJECT_INDEX: hash_color(object_idx, out); return;
    case TGX_SHADE_DEPTH: {
        const float g = smin(1.0f, 8.0f * depth_24b);
        out[0] = g; out[1] = g; out[2] = g; out[3] = 1.0f;
        return;
    }
    case TGX_SHADE_CLUSTER_INDEX:
    case TGX_SHADE_BLOCKS: hash_color(cluster_idx, out); return;
    case TGX_SHADE_VOXEL_INDEX: hash_color(voxel_idx, out); return;
    case TGX_SHADE_COLOR_LUT_INDEX: hash_color(color_lut_idx, out); return;
    case TGX_SHADE_COLOR:
        out[0] = e->albedo[0]; out[1] = e->albedo[1]; out[2] = e->albedo[2]; out[3] = 1.0f;
        return;
    case TGX_SHADE_NORMAL:
        out[0] = normal_ws.x * 0.5f + 0.5f; out[1] = normal_ws.y * 0.5f + 0.5f;
        out[2] = normal_ws.z * 0.5f + 0.5f; out[3] = 1.0f;
        return;
    default: break;
    }
    const float metallic = p->metallic;
    const sv3 v = snormalize((sv3){origin_ws.x - hit_ws.x, origin_ws.y - hit_ws.y, origin_ws.z - hit_ws.z});
    const sv3 l = snormalize((sv3){p->light_dir[0], p->light_dir[1], p->light_dir[2]});
    const sv3 albedo = p->view == TGX_SHADE_SHADING ? (sv3){1.0f, 1.0f, 1.0f}
                                                    : (sv3){e->albedo[0], e->albedo[1], e->albedo[2]};
    /* mix(vec3(0.04), albedo, metallic) */
    const sv3 specular_albedo = {0.04f * (1.0f - metallic) + albedo.x * metallic,
                                 0.04f * (1.0f - metallic) + albedo.y * metallic,
                                 0.04f * (1.0f - metallic) + albedo.z * metallic};
    const float roughness = p->roughness < 0.0f ? e->roughness : p->roughness;
    const sv3 lo = shade(normal_ws, v, l, albedo, specular_albedo, metallic, roughness, p->light_radiance);
    out[0] = p->ambient * albedo.x + lo.x;
    out[1] = p->ambient * albedo.y + lo.y;
    out[2] = p->ambient * albedo.z + lo.z;
    out[3] = 1.0f;
}

void tgxo_shade(const tgxo_scene* sc, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis,
                const tgx_shade_params* p, float* rgba)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t y = 0; y < (int64_t)h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const size_t i = (size_t)y * w + x;
            tgxo_shade_pixel(sc, cam, w, h, x, (uint32_t)y, vis[i], p, &rgba[4 * i]);
        }
}
