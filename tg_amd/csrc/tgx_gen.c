/*
 * tgx_gen.c — deterministic synthetic scenes and cameras (host only).
 * Declarations and stream layout: tgx_gen.h.
 */
#include "tgx_gen.h"
#include "tgx_gen_int.h"
#include "tgx_m4.h"
#include "tgx_noise.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TGX_PI_D 3.14159265358979323846

enum {
    STREAM_EXTENTS   = 1,
    STREAM_TRANSFORM = 2,
    STREAM_SOLID     = 3,
    STREAM_MATERIAL  = 4,
    STREAM_LUT       = 5,
    STREAM_CAMERA    = 6,
    STREAM_NOISE     = 7,
    STREAM_LUT_IDX   = 8
};

#define splitmix64_next tgx_splitmix64_next

uint64_t tgx_gen_stream_seed(uint64_t seed, uint64_t purpose)
{
    uint64_t s = seed ^ (purpose * 0xD1B54A32D192ED03ull);
    (void)splitmix64_next(&s);
    return splitmix64_next(&s);
}

/* Uniform double in [0, 1) with 53 random bits. */
static double rng_unit(uint64_t* s) { return (double)(splitmix64_next(s) >> 11) * (1.0 / 9007199254740992.0); }
/* Uniform float in [0, 1) with 24 random bits. */
static float rng_unit_f(uint64_t* s) { return (float)(splitmix64_next(s) >> 40) * (1.0f / 16777216.0f); }
/* Uniform integer in [0, n). */
static uint32_t rng_below(uint64_t* s, uint32_t n) { return (uint32_t)(((splitmix64_next(s) >> 32) * (uint64_t)n) >> 32); }

/* ------------------------------------------------------------------------ */

static int params_ok(const tgx_gen_params* p)
{
    if (!p) return 0;
    if (p->n_luts == 0) return 0;
    if (p->kind == TGX_GEN_SINGLE_BERNOULLI) return 1;
    if (p->kind == TGX_GEN_NOISE_OBJECTS)
        return p->n_objects > 0 && p->max_extent > 0 && p->max_extent <= 1024 && p->world > 0.0f;
    return 0;
}

/* Per-object extents for NOISE_OBJECTS, drawn from the extents stream. */
static void object_extent(const tgx_gen_params* p, uint64_t* s, uint32_t n[3])
{
    for (int a = 0; a < 3; a++) n[a] = 1u + rng_below(s, p->max_extent);
}

int tgx_gen_scene_sizes(const tgx_gen_params* p, tgx_gen_sizes* out)
{
    if (!params_ok(p) || !out) return TGX_ERR_INVALID_ARG;
    memset(out, 0, sizeof(*out));
    out->n_luts = p->n_luts;
    if (p->kind == TGX_GEN_SINGLE_BERNOULLI) {
        out->n_objects = 1;
        out->n_cluster_pointers = 64;
        out->n_clusters = 64;
        return TGX_OK;
    }
    uint64_t s = tgx_gen_stream_seed(p->seed, STREAM_EXTENTS);
    uint64_t total = 0;
    for (uint32_t o = 0; o < p->n_objects; o++) {
        uint32_t n[3];
        object_extent(p, &s, n);
        total += (uint64_t)n[0] * n[1] * n[2];
    }
    if (total >= (1ull << 31)) return TGX_ERR_CAPACITY;
    out->n_objects = p->n_objects;
    out->n_cluster_pointers = (uint32_t)total;
    out->n_clusters = (uint32_t)total;
    return TGX_OK;
}

/* Rotation matrix of a uniform unit quaternion (Shoemake 1992). */
static void random_rotation(uint64_t* s, float m[16])
{
    const double u1 = rng_unit(s), u2 = rng_unit(s), u3 = rng_unit(s);
    const double a = sqrt(1.0 - u1), b = sqrt(u1);
    const double qx = a * sin(2.0 * TGX_PI_D * u2), qy = a * cos(2.0 * TGX_PI_D * u2);
    const double qz = b * sin(2.0 * TGX_PI_D * u3), qw = b * cos(2.0 * TGX_PI_D * u3);
    double r[3][3];
    r[0][0] = 1 - 2 * (qy * qy + qz * qz); r[0][1] = 2 * (qx * qy - qz * qw); r[0][2] = 2 * (qx * qz + qy * qw);
    r[1][0] = 2 * (qx * qy + qz * qw); r[1][1] = 1 - 2 * (qx * qx + qz * qz); r[1][2] = 2 * (qy * qz - qx * qw);
    r[2][0] = 2 * (qx * qz - qy * qw); r[2][1] = 2 * (qy * qz + qx * qw); r[2][2] = 1 - 2 * (qx * qx + qy * qy);
    for (int i = 0; i < 16; i++) m[i] = 0.0f;
    for (int row = 0; row < 3; row++)
        for (int col = 0; col < 3; col++) m[4 * col + row] = (float)r[row][col];
    m[15] = 1.0f;
}

void tgx_gen_luts(const tgx_gen_params* p, tgx_lut_entry* luts)
{
    uint64_t s = tgx_gen_stream_seed(p->seed, STREAM_LUT);
    for (uint32_t i = 0; i < TGX_LUT_ENTRIES * p->n_luts; i++) {
        luts[i].albedo[0] = rng_unit_f(&s);
        luts[i].albedo[1] = rng_unit_f(&s);
        luts[i].albedo[2] = rng_unit_f(&s);
        luts[i].roughness = rng_unit_f(&s);
    }
}

static void fill_single(const tgx_gen_params* p, tgx_object* objects, uint32_t* cptrs,
                        uint32_t* c2o, uint32_t* bits, uint8_t* materials)
{
    tgx_object* o = &objects[0];
    memset(o, 0, sizeof(*o));
    o->n_cluster_pointers_per_dim[0] = 4;
    o->n_cluster_pointers_per_dim[1] = 4;
    o->n_cluster_pointers_per_dim[2] = 4;
    o->first_cluster_pointer = 0;
    o->lut_idx = 0;
    o->rotation[0] = o->rotation[5] = o->rotation[10] = o->rotation[15] = 1.0f;
    for (uint32_t c = 0; c < 64; c++) { cptrs[c] = c; c2o[c] = 0; }

    /* Bernoulli(0.3) per voxel: a 24-bit uniform below 0.3 * 2^24. */
    uint64_t s = tgx_gen_stream_seed(p->seed, STREAM_SOLID);
    const uint64_t threshold = 5033165ull; /* ceil(0.3 * 2^24) */
    memset(bits, 0, 64 * TGX_BITS_WORDS_PER_CLUSTER * sizeof(uint32_t));
    for (uint32_t i = 0; i < 64 * TGX_VOXELS_PER_CLUSTER; i++)
        if ((splitmix64_next(&s) >> 40) < threshold) bits[i >> 5] |= 1u << (i & 31u);

    uint64_t sm = tgx_gen_stream_seed(p->seed, STREAM_MATERIAL);
    for (uint32_t i = 0; i < 64 * TGX_VOXELS_PER_CLUSTER; i++)
        materials[i] = (uint8_t)(splitmix64_next(&sm) >> 56);
}

int tgx_gen_noise_objects(const tgx_gen_params* p, tgx_object* objects, double* noise_offsets)
{
    if (!params_ok(p) || p->kind != TGX_GEN_NOISE_OBJECTS || !objects || !noise_offsets) return TGX_ERR_INVALID_ARG;
    uint64_t se = tgx_gen_stream_seed(p->seed, STREAM_EXTENTS);
    uint64_t st = tgx_gen_stream_seed(p->seed, STREAM_TRANSFORM);
    uint64_t sl = tgx_gen_stream_seed(p->seed, STREAM_LUT_IDX);
    uint64_t sn = tgx_gen_stream_seed(p->seed, STREAM_NOISE);
    uint32_t cptr = 0;
    for (uint32_t oi = 0; oi < p->n_objects; oi++) {
        tgx_object* o = &objects[oi];
        memset(o, 0, sizeof(*o));
        uint32_t n[3];
        object_extent(p, &se, n);
        o->n_cluster_pointers_per_dim[0] = n[0];
        o->n_cluster_pointers_per_dim[1] = n[1];
        o->n_cluster_pointers_per_dim[2] = n[2];
        o->first_cluster_pointer = cptr;
        random_rotation(&st, o->rotation);
        o->translation[0] = (float)(rng_unit(&st) * (double)p->world);
        o->translation[1] = (float)(rng_unit(&st) * (double)p->world);
        o->translation[2] = (float)(rng_unit(&st) * (double)p->world);
        o->lut_idx = p->n_luts > 1 ? rng_below(&sl, p->n_luts) : 0u;
        /* Per-object offset into the noise field, so objects differ. */
        noise_offsets[3 * oi] = rng_unit(&sn) * 65536.0;
        noise_offsets[3 * oi + 1] = rng_unit(&sn) * 65536.0;
        noise_offsets[3 * oi + 2] = rng_unit(&sn) * 65536.0;
        cptr += n[0] * n[1] * n[2];
    }
    return TGX_OK;
}

uint64_t tgx_gen_noise_key(uint64_t seed, int material)
{
    return tgx_gen_stream_seed(seed, material ? STREAM_MATERIAL : STREAM_SOLID);
}

static void fill_noise(const tgx_gen_params* p, tgx_object* objects, uint32_t* cptrs,
                       uint32_t* c2o, uint32_t* bits, uint8_t* materials)
{
    const uint64_t key_solid = tgx_gen_noise_key(p->seed, 0);
    const uint64_t key_mat = tgx_gen_noise_key(p->seed, 1);
    double* off = (double*)malloc(sizeof(double) * 3 * p->n_objects);
    if (!off) return;
    (void)tgx_gen_noise_objects(p, objects, off);
    for (uint32_t oi = 0; oi < p->n_objects; oi++) {
        const tgx_object* o = &objects[oi];
        const uint32_t* n = o->n_cluster_pointers_per_dim;
        const uint32_t nc = n[0] * n[1] * n[2];
        for (uint32_t rc = 0; rc < nc; rc++) {
            const uint32_t c = o->first_cluster_pointer + rc; /* cluster idx == pointer */
            cptrs[c] = c;
            c2o[c] = oi;
            const uint32_t cx = rc % n[0], cy = (rc / n[0]) % n[1], cz = rc / (n[0] * n[1]);
            uint32_t* w = &bits[(size_t)c * TGX_BITS_WORDS_PER_CLUSTER];
            memset(w, 0, TGX_BITS_WORDS_PER_CLUSTER * sizeof(uint32_t));
            uint8_t* mat = &materials[(size_t)c * TGX_VOXELS_PER_CLUSTER];
            for (uint32_t v = 0; v < TGX_VOXELS_PER_CLUSTER; v++) {
                int solid;
                tgx_noise_voxel(key_solid, key_mat, off[3 * oi], off[3 * oi + 1], off[3 * oi + 2], cx, cy, cz, v,
                                &solid, &mat[v]);
                if (solid) w[v >> 5] |= 1u << (v & 31u);
            }
        }
    }
    free(off);
}

int tgx_gen_scene_fill(const tgx_gen_params* p, tgx_object* objects, uint32_t* cluster_pointers,
                       uint32_t* cluster_idx_to_object, uint32_t* bits, uint8_t* materials,
                       tgx_lut_entry* luts)
{
    if (!params_ok(p) || !objects || !cluster_pointers || !cluster_idx_to_object || !bits ||
        !materials || !luts)
        return TGX_ERR_INVALID_ARG;
    if (p->kind == TGX_GEN_SINGLE_BERNOULLI)
        fill_single(p, objects, cluster_pointers, cluster_idx_to_object, bits, materials);
    else
        fill_noise(p, objects, cluster_pointers, cluster_idx_to_object, bits, materials);
    tgx_gen_luts(p, luts);
    return TGX_OK;
}

/* ------------------------------------------------------------------------ */
/* Camera (tgvk_core.c:382-444, tg_math.c:2020-2028, 2469-2499, 2501-2590)   */
/* ------------------------------------------------------------------------ */

static tgx_m4 rotate_x(float a)
{
    const float c = cosf(a), s = sinf(a);
    tgx_m4 r = tgx_m4_identity();
    TGX_M(r, 1, 1) = c; TGX_M(r, 2, 1) = s; TGX_M(r, 1, 2) = -s; TGX_M(r, 2, 2) = c;
    return r;
}

static tgx_m4 rotate_y(float a)
{
    const float c = cosf(a), s = sinf(a);
    tgx_m4 r = tgx_m4_identity();
    TGX_M(r, 0, 0) = c; TGX_M(r, 2, 0) = -s; TGX_M(r, 0, 2) = s; TGX_M(r, 2, 2) = c;
    return r;
}

static tgx_m4 rotate_z(float a)
{
    const float c = cosf(a), s = sinf(a);
    tgx_m4 r = tgx_m4_identity();
    TGX_M(r, 0, 0) = c; TGX_M(r, 1, 0) = s; TGX_M(r, 0, 1) = -s; TGX_M(r, 1, 1) = c;
    return r;
}

static tgx_m4 euler(float pitch, float yaw, float roll)
{
    return tgx_m4_mul(rotate_z(roll), tgx_m4_mul(rotate_y(yaw), rotate_x(pitch)));
}

static tgx_m4 perspective(float fov_y, float aspect, float n, float f)
{
    const float t = tanf(fov_y / 2.0f);
    const float a = f / (n - f);
    const float b = -(2.0f * f * n) / (f - n);
    tgx_m4 r;
    for (int i = 0; i < 16; i++) r.m[i] = 0.0f;
    TGX_M(r, 0, 0) = 1.0f / (aspect * t);
    TGX_M(r, 1, 1) = -1.0f / t;
    TGX_M(r, 2, 2) = a;
    TGX_M(r, 3, 2) = -1.0f;
    TGX_M(r, 2, 3) = b;
    return r;
}

static void normalized_xyz(tgx_v4 v, float out[4])
{
    const float mag = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    out[0] = v.x / mag; out[1] = v.y / mag; out[2] = v.z / mag; out[3] = 0.0f;
}

int tgx_camera_make(const float position[3], float pitch, float yaw, float roll, float fov_y,
                    float aspect, float near_plane, float far_plane, tgx_camera* out)
{
    if (!position || !out || !(fov_y > 0.0f) || !(aspect > 0.0f) || !(near_plane > 0.0f) ||
        !(far_plane > near_plane))
        return TGX_ERR_INVALID_ARG;
    const tgx_m4 r = tgx_m4_inverse(euler(pitch, yaw, roll));
    const tgx_m4 p = perspective(fov_y, aspect, near_plane, far_plane);
    const tgx_m4 ivp = tgx_m4_inverse(tgx_m4_mul(p, r));
    memset(out, 0, sizeof(*out));
    out->camera[0] = position[0]; out->camera[1] = position[1]; out->camera[2] = position[2];
    normalized_xyz(tgx_m4_mulv4(ivp, -1.0f,  1.0f, 1.0f, 1.0f), out->ray_bl);
    normalized_xyz(tgx_m4_mulv4(ivp,  1.0f,  1.0f, 1.0f, 1.0f), out->ray_br);
    normalized_xyz(tgx_m4_mulv4(ivp,  1.0f, -1.0f, 1.0f, 1.0f), out->ray_tr);
    normalized_xyz(tgx_m4_mulv4(ivp, -1.0f, -1.0f, 1.0f, 1.0f), out->ray_tl);
    out->near_plane = near_plane;
    out->far_plane = far_plane;
    return TGX_OK;
}

void tgx_camera_forward(float pitch, float yaw, float out_forward[3])
{
    /* euler(pitch, yaw, 0) applied to the view-space forward (0, 0, -1). */
    const tgx_v4 f = tgx_m4_mulv4(euler(pitch, yaw, 0.0f), 0.0f, 0.0f, -1.0f, 0.0f);
    out_forward[0] = f.x; out_forward[1] = f.y; out_forward[2] = f.z;
}

int tgx_gen_camera(const tgx_gen_params* p, uint32_t camera_index, uint32_t w, uint32_t h,
                   tgx_camera* out)
{
    if (!params_ok(p) || !out || w == 0 || h == 0) return TGX_ERR_INVALID_ARG;
    uint64_t s = tgx_gen_stream_seed(p->seed, STREAM_CAMERA);
    float yaw = 0.0f, pitch = 0.0f;
    const float pitch_range = p->kind == TGX_GEN_SINGLE_BERNOULLI ? 0.6f : 0.5f;
    for (uint32_t i = 0; i <= camera_index; i++) {
        yaw = (float)(rng_unit(&s) * 2.0 * TGX_PI_D);
        pitch = (float)((rng_unit(&s) * 2.0 - 1.0) * (double)pitch_range);
    }
    const float fov = (float)(60.0 * TGX_PI_D / 180.0);
    const float aspect = (float)w / (float)h;
    float pos[3];
    if (p->kind == TGX_GEN_SINGLE_BERNOULLI) {
        float fwd[3];
        tgx_camera_forward(pitch, yaw, fwd);
        pos[0] = -64.0f * fwd[0]; pos[1] = -64.0f * fwd[1]; pos[2] = -64.0f * fwd[2];
    } else {
        pos[0] = pos[1] = pos[2] = 0.5f * p->world;
    }
    return tgx_camera_make(pos, pitch, yaw, 0.0f, fov, aspect, 0.1f, 1000.0f, out);
}
