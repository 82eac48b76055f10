/*
 * tgx_oracle.c — CPU oracle of tg's primary visibility (TEST INFRASTRUCTURE
 * ONLY; see tgx_oracle.h).
 *
 * Strict IEEE fp32: every expression below is evaluated in the order
 * written (-ffp-contract=off, SSE2, FLT_EVAL_METHOD == 0). min/max/clamp
 * are the comparisons written out, never fminf/fmaxf.
 *
 * Parity pins (tests/test_oracle_visibility.py):
 *   - slab test, DDA and packing against closed forms: axis-aligned rays
 *     through single voxels at known distances, full and empty clusters;
 *   - the culled frame equals the exhaustive min on every tested scene;
 *   - m4 inverse and product against numpy in float64 (to fp32 rounding);
 *   - cluster-pointer permutation invariance of the frame.
 */
#include "tgx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define F32_MAX 3.402823466e+38f
#define F32_MIN (-F32_MAX)

typedef struct { float x, y, z; } v3;
typedef struct { float m[16]; } m4; /* column-major, m[4*col + row] */
#define E(mat, r, c) ((mat).m[4 * (c) + (r)])

static float fmin2(float a, float b) { return a < b ? a : b; }          /* tg_math.h:20 */
static float fmax2(float a, float b) { return a > b ? a : b; }          /* tg_math.h:17 */
/* GLSL max(x, y) = (x < y) ? y : x */
static float glsl_max(float x, float y) { return x < y ? y : x; }
/* GLSL clamp(x, lo, hi) = min(max(x, lo), hi), with GLSL min/max. */
static float glsl_clamp(float x, float lo, float hi)
{
    const float t = x < lo ? lo : x;
    return hi < t ? hi : t;
}

/* ------------------------------------------------------------------------ */
/* m4 (tg_math.c:2185-2234, 2374-2438, 2669-2700)                            */
/* ------------------------------------------------------------------------ */

static m4 m4_identity(void)
{
    m4 r;
    memset(&r, 0, sizeof(r));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}

static m4 m4_translate(float x, float y, float z)
{
    m4 r = m4_identity();
    E(r, 0, 3) = x;
    E(r, 1, 3) = y;
    E(r, 2, 3) = z;
    return r;
}

static m4 m4_mul(m4 a, m4 b)
{
    m4 r;
    for (int c = 0; c < 4; c++)
        for (int i = 0; i < 4; i++)
            E(r, i, c) = E(a, i, 0) * E(b, 0, c) + E(a, i, 1) * E(b, 1, c) +
                         E(a, i, 2) * E(b, 2, c) + E(a, i, 3) * E(b, 3, c);
    return r;
}

static m4 m4_inverse(m4 m)
{
    const float m2323 = E(m,2,2) * E(m,3,3) - E(m,2,3) * E(m,3,2);
    const float m1323 = E(m,2,1) * E(m,3,3) - E(m,2,3) * E(m,3,1);
    const float m1223 = E(m,2,1) * E(m,3,2) - E(m,2,2) * E(m,3,1);
    const float m0323 = E(m,2,0) * E(m,3,3) - E(m,2,3) * E(m,3,0);
    const float m0223 = E(m,2,0) * E(m,3,2) - E(m,2,2) * E(m,3,0);
    const float m0123 = E(m,2,0) * E(m,3,1) - E(m,2,1) * E(m,3,0);
    const float m2313 = E(m,1,2) * E(m,3,3) - E(m,1,3) * E(m,3,2);
    const float m1313 = E(m,1,1) * E(m,3,3) - E(m,1,3) * E(m,3,1);
    const float m1213 = E(m,1,1) * E(m,3,2) - E(m,1,2) * E(m,3,1);
    const float m2312 = E(m,1,2) * E(m,2,3) - E(m,1,3) * E(m,2,2);
    const float m1312 = E(m,1,1) * E(m,2,3) - E(m,1,3) * E(m,2,1);
    const float m1212 = E(m,1,1) * E(m,2,2) - E(m,1,2) * E(m,2,1);
    const float m0313 = E(m,1,0) * E(m,3,3) - E(m,1,3) * E(m,3,0);
    const float m0213 = E(m,1,0) * E(m,3,2) - E(m,1,2) * E(m,3,0);
    const float m0312 = E(m,1,0) * E(m,2,3) - E(m,1,3) * E(m,2,0);
    const float m0212 = E(m,1,0) * E(m,2,2) - E(m,1,2) * E(m,2,0);
    const float m0113 = E(m,1,0) * E(m,3,1) - E(m,1,1) * E(m,3,0);
    const float m0112 = E(m,1,0) * E(m,2,1) - E(m,1,1) * E(m,2,0);

    const float det = 1.0f / (
        E(m,0,0) * (E(m,1,1) * m2323 - E(m,1,2) * m1323 + E(m,1,3) * m1223) -
        E(m,0,1) * (E(m,1,0) * m2323 - E(m,1,2) * m0323 + E(m,1,3) * m0223) +
        E(m,0,2) * (E(m,1,0) * m1323 - E(m,1,1) * m0323 + E(m,1,3) * m0123) -
        E(m,0,3) * (E(m,1,0) * m1223 - E(m,1,1) * m0223 + E(m,1,2) * m0123));

    m4 r;
    E(r,0,0) = det *  (E(m,1,1) * m2323 - E(m,1,2) * m1323 + E(m,1,3) * m1223);
    E(r,0,1) = det * -(E(m,0,1) * m2323 - E(m,0,2) * m1323 + E(m,0,3) * m1223);
    E(r,0,2) = det *  (E(m,0,1) * m2313 - E(m,0,2) * m1313 + E(m,0,3) * m1213);
    E(r,0,3) = det * -(E(m,0,1) * m2312 - E(m,0,2) * m1312 + E(m,0,3) * m1212);
    E(r,1,0) = det * -(E(m,1,0) * m2323 - E(m,1,2) * m0323 + E(m,1,3) * m0223);
    E(r,1,1) = det *  (E(m,0,0) * m2323 - E(m,0,2) * m0323 + E(m,0,3) * m0223);
    E(r,1,2) = det * -(E(m,0,0) * m2313 - E(m,0,2) * m0313 + E(m,0,3) * m0213);
    E(r,1,3) = det *  (E(m,0,0) * m2312 - E(m,0,2) * m0312 + E(m,0,3) * m0212);
    E(r,2,0) = det *  (E(m,1,0) * m1323 - E(m,1,1) * m0323 + E(m,1,3) * m0123);
    E(r,2,1) = det * -(E(m,0,0) * m1323 - E(m,0,1) * m0323 + E(m,0,3) * m0123);
    E(r,2,2) = det *  (E(m,0,0) * m1313 - E(m,0,1) * m0313 + E(m,0,3) * m0113);
    E(r,2,3) = det * -(E(m,0,0) * m1312 - E(m,0,1) * m0312 + E(m,0,3) * m0112);
    E(r,3,0) = det * -(E(m,1,0) * m1223 - E(m,1,1) * m0223 + E(m,1,2) * m0123);
    E(r,3,1) = det *  (E(m,0,0) * m1223 - E(m,0,1) * m0223 + E(m,0,2) * m0123);
    E(r,3,2) = det * -(E(m,0,0) * m1213 - E(m,0,1) * m0213 + E(m,0,2) * m0113);
    E(r,3,3) = det *  (E(m,0,0) * m1212 - E(m,0,1) * m0212 + E(m,0,2) * m0112);
    return r;
}

/* mulv4 with w given (tg_math.c:2428-2438). */
static v3 m4_mulv(m4 m, v3 v, float w)
{
    v3 r;
    r.x = v.x * E(m,0,0) + v.y * E(m,0,1) + v.z * E(m,0,2) + w * E(m,0,3);
    r.y = v.x * E(m,1,0) + v.y * E(m,1,1) + v.z * E(m,1,2) + w * E(m,1,3);
    r.z = v.x * E(m,2,0) + v.y * E(m,2,1) + v.z * E(m,2,2) + w * E(m,2,3);
    return r;
}

void tgxo_m4_mul(const float a[16], const float b[16], float out[16])
{
    m4 ma, mb;
    memcpy(ma.m, a, sizeof(ma.m));
    memcpy(mb.m, b, sizeof(mb.m));
    const m4 r = m4_mul(ma, mb);
    memcpy(out, r.m, sizeof(r.m));
}

void tgxo_m4_inverse(const float a[16], float out[16])
{
    m4 ma;
    memcpy(ma.m, a, sizeof(ma.m));
    const m4 r = m4_inverse(ma);
    memcpy(out, r.m, sizeof(r.m));
}

/* World -> cluster space (visibility.frag:35-62): ws2ms3*ws2ms2*ws2ms1*ws2ms0
 * = T(-offset) * T(half_object) * inverse(R) * T(-t), left-associated. */
static m4 world_to_cluster(const tgx_object* o, uint32_t rel)
{
    const uint32_t nx = o->n_cluster_pointers_per_dim[0], ny = o->n_cluster_pointers_per_dim[1];
    const uint32_t rx = rel % nx, ry = (rel / nx) % ny, rz = rel / (nx * ny);
    const float offx = (float)(rx * 8u), offy = (float)(ry * 8u), offz = (float)(rz * 8u);
    const float half = 8.0f * 0.5f; /* cluster_extent * 0.5 */
    const float hx = half * (float)o->n_cluster_pointers_per_dim[0];
    const float hy = half * (float)o->n_cluster_pointers_per_dim[1];
    const float hz = half * (float)o->n_cluster_pointers_per_dim[2];
    m4 rot;
    memcpy(rot.m, o->rotation, sizeof(rot.m));
    const m4 w0 = m4_translate(-o->translation[0], -o->translation[1], -o->translation[2]);
    const m4 w1 = m4_inverse(rot);
    const m4 w2 = m4_translate(hx, hy, hz);
    const m4 w3 = m4_translate(-offx, -offy, -offz);
    return m4_mul(m4_mul(m4_mul(w3, w2), w1), w0);
}

void tgxo_world_to_cluster(const tgx_object* o, uint32_t rel_cptr, float out[16])
{
    const m4 r = world_to_cluster(o, rel_cptr);
    memcpy(out, r.m, sizeof(r.m));
}

/* ------------------------------------------------------------------------ */
/* Ray helpers                                                              */
/* ------------------------------------------------------------------------ */

/* collide.inc:3-24 (identical to tg_physics.c:394-406 with
 * tgm_v3_div_zero_check, tg_math.c:1044-1051). */
static int intersect_ray_aabb(v3 o, v3 d, v3 mn, v3 mx, float* enter, float* exit_)
{
    const float v0x = (d.x == 0.0f) ? F32_MIN : ((mn.x - o.x) / d.x);
    const float v0y = (d.y == 0.0f) ? F32_MIN : ((mn.y - o.y) / d.y);
    const float v0z = (d.z == 0.0f) ? F32_MIN : ((mn.z - o.z) / d.z);
    const float v1x = (d.x == 0.0f) ? F32_MAX : ((mx.x - o.x) / d.x);
    const float v1y = (d.y == 0.0f) ? F32_MAX : ((mx.y - o.y) / d.y);
    const float v1z = (d.z == 0.0f) ? F32_MAX : ((mx.z - o.z) / d.z);
    const float nx = fmin2(v0x, v1x), ny = fmin2(v0y, v1y), nz = fmin2(v0z, v1z);
    const float fx = fmax2(v0x, v1x), fy = fmax2(v0y, v1y), fz = fmax2(v0z, v1z);
    *enter = fmax2(fmax2(nx, ny), nz);
    *exit_ = fmin2(fmin2(fx, fy), fz);
    return *exit_ > 0.0f && *enter <= *exit_;
}

int tgxo_intersect_ray_aabb(const float o[3], const float d[3], const float mn[3],
                            const float mx[3], float* enter, float* exit_)
{
    const v3 vo = {o[0], o[1], o[2]}, vd = {d[0], d[1], d[2]};
    const v3 vmn = {mn[0], mn[1], mn[2]}, vmx = {mx[0], mx[1], mx[2]};
    return intersect_ray_aabb(vo, vd, vmn, vmx, enter, exit_);
}

/* mix(a, b, t) = a*(1-t) + b*t (common.inc:40-46; tg_math.c:1091-1098). */
static v3 mix3(v3 a, v3 b, float t)
{
    const float omt = 1.0f - t;
    v3 r;
    r.x = a.x * omt + b.x * t;
    r.y = a.y * omt + b.y * t;
    r.z = a.z * omt + b.z * t;
    return r;
}

/* tg_lerp_corner_ray_directions_nn (common.inc:40-46). */
static v3 pixel_direction(const tgx_camera* c, float fx, float fy)
{
    const v3 bl = {c->ray_bl[0], c->ray_bl[1], c->ray_bl[2]};
    const v3 br = {c->ray_br[0], c->ray_br[1], c->ray_br[2]};
    const v3 tr = {c->ray_tr[0], c->ray_tr[1], c->ray_tr[2]};
    const v3 tl = {c->ray_tl[0], c->ray_tl[1], c->ray_tl[2]};
    return mix3(mix3(bl, tl, fy), mix3(br, tr, fy), fx);
}

/* v / sqrt(dot(v, v)) (tg_math.c:1106-1110, 1175-1184). */
static v3 normalize3(v3 v)
{
    const float mag = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    v3 r = {v.x / mag, v.y / mag, v.z / mag};
    return r;
}

/* Pixel sample position: gl_FragCoord = (x + 0.5, y + 0.5)
 * (visibility.frag:29-33). */
static float pixel_fx(uint32_t x, uint32_t w) { return ((float)x + 0.5f) / (float)w; }
static float pixel_fy(uint32_t y, uint32_t h) { return 1.0f - ((float)y + 0.5f) / (float)h; }

/* ------------------------------------------------------------------------ */
/* Per-cluster setup and per-pair evaluation                                */
/* ------------------------------------------------------------------------ */

typedef struct cluster_ray_frame {
    m4       ws2ms;
    v3       origin_ms;
    uint32_t cidx;
} cluster_ray_frame;

static cluster_ray_frame cluster_frame(const tgxo_scene* s, const tgx_camera* cam, uint32_t cptr)
{
    cluster_ray_frame f;
    f.cidx = s->cluster_pointers[cptr];
    const uint32_t obj = s->cluster_idx_to_object[f.cidx];
    const tgx_object* o = &s->objects[obj];
    f.ws2ms = world_to_cluster(o, cptr - o->first_cluster_pointer);
    const v3 cpos = {cam->camera[0], cam->camera[1], cam->camera[2]};
    f.origin_ms = m4_mulv(f.ws2ms, cpos, 1.0f);
    return f;
}

/* visibility.frag:65-206 for one pixel direction (world space). */
static uint64_t pair_word(const tgxo_scene* s, const cluster_ray_frame* f, uint32_t cptr,
                          v3 dir_ws, float far_plane, int* slab_hit)
{
    const v3 o = f->origin_ms;
    const v3 d = normalize3(m4_mulv(f->ws2ms, dir_ws, 0.0f));
    const v3 cmin = {0.0f, 0.0f, 0.0f};
    const v3 cmax = {8.0f, 8.0f, 8.0f};
    float enter, exit_;
    if (slab_hit) *slab_hit = 0;
    /* A zero direction component leaves that axis unconstrained in the
     * slab test (collide.inc:5-11), so a ray parallel to a slab but outside
     * it would "hit". tg never evaluates such a pair (the box's raster does
     * not cover the pixel); the frame definition rejects it explicitly. */
    if ((d.x == 0.0f && (o.x < 0.0f || o.x > 8.0f)) || (d.y == 0.0f && (o.y < 0.0f || o.y > 8.0f)) ||
        (d.z == 0.0f && (o.z < 0.0f || o.z > 8.0f)))
        return TGX_VISIBILITY_EMPTY;
    if (!intersect_ray_aabb(o, d, cmin, cmax, &enter, &exit_)) return TGX_VISIBILITY_EMPTY;
    if (slab_hit) *slab_hit = 1;

    v3 hit;
    if (enter > 0.0f) {
        hit.x = o.x + enter * d.x;
        hit.y = o.y + enter * d.y;
        hit.z = o.z + enter * d.z;
    } else {
        hit = o;
    }
    const float xf = glsl_clamp(floorf(hit.x), 0.0f, 7.0f);
    const float yf = glsl_clamp(floorf(hit.y), 0.0f, 7.0f);
    const float zf = glsl_clamp(floorf(hit.z), 0.0f, 7.0f);
    int x = (int)xf, y = (int)yf, z = (int)zf;

    int step_x = 0, step_y = 0, step_z = 0;
    float tmx = F32_MAX, tmy = F32_MAX, tmz = F32_MAX;
    float tdx = F32_MAX, tdy = F32_MAX, tdz = F32_MAX;
    if (d.x > 0.0f) { step_x = 1; tmx = enter + ((float)(x + 1) - hit.x) / d.x; tdx = 1.0f / d.x; }
    else if (d.x < 0.0f) { step_x = -1; tmx = enter + (hit.x - (float)x) / -d.x; tdx = 1.0f / -d.x; }
    if (d.y > 0.0f) { step_y = 1; tmy = enter + ((float)(y + 1) - hit.y) / d.y; tdy = 1.0f / d.y; }
    else if (d.y < 0.0f) { step_y = -1; tmy = enter + (hit.y - (float)y) / -d.y; tdy = 1.0f / -d.y; }
    if (d.z > 0.0f) { step_z = 1; tmz = enter + ((float)(z + 1) - hit.z) / d.z; tdz = 1.0f / d.z; }
    else if (d.z < 0.0f) { step_z = -1; tmz = enter + (hit.z - (float)z) / -d.z; tdz = 1.0f / -d.z; }

    const uint64_t cluster_offset = (uint64_t)f->cidx * 512u;
    for (uint32_t it = 0; it < TGX_DDA_MAX_ITERATIONS; it++) {
        const uint32_t rel = 64u * (uint32_t)z + 8u * (uint32_t)y + (uint32_t)x;
        const uint64_t vox = cluster_offset + rel;
        const uint32_t word = s->bits[vox / 32u];
        if (word & (1u << (uint32_t)(vox % 32u))) {
            const v3 vmin = {0.0f + (float)x, 0.0f + (float)y, 0.0f + (float)z};
            const v3 vmax = {0.0f + (float)(x + 1), 0.0f + (float)(y + 1), 0.0f + (float)(z + 1)};
            float ve, vx;
            (void)intersect_ray_aabb(o, d, vmin, vmax, &ve, &vx);
            const float dd = glsl_max(0.0f, ve / far_plane);
            if (dd <= 1.0f)
                return ((uint64_t)(dd * 16777215.0f) << 40) | ((uint64_t)cptr << 9) | (uint64_t)rel;
            return TGX_VISIBILITY_EMPTY;
        }
        if (tmx < tmy) {
            if (tmx < tmz) { tmx += tdx; x += step_x; if (x < 0 || x >= 8) break; }
            else           { tmz += tdz; z += step_z; if (z < 0 || z >= 8) break; }
        } else {
            if (tmy < tmz) { tmy += tdy; y += step_y; if (y < 0 || y >= 8) break; }
            else           { tmz += tdz; z += step_z; if (z < 0 || z >= 8) break; }
        }
    }
    return TGX_VISIBILITY_EMPTY;
}

uint64_t tgxo_vis_pair(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                       uint32_t px, uint32_t py, uint32_t cptr)
{
    const cluster_ray_frame f = cluster_frame(s, cam, cptr);
    const v3 dir = pixel_direction(cam, pixel_fx(px, w), pixel_fy(py, h));
    return pair_word(s, &f, cptr, dir, cam->far_plane, NULL);
}

void tgxo_visibility_brute(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                           uint64_t* out)
{
    cluster_ray_frame* frames = (cluster_ray_frame*)malloc(sizeof(*frames) * (s->n_cluster_pointers + 1u));
    for (uint32_t c = 0; c < s->n_cluster_pointers; c++) frames[c] = cluster_frame(s, cam, c);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t yy = 0; yy < (int64_t)h; yy++) {
        const uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < w; x++) {
            const v3 dir = pixel_direction(cam, pixel_fx(x, w), pixel_fy(y, h));
            uint64_t best = TGX_VISIBILITY_EMPTY;
            for (uint32_t c = 0; c < s->n_cluster_pointers; c++) {
                const uint64_t word = pair_word(s, &frames[c], c, dir, cam->far_plane, NULL);
                if (word < best) best = word;
            }
            out[(size_t)y * w + x] = best;
        }
    }
    free(frames);
}

/* ------------------------------------------------------------------------ */
/* Conservative screen rectangle (float64)                                  */
/*                                                                          */
/* The pixel ray is dir(fx, fy) = bl + fx U + fy V with U = br - bl,        */
/* V = tl - bl (the corner lerp of a planar corner quad). A point P maps to */
/* w = [bl U V]^-1 (P - cam) = t (1, fx, fy). The cluster box, expanded by  */
/* GAMMA voxels to absorb fp32 rounding of the frame and of the pair test,  */
/* is clipped to w0 >= EPS and its vertices projected; points with          */
/* 0 < w0 < EPS can only land on screen inside the cube [0, EPS]^3, which   */
/* forces the whole screen.                                                 */
/* ------------------------------------------------------------------------ */

#define RECT_GAMMA 0.03125
#define RECT_EPS   0.25
#define RECT_DELTA 0.01
#define RECT_QUAD_TOL 1e-5

typedef struct cam_basis {
    double inv[3][3];
    double cam[3];
    int    planar;
} cam_basis;

static cam_basis make_basis(const tgx_camera* c)
{
    cam_basis b;
    double m[3][3];
    for (int i = 0; i < 3; i++) {
        m[i][0] = c->ray_bl[i];
        m[i][1] = (double)c->ray_br[i] - (double)c->ray_bl[i];
        m[i][2] = (double)c->ray_tl[i] - (double)c->ray_bl[i];
        b.cam[i] = c->camera[i];
    }
    double quad = 0.0;
    for (int i = 0; i < 3; i++) {
        const double q = (double)c->ray_tr[i] - (double)c->ray_br[i] - (double)c->ray_tl[i] + (double)c->ray_bl[i];
        quad += q * q;
    }
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    b.planar = sqrt(quad) <= RECT_QUAD_TOL && fabs(det) > 1e-12;
    if (!b.planar) return b;
    b.inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    b.inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    b.inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    b.inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    b.inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    b.inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    b.inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    b.inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    b.inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    return b;
}

typedef struct proj_range { double fxmin, fxmax, fymin, fymax; int n; } proj_range;

static void add_vertex(proj_range* r, double w0, double w1, double w2)
{
    const double fx = w1 / w0, fy = w2 / w0;
    if (fx < r->fxmin) r->fxmin = fx;
    if (fx > r->fxmax) r->fxmax = fx;
    if (fy < r->fymin) r->fymin = fy;
    if (fy > r->fymax) r->fymax = fy;
    r->n++;
}

/* Returns 0 (culled) or 1 with rect = {x0, y0, x1, y1} inclusive. */
static int rect_of(const cam_basis* b, const tgx_object* o, uint32_t rel, uint32_t w, uint32_t h,
                   uint32_t rect[4])
{
    if (!b->planar) { rect[0] = 0; rect[1] = 0; rect[2] = w - 1; rect[3] = h - 1; return 1; }
    const uint32_t nx = o->n_cluster_pointers_per_dim[0], ny = o->n_cluster_pointers_per_dim[1];
    const uint32_t rx = rel % nx, ry = (rel / nx) % ny, rz = rel / (nx * ny);
    const double base[3] = {-4.0 * nx + 8.0 * rx, -4.0 * ny + 8.0 * ry,
                            -4.0 * o->n_cluster_pointers_per_dim[2] + 8.0 * rz};
    double wv[8][3];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < 8; k++) {
        const double l[3] = {base[0] + ((k & 1) ? 8.0 + RECT_GAMMA : -RECT_GAMMA),
                             base[1] + ((k & 2) ? 8.0 + RECT_GAMMA : -RECT_GAMMA),
                             base[2] + ((k & 4) ? 8.0 + RECT_GAMMA : -RECT_GAMMA)};
        double p[3];
        for (int i = 0; i < 3; i++)
            p[i] = (double)o->rotation[i] * l[0] + (double)o->rotation[4 + i] * l[1] +
                   (double)o->rotation[8 + i] * l[2] + (double)o->translation[i] - b->cam[i];
        for (int i = 0; i < 3; i++) {
            wv[k][i] = b->inv[i][0] * p[0] + b->inv[i][1] * p[1] + b->inv[i][2] * p[2];
            if (wv[k][i] < lo[i]) lo[i] = wv[k][i];
            if (wv[k][i] > hi[i]) hi[i] = wv[k][i];
        }
    }
    /* Part of the box within EPS of the camera plane that may be on screen. */
    int near_cube = 1;
    for (int i = 0; i < 3; i++)
        if (hi[i] < -RECT_DELTA || lo[i] > RECT_EPS + RECT_DELTA) near_cube = 0;
    if (near_cube) { rect[0] = 0; rect[1] = 0; rect[2] = w - 1; rect[3] = h - 1; return 1; }
    if (hi[0] < RECT_EPS) return 0;

    proj_range pr = {1e300, -1e300, 1e300, -1e300, 0};
    for (int k = 0; k < 8; k++)
        if (wv[k][0] >= RECT_EPS) add_vertex(&pr, wv[k][0], wv[k][1], wv[k][2]);
    /* 12 box edges: corner pairs differing in one bit. */
    for (int k = 0; k < 8; k++) {
        for (int bit = 1; bit < 8; bit <<= 1) {
            const int j = k | bit;
            if (j == k) continue;
            const double a0 = wv[k][0], a1 = wv[j][0];
            if ((a0 < RECT_EPS) == (a1 < RECT_EPS)) continue;
            const double t = (RECT_EPS - a0) / (a1 - a0);
            add_vertex(&pr, RECT_EPS, wv[k][1] + t * (wv[j][1] - wv[k][1]),
                       wv[k][2] + t * (wv[j][2] - wv[k][2]));
        }
    }
    const int nv = pr.n;
    const double fxmin = pr.fxmin, fxmax = pr.fxmax, fymin = pr.fymin, fymax = pr.fymax;
    if (nv == 0) return 0;
    /* Pixel centres: fx = (x + 0.5)/w, fy = 1 - (y + 0.5)/h. One extra pixel
     * covers the rounding of the pixel position itself. */
    const double x0 = floor(fxmin * w - 0.5) - 1.0, x1 = ceil(fxmax * w - 0.5) + 1.0;
    const double y0 = floor((1.0 - fymax) * h - 0.5) - 1.0, y1 = ceil((1.0 - fymin) * h - 0.5) + 1.0;
    if (x1 < 0.0 || y1 < 0.0 || x0 > (double)(w - 1) || y0 > (double)(h - 1)) return 0;
    rect[0] = x0 < 0.0 ? 0u : (uint32_t)x0;
    rect[1] = y0 < 0.0 ? 0u : (uint32_t)y0;
    rect[2] = x1 > (double)(w - 1) ? w - 1 : (uint32_t)x1;
    rect[3] = y1 > (double)(h - 1) ? h - 1 : (uint32_t)y1;
    return 1;
}

int tgxo_cluster_rect(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                      uint32_t cptr, uint32_t rect[4])
{
    const cam_basis b = make_basis(cam);
    const uint32_t cidx = s->cluster_pointers[cptr];
    const tgx_object* o = &s->objects[s->cluster_idx_to_object[cidx]];
    return rect_of(&b, o, cptr - o->first_cluster_pointer, w, h, rect);
}

typedef struct culled_cluster {
    cluster_ray_frame frame;
    uint32_t          cptr;
    uint32_t          rect[4];
} culled_cluster;

static int cluster_is_empty(const tgxo_scene* s, uint32_t cidx)
{
    const uint32_t* wds = &s->bits[(size_t)cidx * 16u];
    for (int i = 0; i < 16; i++) if (wds[i]) return 0;
    return 1;
}

/* The clusters that can produce a hit and their pixel rectangles, in
 * pointer order. Each pointer is tested independently (in parallel); the
 * survivors are then compacted in order. */
static uint32_t collect_clusters(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                                 culled_cluster* list)
{
    const cam_basis b = make_basis(cam);
    const int64_t n_cptr = (int64_t)s->n_cluster_pointers;
#pragma omp parallel for schedule(static, 4096)
    for (int64_t ci = 0; ci < n_cptr; ci++) {
        const uint32_t c = (uint32_t)ci;
        const uint32_t cidx = s->cluster_pointers[c];
        list[c].cptr = UINT32_MAX;
        if (cluster_is_empty(s, cidx)) continue; /* no solid bit: never a hit */
        const tgx_object* o = &s->objects[s->cluster_idx_to_object[cidx]];
        uint32_t r[4];
        if (!rect_of(&b, o, c - o->first_cluster_pointer, w, h, r)) continue;
        list[c].frame = cluster_frame(s, cam, c);
        list[c].cptr = c;
        memcpy(list[c].rect, r, sizeof(r));
    }
    uint32_t n = 0;
    for (uint32_t c = 0; c < s->n_cluster_pointers; c++)
        if (list[c].cptr != UINT32_MAX) list[n++] = list[c];
    return n;
}

uint64_t tgxo_visibility_culled(const tgxo_scene* s, const tgx_camera* cam, uint32_t w,
                                uint32_t h, uint32_t row_step, int n_threads, uint64_t* out)
{
    if (row_step == 0) row_step = 1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    culled_cluster* list = (culled_cluster*)malloc(sizeof(*list) * (s->n_cluster_pointers + 1u));
    const uint32_t n = collect_clusters(s, cam, w, h, list);
    uint64_t pairs = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : pairs)
    for (int64_t yy = 0; yy < (int64_t)h; yy += row_step) {
        const uint32_t y = (uint32_t)yy;
        uint64_t* row = &out[(size_t)y * w];
        for (uint32_t x = 0; x < w; x++) row[x] = TGX_VISIBILITY_EMPTY;
        const float fy = pixel_fy(y, h);
        for (uint32_t i = 0; i < n; i++) {
            const culled_cluster* c = &list[i];
            if (y < c->rect[1] || y > c->rect[3]) continue;
            for (uint32_t x = c->rect[0]; x <= c->rect[2]; x++) {
                const v3 dir = pixel_direction(cam, pixel_fx(x, w), fy);
                const uint64_t word = pair_word(s, &c->frame, c->cptr, dir, cam->far_plane, NULL);
                if (word < row[x]) row[x] = word;
                pairs++;
            }
        }
    }
    free(list);
    return pairs;
}

/* ------------------------------------------------------------------------ */
/* Algorithmic bytes                                                        */
/* ------------------------------------------------------------------------ */

uint64_t tgxo_visibility_alg_bytes(const tgxo_scene* s, const tgx_camera* cam, uint32_t w,
                                   uint32_t h)
{
    culled_cluster* list = (culled_cluster*)malloc(sizeof(*list) * (s->n_cluster_pointers + 1u));
    const uint32_t n = collect_clusters(s, cam, w, h, list);
    /* Line bitmaps per input array (64-B lines, arrays assumed 64-B aligned). */
    const size_t n_cptr_lines = ((size_t)s->n_cluster_pointers * 4u + 63u) / 64u;
    const size_t n_c2o_lines = ((size_t)s->n_clusters * 4u + 63u) / 64u;
    const size_t n_obj_lines = ((size_t)s->n_objects * 96u + 63u) / 64u;
    unsigned char* cptr_l = (unsigned char*)calloc(n_cptr_lines + 1, 1);
    unsigned char* c2o_l = (unsigned char*)calloc(n_c2o_lines + 1, 1);
    unsigned char* obj_l = (unsigned char*)calloc(n_obj_lines + 1, 1);
    unsigned char* used = (unsigned char*)calloc((size_t)n + 1, 1);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const culled_cluster* c = &list[ii];
        int any = 0;
        for (uint32_t y = c->rect[1]; y <= c->rect[3] && !any; y++) {
            const float fy = pixel_fy(y, h);
            for (uint32_t x = c->rect[0]; x <= c->rect[2]; x++) {
                int slab = 0;
                const v3 dir = pixel_direction(cam, pixel_fx(x, w), fy);
                (void)pair_word(s, &c->frame, c->cptr, dir, cam->far_plane, &slab);
                if (slab) { any = 1; break; }
            }
        }
        used[ii] = (unsigned char)any;
    }
    uint64_t bits_lines = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (!used[i]) continue;
        const uint32_t cptr = list[i].cptr, cidx = list[i].frame.cidx;
        const uint32_t obj = s->cluster_idx_to_object[cidx];
        cptr_l[(size_t)cptr * 4u / 64u] = 1;
        c2o_l[(size_t)cidx * 4u / 64u] = 1;
        obj_l[(size_t)obj * 96u / 64u] = 1;
        obj_l[((size_t)obj * 96u + 95u) / 64u] = 1;
        bits_lines++; /* 64 B of solid bits per cluster = one line */
    }
    uint64_t lines = bits_lines;
    for (size_t i = 0; i < n_cptr_lines; i++) lines += cptr_l[i];
    for (size_t i = 0; i < n_c2o_lines; i++) lines += c2o_l[i];
    for (size_t i = 0; i < n_obj_lines; i++) lines += obj_l[i];
    free(cptr_l); free(c2o_l); free(obj_l); free(used); free(list);
    return 64u * lines + 8u * (uint64_t)w * h;
}
