/*
 * tgx_visibility.hip — primary visibility on gfx950.
 *
 * Replaces tg's instanced raster of cluster boxes (visibility.vert:19-26)
 * and the per-fragment DDA (visibility.frag:22-208) with three compute
 * passes per frame:
 *
 *   1. vis_setup     one thread per (camera, cluster pointer): drops empty
 *                    clusters and clusters no pixel ray can reach, builds
 *                    the cluster's world->cluster frame (exact fp32 order of
 *                    visibility.frag:57-66), a conservative pixel rectangle
 *                    and a lower bound of its depth; counts 16x16 tiles.
 *   2. vis_scan +    tile lists: exclusive scan of the tile counts and a
 *      vis_scatter   scatter of record indices into each covered tile.
 *   3. vis_traverse  one 256-thread workgroup per tile, one pixel per lane,
 *                    each wave an 8x8 pixel block. For each listed cluster
 *                    the wave skips it unless some lane's pixel lies in the
 *                    rectangle and that lane's current best depth is not
 *                    below the cluster's depth bound; the cluster's 64 B of
 *                    solid bits are staged in LDS and every active lane runs
 *                    the slab test and the 8^3 Amanatides-Woo walk, keeping
 *                    the minimum packed word. A tile is owned by one
 *                    workgroup, so the min needs no atomics and the buffer
 *                    needs no clear.
 *
 * Bit-exactness: all code here is compiled with -ffp-contract=off, IEEE
 * fp32 division and square root, denormals preserved; the per-pair
 * arithmetic follows visibility.frag operation by operation.
 */
#include "tgx_internal.h"
#include "tgx_m4.h"

#include <math.h>

#include <mutex>

#define VIS_GLOBAL_PIXELS 65536u
#define VIS_F32_MAX 3.402823466e+38f

/* Conservative-rectangle constants: the cluster box is grown by GAMMA
 * voxels, clipped at the camera-plane distance EPS; a box reaching within
 * EPS + DELTA of the camera in every basis coordinate covers the screen. */
#define VIS_RECT_GAMMA 0.03125f
#define VIS_RECT_EPS   0.25f
#define VIS_RECT_DELTA 0.01f

struct vis_cam {
    float    pos[3];
    float    far_plane;
    float    bl[3], br[3], tr[3], tl[3];
    float    inv[9];      /* [bl | br-bl | tl-bl]^-1, row-major            */
    uint32_t full_screen; /* corner quad not planar: no rectangle cull    */
    uint32_t rec_base;    /* first record slot of this camera             */
    uint32_t tile_base;   /* first tile of this camera                    */
    uint32_t pad;
};

/* Per-(camera, cluster) record, 80 B. */
struct vis_cluster {
    float    m[12];       /* world->cluster rows 0..2, columns 0..3        */
    float    o[3];        /* camera position in cluster space              */
    uint32_t cptr;
    uint32_t cidx;
    uint32_t rect_lo;     /* x0 | y0 << 16                                 */
    uint32_t rect_hi;     /* x1 | y1 << 16                                 */
    uint32_t depth_lb24;  /* no hit of this cluster has a smaller depth24  */
};
static_assert(sizeof(vis_cluster) == 80, "vis_cluster layout");

/* ------------------------------------------------------------------------ */
/* Pair evaluation (visibility.frag:65-206)                                 */
/* ------------------------------------------------------------------------ */

__device__ static inline float glsl_max(float x, float y) { return x < y ? y : x; }
__device__ static inline float glsl_clamp(float x, float lo, float hi)
{
    const float t = x < lo ? lo : x;
    return hi < t ? hi : t;
}
__device__ static inline float min2(float a, float b) { return a < b ? a : b; }
__device__ static inline float max2(float a, float b) { return a > b ? a : b; }

/* collide.inc:3-24 */
__device__ static inline bool slab(float ox, float oy, float oz, float dx, float dy, float dz,
                                   float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                   float& enter, float& exit_)
{
    const float v0x = (dx == 0.0f) ? -VIS_F32_MAX : ((mnx - ox) / dx);
    const float v0y = (dy == 0.0f) ? -VIS_F32_MAX : ((mny - oy) / dy);
    const float v0z = (dz == 0.0f) ? -VIS_F32_MAX : ((mnz - oz) / dz);
    const float v1x = (dx == 0.0f) ? VIS_F32_MAX : ((mxx - ox) / dx);
    const float v1y = (dy == 0.0f) ? VIS_F32_MAX : ((mxy - oy) / dy);
    const float v1z = (dz == 0.0f) ? VIS_F32_MAX : ((mxz - oz) / dz);
    const float nx = min2(v0x, v1x), ny = min2(v0y, v1y), nz = min2(v0z, v1z);
    const float fx = max2(v0x, v1x), fy = max2(v0y, v1y), fz = max2(v0z, v1z);
    enter = max2(max2(nx, ny), nz);
    exit_ = min2(min2(fx, fy), fz);
    return exit_ > 0.0f && enter <= exit_;
}

/* Candidate word of one (pixel, cluster) pair; lds_bits holds the 16 solid
 * words of the cluster. */
__device__ static inline uint64_t pair_word(const vis_cluster& c, const uint32_t* lds_bits,
                                            float dwx, float dwy, float dwz, float far_plane)
{
    /* normalize((ws2ms * v4(dir, 0)).xyz) */
    const float rx = dwx * c.m[0] + dwy * c.m[1] + dwz * c.m[2] + 0.0f * c.m[3];
    const float ry = dwx * c.m[4] + dwy * c.m[5] + dwz * c.m[6] + 0.0f * c.m[7];
    const float rz = dwx * c.m[8] + dwy * c.m[9] + dwz * c.m[10] + 0.0f * c.m[11];
    const float mag = sqrtf(rx * rx + ry * ry + rz * rz);
    const float dx = rx / mag, dy = ry / mag, dz = rz / mag;
    const float ox = c.o[0], oy = c.o[1], oz = c.o[2];

    /* A ray parallel to a slab but outside it is a miss (the slab test
     * alone leaves a zero-direction axis unconstrained, collide.inc:5-11). */
    if ((dx == 0.0f && (ox < 0.0f || ox > 8.0f)) || (dy == 0.0f && (oy < 0.0f || oy > 8.0f)) ||
        (dz == 0.0f && (oz < 0.0f || oz > 8.0f)))
        return TGX_VISIBILITY_EMPTY;
    float enter, exit_;
    if (!slab(ox, oy, oz, dx, dy, dz, 0.0f, 0.0f, 0.0f, 8.0f, 8.0f, 8.0f, enter, exit_))
        return TGX_VISIBILITY_EMPTY;

    float hx = ox, hy = oy, hz = oz;
    if (enter > 0.0f) {
        hx = ox + enter * dx;
        hy = oy + enter * dy;
        hz = oz + enter * dz;
    }
    int x = (int)glsl_clamp(floorf(hx), 0.0f, 7.0f);
    int y = (int)glsl_clamp(floorf(hy), 0.0f, 7.0f);
    int z = (int)glsl_clamp(floorf(hz), 0.0f, 7.0f);

    int sx = 0, sy = 0, sz = 0;
    float tmx = VIS_F32_MAX, tmy = VIS_F32_MAX, tmz = VIS_F32_MAX;
    float tdx = VIS_F32_MAX, tdy = VIS_F32_MAX, tdz = VIS_F32_MAX;
    if (dx > 0.0f) { sx = 1; tmx = enter + ((float)(x + 1) - hx) / dx; tdx = 1.0f / dx; }
    else if (dx < 0.0f) { sx = -1; tmx = enter + (hx - (float)x) / -dx; tdx = 1.0f / -dx; }
    if (dy > 0.0f) { sy = 1; tmy = enter + ((float)(y + 1) - hy) / dy; tdy = 1.0f / dy; }
    else if (dy < 0.0f) { sy = -1; tmy = enter + (hy - (float)y) / -dy; tdy = 1.0f / -dy; }
    if (dz > 0.0f) { sz = 1; tmz = enter + ((float)(z + 1) - hz) / dz; tdz = 1.0f / dz; }
    else if (dz < 0.0f) { sz = -1; tmz = enter + (hz - (float)z) / -dz; tdz = 1.0f / -dz; }

    for (uint32_t it = 0; it < TGX_DDA_MAX_ITERATIONS; it++) {
        /* cidx*512 is a multiple of 32: the word within the cluster is
         * rel/32 and the bit rel%32 (visibility.frag:143-148). */
        const uint32_t rel = 64u * (uint32_t)z + 8u * (uint32_t)y + (uint32_t)x;
        if (lds_bits[rel >> 5] & (1u << (rel & 31u))) {
            float ve, vx;
            (void)slab(ox, oy, oz, dx, dy, dz, 0.0f + (float)x, 0.0f + (float)y, 0.0f + (float)z,
                       0.0f + (float)(x + 1), 0.0f + (float)(y + 1), 0.0f + (float)(z + 1), ve, vx);
            const float d = glsl_max(0.0f, ve / far_plane);
            if (d <= 1.0f)
                return ((uint64_t)(d * 16777215.0f) << 40) | ((uint64_t)c.cptr << 9) | (uint64_t)rel;
            return TGX_VISIBILITY_EMPTY;
        }
        if (tmx < tmy) {
            if (tmx < tmz) { tmx += tdx; x += sx; if (x < 0 || x >= 8) break; }
            else           { tmz += tdz; z += sz; if (z < 0 || z >= 8) break; }
        } else {
            if (tmy < tmz) { tmy += tdy; y += sy; if (y < 0 || y >= 8) break; }
            else           { tmz += tdz; z += sz; if (z < 0 || z >= 8) break; }
        }
    }
    return TGX_VISIBILITY_EMPTY;
}

/* ------------------------------------------------------------------------ */
/* Pass 1: per-cluster setup and conservative rectangle                     */
/* ------------------------------------------------------------------------ */

struct rect_acc { float fxmin, fxmax, fymin, fymax; int n; };

__device__ static inline void rect_add(rect_acc& r, float w0, float w1, float w2)
{
    const float fx = w1 / w0, fy = w2 / w0;
    r.fxmin = fminf(r.fxmin, fx); r.fxmax = fmaxf(r.fxmax, fx);
    r.fymin = fminf(r.fymin, fy); r.fymax = fmaxf(r.fymax, fy);
    r.n++;
}

/* Returns false if no pixel ray can hit the grown box. */
__device__ static bool cluster_rect(const vis_cam& cam, const tgx_object& ob, uint32_t rx,
                                    uint32_t ry, uint32_t rz, uint32_t W, uint32_t H,
                                    uint32_t& x0, uint32_t& y0, uint32_t& x1, uint32_t& y1)
{
    if (cam.full_screen) { x0 = 0; y0 = 0; x1 = W - 1; y1 = H - 1; return true; }
    const float bx = -4.0f * (float)ob.n_cluster_pointers_per_dim[0] + 8.0f * (float)rx;
    const float by = -4.0f * (float)ob.n_cluster_pointers_per_dim[1] + 8.0f * (float)ry;
    const float bz = -4.0f * (float)ob.n_cluster_pointers_per_dim[2] + 8.0f * (float)rz;
    const float tcx = ob.translation[0] - cam.pos[0];
    const float tcy = ob.translation[1] - cam.pos[1];
    const float tcz = ob.translation[2] - cam.pos[2];
    float w[8][3];
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float lx = bx + ((k & 1) ? 8.0f + VIS_RECT_GAMMA : -VIS_RECT_GAMMA);
        const float ly = by + ((k & 2) ? 8.0f + VIS_RECT_GAMMA : -VIS_RECT_GAMMA);
        const float lz = bz + ((k & 4) ? 8.0f + VIS_RECT_GAMMA : -VIS_RECT_GAMMA);
        const float px = ob.rotation[0] * lx + ob.rotation[4] * ly + ob.rotation[8] * lz + tcx;
        const float py = ob.rotation[1] * lx + ob.rotation[5] * ly + ob.rotation[9] * lz + tcy;
        const float pz = ob.rotation[2] * lx + ob.rotation[6] * ly + ob.rotation[10] * lz + tcz;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float v = cam.inv[3 * i + 0] * px + cam.inv[3 * i + 1] * py + cam.inv[3 * i + 2] * pz;
            w[k][i] = v;
            lo[i] = fminf(lo[i], v);
            hi[i] = fmaxf(hi[i], v);
        }
    }
    bool near_cube = true;
#pragma unroll
    for (int i = 0; i < 3; i++)
        if (hi[i] < -VIS_RECT_DELTA || lo[i] > VIS_RECT_EPS + VIS_RECT_DELTA) near_cube = false;
    if (near_cube) { x0 = 0; y0 = 0; x1 = W - 1; y1 = H - 1; return true; }
    if (hi[0] < VIS_RECT_EPS) return false;

    rect_acc r = {3e38f, -3e38f, 3e38f, -3e38f, 0};
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (w[k][0] >= VIS_RECT_EPS) rect_add(r, w[k][0], w[k][1], w[k][2]);
#pragma unroll
    for (int k = 0; k < 8; k++) {
#pragma unroll
        for (int bit = 1; bit < 8; bit <<= 1) {
            if (k & bit) continue;
            const int j = k | bit;
            const float a0 = w[k][0], a1 = w[j][0];
            if ((a0 < VIS_RECT_EPS) == (a1 < VIS_RECT_EPS)) continue;
            const float t = (VIS_RECT_EPS - a0) / (a1 - a0);
            rect_add(r, VIS_RECT_EPS, w[k][1] + t * (w[j][1] - w[k][1]), w[k][2] + t * (w[j][2] - w[k][2]));
        }
    }
    if (r.n == 0) return false;
    const float fw = (float)W, fh = (float)H;
    const float fx0 = floorf(r.fxmin * fw - 0.5f) - 1.0f, fx1 = ceilf(r.fxmax * fw - 0.5f) + 1.0f;
    const float fy0 = floorf((1.0f - r.fymax) * fh - 0.5f) - 1.0f, fy1 = ceilf((1.0f - r.fymin) * fh - 0.5f) + 1.0f;
    if (fx1 < 0.0f || fy1 < 0.0f || fx0 > fw - 1.0f || fy0 > fh - 1.0f) return false;
    x0 = fx0 < 0.0f ? 0u : (uint32_t)fx0;
    y0 = fy0 < 0.0f ? 0u : (uint32_t)fy0;
    x1 = fx1 > fw - 1.0f ? W - 1 : (uint32_t)fx1;
    y1 = fy1 > fh - 1.0f ? H - 1 : (uint32_t)fy1;
    return true;
}

/* Object-level cull, one count per (camera, object): the object's cluster
 * pointers when kept, 0 when no cluster of the object can get a rectangle
 * from cluster_rect or a word with depth <= 1. The
 * object's box, grown by one voxel, goes through cluster_rect's transform
 * into the camera's corner-ray basis w (fx = w1 / w0, fy = w2 / w0). A
 * half-space test that holds at all eight corners holds on the whole box,
 * so on every cluster box in it and every clipped point cluster_rect
 * projects: all corners behind the near plane, or beyond one screen edge by
 * 3 pixels (the rectangle's padding is 1.5). A box that reaches the cube
 * around the camera, where cluster_rect accepts clusters as full-screen, is
 * kept. The one-voxel growth covers the clusters' own growth and the
 * rounding of both computations. An object whose bounding sphere (radius
 * of the grown box) lies beyond the far plane, by a margin of 0.1 % and one
 * voxel for the rounding of the pair's slab distance, holds no hit with
 * depth <= 1 and is culled for every camera. */
__global__ __launch_bounds__(256) void vis_objcull_k(const tgx_object* __restrict__ objects, uint32_t n_objects,
                                                     const uint2* __restrict__ obj_occ,
                                                     const vis_cam* __restrict__ cams, uint32_t W, uint32_t H,
                                                     uint32_t* __restrict__ objcnt)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ci = blockIdx.y;
    if (o >= n_objects) return;
    const vis_cam& cam = cams[ci];
    const tgx_object ob = objects[o];
    /* the object's cluster pointers to clusters with a solid voxel */
    const uint32_t n = obj_occ[o].y;
    uint8_t keep = n ? 1u : 0u;
    {
        const float ex = 4.0f * (float)ob.n_cluster_pointers_per_dim[0] + 1.0f;
        const float ey = 4.0f * (float)ob.n_cluster_pointers_per_dim[1] + 1.0f;
        const float ez = 4.0f * (float)ob.n_cluster_pointers_per_dim[2] + 1.0f;
        const float tcx = ob.translation[0] - cam.pos[0];
        const float tcy = ob.translation[1] - cam.pos[1];
        const float tcz = ob.translation[2] - cam.pos[2];
        const float dist = sqrtf(tcx * tcx + tcy * tcy + tcz * tcz);
        const float rad = sqrtf(ex * ex + ey * ey + ez * ez);
        if (dist - rad > cam.far_plane * 1.001f + 1.0f) keep = 0u;
    }
    if (keep && !cam.full_screen) {
        const float ex = 4.0f * (float)ob.n_cluster_pointers_per_dim[0] + 1.0f;
        const float ey = 4.0f * (float)ob.n_cluster_pointers_per_dim[1] + 1.0f;
        const float ez = 4.0f * (float)ob.n_cluster_pointers_per_dim[2] + 1.0f;
        const float tcx = ob.translation[0] - cam.pos[0];
        const float tcy = ob.translation[1] - cam.pos[1];
        const float tcz = ob.translation[2] - cam.pos[2];
        const float mx = 3.0f / (float)W, my = 3.0f / (float)H;
        float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
        bool left = true, right = true, below = true, above = true;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float lx = (k & 1) ? ex : -ex, ly = (k & 2) ? ey : -ey, lz = (k & 4) ? ez : -ez;
            const float px = ob.rotation[0] * lx + ob.rotation[4] * ly + ob.rotation[8] * lz + tcx;
            const float py = ob.rotation[1] * lx + ob.rotation[5] * ly + ob.rotation[9] * lz + tcy;
            const float pz = ob.rotation[2] * lx + ob.rotation[6] * ly + ob.rotation[10] * lz + tcz;
            float w[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                w[i] = cam.inv[3 * i + 0] * px + cam.inv[3 * i + 1] * py + cam.inv[3 * i + 2] * pz;
                lo[i] = fminf(lo[i], w[i]);
                hi[i] = fmaxf(hi[i], w[i]);
            }
            left = left && w[1] + mx * w[0] < 0.0f;
            right = right && w[1] - (1.0f + mx) * w[0] > 0.0f;
            below = below && w[2] + my * w[0] < 0.0f;
            above = above && w[2] - (1.0f + my) * w[0] > 0.0f;
        }
        bool near_cube = true;
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (hi[i] < -VIS_RECT_DELTA - 1.0f || lo[i] > VIS_RECT_EPS + VIS_RECT_DELTA + 1.0f) near_cube = false;
        const bool behind = hi[0] < VIS_RECT_EPS - 1.0f;
        if (!near_cube && (behind || left || right || below || above)) keep = 0u;
    }
    objcnt[(size_t)ci * 3u * n_objects + o] = keep ? n : 0u;
}

/* Per camera (one workgroup each): the kept objects (nonzero counts) in
 * object order, compacted into ids and exclusive offsets of their cluster
 * pointers (the camera's block of objcnt holds counts, ids, offsets, n_objects
 * u32 each); total[ci] the number of cluster pointers of kept objects and
 * n_kept[ci] the number of kept objects. */
__global__ __launch_bounds__(1024) void vis_objscan_k(uint32_t* __restrict__ objcnt, uint32_t n_objects,
                                                      uint32_t* __restrict__ total, uint32_t* __restrict__ n_kept)
{
    __shared__ uint32_t wsum[16], wnum[16];
    __shared__ uint32_t carry, carry_n;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t* cnt = objcnt + (size_t)blockIdx.x * 3u * n_objects;
    uint32_t* kid = objcnt + (size_t)blockIdx.x * 3u * n_objects + n_objects;
    uint32_t* koff = kid + n_objects;
    if (t == 0) { carry = 0u; carry_n = 0u; }
    __syncthreads();
    for (uint32_t base = 0; base < n_objects; base += 1024u) {
        const uint32_t i = base + t;
        const uint32_t v = i < n_objects ? cnt[i] : 0u;
        const uint32_t f = v ? 1u : 0u;
        uint32_t incl = v, ninc = f;
#pragma unroll
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint32_t u = __shfl_up(incl, d), un = __shfl_up(ninc, d);
            if (lane >= d) { incl += u; ninc += un; }
        }
        if (lane == 63u) { wsum[w] = incl; wnum[w] = ninc; }
        __syncthreads();
        uint32_t wofs = 0u, blk = 0u, wn = 0u, blkn = 0u;
        for (uint32_t k = 0; k < 16u; k++) {
            wofs += k < w ? wsum[k] : 0u;
            blk += wsum[k];
            wn += k < w ? wnum[k] : 0u;
            blkn += wnum[k];
        }
        const uint32_t c0 = carry, n0 = carry_n;
        if (f) {
            const uint32_t pos = n0 + wn + ninc - 1u;
            kid[pos] = i;
            koff[pos] = c0 + wofs + incl - v;
        }
        __syncthreads();
        if (t == 0) { carry = c0 + blk; carry_n = n0 + blkn; }
        __syncthreads();
    }
    if (t == 0) { total[blockIdx.x] = carry; n_kept[blockIdx.x] = carry_n; }
}

/* Builds the record of one (camera, cluster pointer) whose cluster holds a
 * solid voxel; false if it is out of every pixel ray's reach or beyond the
 * far plane. */
__device__ static bool setup_one(const tgx_object* __restrict__ objects, const uint32_t* __restrict__ cptrs,
                                 uint32_t oi, uint32_t cptr, const vis_cam& cam, uint32_t W, uint32_t H,
                                 vis_cluster& rec)
{
    const uint32_t cidx = cptrs[cptr];

    const tgx_object ob = objects[oi];
    const uint32_t rel = cptr - ob.first_cluster_pointer;
    const uint32_t nx = ob.n_cluster_pointers_per_dim[0], ny = ob.n_cluster_pointers_per_dim[1];
    const uint32_t rx = rel % nx, ry = (rel / nx) % ny, rz = rel / (nx * ny);
    uint32_t x0, y0, x1, y1;
    if (!cluster_rect(cam, ob, rx, ry, rz, W, H, x0, y0, x1, y1)) return false;

    const tgx_m4 m = tgx_world_to_cluster(ob.rotation, ob.translation, ob.n_cluster_pointers_per_dim, rx, ry, rz);
    const tgx_v4 o = tgx_m4_mulv4(m, cam.pos[0], cam.pos[1], cam.pos[2], 1.0f);

    /* Lower bound of the slab entry of any ray from o: its distance to the
     * box [0,8]^3 (unit direction), shrunk for fp32 rounding. */
    const float ddx = o.x < 0.0f ? -o.x : (o.x > 8.0f ? o.x - 8.0f : 0.0f);
    const float ddy = o.y < 0.0f ? -o.y : (o.y > 8.0f ? o.y - 8.0f : 0.0f);
    const float ddz = o.z < 0.0f ? -o.z : (o.z > 8.0f ? o.z - 8.0f : 0.0f);
    float lb = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) * 0.9999f - 1e-3f;
    lb = lb > 0.0f ? lb : 0.0f;
    const float dlb = lb / cam.far_plane;
    if (dlb > 1.0f) return false; /* every hit would have depth > 1 */
    uint32_t depth_lb24 = (uint32_t)(dlb * 16777215.0f);
    depth_lb24 = depth_lb24 > 0u ? depth_lb24 - 1u : 0u;

#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) rec.m[4 * r + c] = TGX_M(m, r, c);
    rec.o[0] = o.x; rec.o[1] = o.y; rec.o[2] = o.z;
    rec.cptr = cptr;
    rec.cidx = cidx;
    rec.rect_lo = x0 | (y0 << 16);
    rec.rect_hi = x1 | (y1 << 16);
    rec.depth_lb24 = depth_lb24;
    return true;
}

/* The wave walks the tiles of each lane's rectangle together, 64 tiles per
 * step, so one screen-sized rectangle does not serialise on one lane. */
template <typename F>
__device__ static inline void for_each_rect_tile(bool has, uint32_t rect_lo, uint32_t rect_hi,
                                                 uint64_t payload, uint32_t tshift, F&& f)
{
    const uint32_t lane = threadIdx.x & 63u;
    unsigned long long mask = __ballot(has);
    while (mask) { /* uniform: every lane takes part in the broadcasts */
        const int l = __ffsll(mask) - 1;
        mask &= mask - 1ull;
        const uint32_t lo = __shfl(rect_lo, l), hi = __shfl(rect_hi, l);
        const uint64_t pl = __shfl(payload, l);
        const uint32_t tx0 = (lo & 0xFFFFu) >> tshift, ty0 = (lo >> 16) >> tshift;
        const uint32_t tx1 = (hi & 0xFFFFu) >> tshift, ty1 = (hi >> 16) >> tshift;
        const uint32_t ntx = tx1 - tx0 + 1u, nt = ntx * (ty1 - ty0 + 1u);
        for (uint32_t j = lane; j < nt; j += 64u) f(pl, tx0 + j % ntx, ty0 + j / ntx);
    }
}

/* One byte per cluster index: 1 when its 8^3 grid holds a solid voxel.
 * The cull tests it for every (camera, cluster pointer) instead of the
 * cluster's 64 B of bits. */
__global__ __launch_bounds__(256) void vis_solid_k(const uint4* __restrict__ bits, uint32_t n_clusters,
                                                   uint8_t* __restrict__ solid)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_clusters) return;
    const uint4* b4 = bits + 4ull * c;
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) { const uint4 q = b4[i]; any |= q.x | q.y | q.z | q.w; }
    solid[c] = any ? 1u : 0u;
}

/* Occupied cluster pointers (those whose cluster holds a solid voxel), in
 * pointer order: flag[p], its exclusive scan pos[p], then the list and
 * each object's (first entry, count). An object's pointers are one
 * contiguous range, so its occupied ones are one contiguous run. */
__global__ __launch_bounds__(256) void vis_occ_flag_k(const uint32_t* __restrict__ cptrs, const uint8_t* __restrict__ solid,
                                                      uint32_t n_cptr, uint32_t* __restrict__ flag)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n_cptr) flag[p] = solid[cptrs[p]] ? 1u : 0u;
}

__global__ __launch_bounds__(256) void vis_occ_list_k(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                                      uint32_t n_cptr, uint32_t* __restrict__ list)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n_cptr && flag[p]) list[pos[p]] = p;
}

__global__ __launch_bounds__(256) void vis_occ_obj_k(const tgx_object* __restrict__ objects, uint32_t n_objects,
                                                     const uint32_t* __restrict__ pos, const uint32_t* __restrict__ total,
                                                     uint32_t n_cptr, uint2* __restrict__ obj_occ)
{
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_objects) return;
    const tgx_object ob = objects[o];
    const uint32_t n = ob.n_cluster_pointers_per_dim[0] * ob.n_cluster_pointers_per_dim[1] *
                       ob.n_cluster_pointers_per_dim[2];
    const uint32_t a = ob.first_cluster_pointer, e = a + n;
    const uint32_t pa = a < n_cptr ? pos[a] : *total, pe = e < n_cptr ? pos[e] : *total;
    obj_occ[o] = make_uint2(pa, pe - pa);
}

/* Device-wide exclusive scan in three passes: each workgroup scans 4096
 * elements (4 per lane) and leaves its total in sums; one workgroup scans
 * the totals (and writes the grand total); each workgroup then adds its
 * offset, optionally writing the result to a second array as well. */
#define VIS_SCAN_BLOCK 4096u

__device__ static inline uint32_t block_excl_scan(uint32_t s, uint32_t* wsum, uint32_t& block_total)
{
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    uint32_t x = s;
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63u) wsum[w] = x;
    __syncthreads();
    if (w == 0u) {
        uint32_t v = lane < 16u ? wsum[lane] : 0u;
        for (uint32_t d = 1; d < 16u; d <<= 1) {
            const uint32_t y = __shfl_up(v, d);
            if (lane >= d) v += y;
        }
        if (lane < 16u) wsum[lane] = v;
    }
    __syncthreads();
    block_total = wsum[15];
    const uint32_t r = x - s + (w ? wsum[w - 1] : 0u);
    __syncthreads(); /* wsum is reused by the caller's next round */
    return r;
}

__global__ __launch_bounds__(1024) void vis_scan_local_k(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                         uint32_t n, uint32_t* __restrict__ sums)
{
    __shared__ uint32_t wsum[16];
    const uint32_t i0 = blockIdx.x * VIS_SCAN_BLOCK + 4u * threadIdx.x;
    uint32_t v[4], s = 0u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; k++) {
        v[k] = i0 + k < n ? in[i0 + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(s, wsum, tot);
#pragma unroll
    for (uint32_t k = 0; k < 4u; k++) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

/* Scans sums[0..nb) in place. *total gets the grand total; with stats,
 * thread 0 also accumulates the visibility frame statistics: stats[0] +=
 * records of the n_cams cameras, stats[1] += list entries, stats[2] = max
 * list entries. */
__global__ __launch_bounds__(1024) void vis_scan_top_k(uint32_t* __restrict__ sums, uint32_t nb,
                                                       uint32_t* __restrict__ total,
                                                       const uint32_t* __restrict__ rec_count, uint32_t n_cams,
                                                       unsigned long long* __restrict__ stats)
{
    __shared__ uint32_t wsum[16];
    uint32_t carry = 0u;
    for (uint32_t b0 = 0; b0 < nb; b0 += VIS_SCAN_BLOCK) {
        const uint32_t i0 = b0 + 4u * threadIdx.x;
        uint32_t v[4], s = 0u;
#pragma unroll
        for (uint32_t k = 0; k < 4u; k++) {
            v[k] = i0 + k < nb ? sums[i0 + k] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t run = carry + block_excl_scan(s, wsum, tot);
#pragma unroll
        for (uint32_t k = 0; k < 4u; k++) {
            if (i0 + k < nb) sums[i0 + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        if (stats) {
            unsigned long long recs = 0;
            for (uint32_t c = 0; c < n_cams; c++) recs += rec_count[c];
            stats[0] += recs;
            stats[1] += carry;
            if (carry > stats[2]) stats[2] = carry;
        }
    }
}

__global__ __launch_bounds__(1024) void vis_scan_add_k(uint32_t* __restrict__ out, uint32_t n,
                                                       const uint32_t* __restrict__ sums, uint32_t* __restrict__ copy)
{
    const uint32_t i0 = blockIdx.x * VIS_SCAN_BLOCK + 4u * threadIdx.x;
    const uint32_t add = sums[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < 4u; k++) {
        if (i0 + k < n) {
            const uint32_t o = out[i0 + k] + add;
            out[i0 + k] = o;
            if (copy) copy[i0 + k] = o;
        }
    }
}

/* One lane per occupied cluster pointer (one whose cluster holds a solid
 * voxel) of the camera's kept objects (the compact lists of vis_objscan_k;
 * the object's occupied pointers are obj_occ[o].y entries of occ_list from
 * obj_occ[o].x), grid-stride: the lane's object is the last kept
 * one whose offset is not above the lane's index. A wave finds the kept
 * object of its first index with a 64-way search (one load per lane and a
 * ballot per round), then each lane steps over the objects its own index
 * has passed (kept objects have at least one cluster pointer). */
__global__ __launch_bounds__(256) void vis_setup(const tgx_object* __restrict__ objects,
                                                 const uint32_t* __restrict__ cptrs,
                                                 const uint32_t* __restrict__ occ_list,
                                                 const uint2* __restrict__ obj_occ,
                                                 const uint32_t* __restrict__ objlists, uint32_t n_objects,
                                                 const uint32_t* __restrict__ n_kept,
                                                 const uint32_t* __restrict__ kept_total,
                                                 const vis_cam* __restrict__ cams, uint32_t W, uint32_t H,
                                                 uint32_t tiles_x, uint32_t tshift, vis_cluster* __restrict__ recs,
                                                 uint32_t* __restrict__ rec_flags,
                                                 uint32_t* __restrict__ rec_count,
                                                 uint32_t* __restrict__ global_count,
                                                 uint64_t* __restrict__ global_keys,
                                                 uint32_t* __restrict__ tile_count)
{
    const uint32_t ci = blockIdx.y;
    const vis_cam& cam = cams[ci];
    const uint32_t total = kept_total[ci], nk = n_kept[ci];
    const uint32_t* kid = objlists + (size_t)ci * 3u * n_objects + n_objects;
    const uint32_t* koff = kid + n_objects;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = threadIdx.x & 63u;
    /* wave-uniform loop: the rectangle walk below shares tiles across lanes */
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < total; base += stride) {
    const uint32_t t = base + lane;
    /* kept object of the wave's first index: koff[lo] <= base holds
     * throughout (koff[0] = 0), so lane 0's test always passes */
    uint32_t lo = 0u, span = nk;
    while (span > 1u) {
        const uint32_t step = (span + 63u) >> 6;
        const bool p = lane * step < span && koff[lo + lane * step] <= base;
        const uint32_t k = (uint32_t)__popcll(__ballot(p)) - 1u;
        lo += k * step;
        span = min(step, span - k * step);
    }
    vis_cluster rec;
    bool has = false;
    if (t < total) {
        uint32_t k = lo;
        while (k + 1u < nk && koff[k + 1u] <= t) k++;
        const uint32_t oi = kid[k];
        const uint32_t cptr = occ_list[obj_occ[oi].x + (t - koff[k])];
        has = setup_one(objects, cptrs, oi, cptr, cam, W, H, rec);
    }
    bool binned = false;
    if (has) {
        const uint32_t k = atomicAdd(&rec_count[ci], 1u);
        recs[cam.rec_base + k] = rec;
        /* Rectangles whose tiles cover more than VIS_GLOBAL_PIXELS pixels
         * skip the tile lists: every tile of the camera merges them from
         * the global list. */
        const uint32_t ntx = ((rec.rect_hi & 0xFFFFu) >> tshift) - ((rec.rect_lo & 0xFFFFu) >> tshift) + 1u;
        const uint32_t nty = ((rec.rect_hi >> 16) >> tshift) - ((rec.rect_lo >> 16) >> tshift) + 1u;
        if (((uint64_t)ntx * nty << (2u * tshift)) > VIS_GLOBAL_PIXELS) {
            const uint32_t g = atomicAdd(&global_count[ci], 1u);
            global_keys[cam.rec_base + g] = ((uint64_t)rec.depth_lb24 << 32) | (cam.rec_base + k);
        } else {
            binned = true;
            rec_flags[cam.rec_base + k] = 1u;
        }
        if (!binned) rec_flags[cam.rec_base + k] = 0u;
    }
    const uint32_t tb = cam.tile_base;
    for_each_rect_tile(binned, binned ? rec.rect_lo : 0u, binned ? rec.rect_hi : 0u, 0ull, tshift,
                       [&](uint64_t, uint32_t tx, uint32_t ty) { atomicAdd(&tile_count[tb + ty * tiles_x + tx], 1u); });
    }
}

/* ------------------------------------------------------------------------ */
/* Pass 2: tile lists                                                       */
/* ------------------------------------------------------------------------ */

__global__ void vis_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t value)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

/* Camera records passed by value (kernel arguments), so that a call can
 * place them in device memory without a copy from host memory. */
#define VIS_ARG_CAMS 16u
struct vis_cam_pack {
    vis_cam c[VIS_ARG_CAMS];
};

__global__ void vis_cams_k(vis_cam_pack pack, uint32_t n, vis_cam* __restrict__ out)
{
    const uint32_t i = threadIdx.x;
    if (i < n) out[i] = pack.c[i];
}

/* Each list entry is the sort key (depth_lb24 << 32) | record slot. Eight
 * lanes share one record and stride over its tiles, so the latency of the
 * returning atomics is spread over many waves. Entries beyond `capacity`
 * are dropped; vis_traverse detects such tiles and walks every record. */
#define VIS_SCATTER_LANES 8u
__global__ __launch_bounds__(256) void vis_scatter(const vis_cluster* __restrict__ recs,
                                                   const uint32_t* __restrict__ rec_flags,
                                                   const uint32_t* __restrict__ rec_count,
                                                   const vis_cam* __restrict__ cams, uint32_t tiles_x,
                                                   uint32_t tshift, uint32_t* __restrict__ cursor,
                                                   uint64_t* __restrict__ list, uint32_t capacity)
{
    const uint32_t ci = blockIdx.y;
    const uint32_t n_work = rec_count[ci] * VIS_SCATTER_LANES;
    for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid < n_work; gid += gridDim.x * blockDim.x) {
    const uint32_t k = gid / VIS_SCATTER_LANES, sub = gid % VIS_SCATTER_LANES;
    const uint32_t slot = cams[ci].rec_base + k;
    if (rec_flags[slot] == 0u) continue;
    const uint32_t lo = recs[slot].rect_lo, hi = recs[slot].rect_hi;
    const uint64_t key = ((uint64_t)recs[slot].depth_lb24 << 32) | slot;
    const uint32_t tx0 = (lo & 0xFFFFu) >> tshift, ty0 = (lo >> 16) >> tshift;
    const uint32_t tx1 = (hi & 0xFFFFu) >> tshift, ty1 = (hi >> 16) >> tshift;
    const uint32_t ntx = tx1 - tx0 + 1u, nt = ntx * (ty1 - ty0 + 1u);
    const uint32_t tb = cams[ci].tile_base;
    for (uint32_t j = sub; j < nt; j += VIS_SCATTER_LANES) {
        const uint32_t pos = atomicAdd(&cursor[tb + (ty0 + j / ntx) * tiles_x + tx0 + j % ntx], 1u);
        if (pos < capacity) list[pos] = key;
    }
    }
}

/* ------------------------------------------------------------------------ */
/* Pass 3: per-tile traversal                                               */
/* ------------------------------------------------------------------------ */

#define VIS_SORT_CAP 1024u

/* Sorts keys[0..n2) ascending (n2 a power of two) with the workgroup. */
__device__ static inline void lds_bitonic_sort(uint64_t* keys, uint32_t n2)
{
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t a = keys[i], b = keys[p];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { keys[i] = b; keys[p] = a; }
                }
            }
            __syncthreads();
        }
    }
}

/* One workgroup of T x T lanes (T = 1 << kShift) per tile, each wave an
 * 8 x 8 pixel block. The sorted entries are staged VIS_STAGE at a time:
 * the workgroup loads their records, then their clusters' bit words, into
 * LDS with all lanes at once, so the walk over the entries reads no
 * global memory. */
#define VIS_STAGE 64u
template <uint32_t kShift>
__global__ __launch_bounds__(1u << (2u * kShift)) void vis_traverse(const vis_cam* __restrict__ cams, uint32_t W, uint32_t H,
                                                    uint32_t tiles_x, uint32_t tiles_per_cam,
                                                    const uint32_t* __restrict__ tile_off,
                                                    const uint32_t* __restrict__ tile_end,
                                                    const uint64_t* __restrict__ list, uint32_t capacity,
                                                    const uint32_t* __restrict__ rec_count,
                                                    const uint32_t* __restrict__ global_count,
                                                    const uint64_t* __restrict__ global_keys,
                                                    const vis_cluster* __restrict__ recs,
                                                    const uint32_t* __restrict__ bits,
                                                    uint64_t* __restrict__ out)
{
    __shared__ uint64_t keys[VIS_SORT_CAP];
    constexpr uint32_t kT = 1u << kShift, kWx = kT / 8u;
    __shared__ vis_cluster s_rec[VIS_STAGE];
    __shared__ uint32_t s_bits[VIS_STAGE][16];
    __shared__ uint32_t s_more;
    const uint32_t tile = blockIdx.x;
    const uint32_t ci = tile / tiles_per_cam;
    const uint32_t t = tile - ci * tiles_per_cam;
    const uint32_t tx = t % tiles_x, ty = t / tiles_x;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t px = tx * kT + (wave % kWx) * 8u + (lane & 7u);
    const uint32_t py = ty * kT + (wave / kWx) * 8u + (lane >> 3);
    const bool valid = px < W && py < H;

    const vis_cam& cam = cams[ci];
    /* tg_lerp_corner_ray_directions_nn: mix(mix(bl, tl, fy), mix(br, tr, fy), fx) */
    const float fx = ((float)px + 0.5f) / (float)W;
    const float fy = 1.0f - ((float)py + 0.5f) / (float)H;
    const float omy = 1.0f - fy, omx = 1.0f - fx;
    const float lx = cam.bl[0] * omy + cam.tl[0] * fy, ly = cam.bl[1] * omy + cam.tl[1] * fy, lz = cam.bl[2] * omy + cam.tl[2] * fy;
    const float rx = cam.br[0] * omy + cam.tr[0] * fy, ry = cam.br[1] * omy + cam.tr[1] * fy, rz = cam.br[2] * omy + cam.tr[2] * fy;
    const float dwx = lx * omx + rx * fx, dwy = ly * omx + ry * fx, dwz = lz * omx + rz * fx;
    const float far_plane = cam.far_plane;

    uint64_t best = TGX_VISIBILITY_EMPTY;
    /* The tile's own list followed by the camera's global list; if the
     * frame's lists overflowed the list buffer, every record instead. */
    const bool overflow = tile_end[tile] > capacity;
    const uint32_t b = tile_off[tile];
    const uint32_t n_own = overflow ? rec_count[ci] : tile_end[tile] - b;
    const uint32_t n_all = n_own + (overflow ? 0u : global_count[ci]);
    const uint64_t* gk = global_keys + cam.rec_base;
    for (uint32_t cb = 0; cb < n_all; cb += VIS_SORT_CAP) {
        /* Front to back: sort this chunk of the tile's list by depth bound. */
        const uint32_t n = min(VIS_SORT_CAP, n_all - cb);
        uint32_t n2 = 1;
        while (n2 < n) n2 <<= 1;
        for (uint32_t j = threadIdx.x; j < n2; j += blockDim.x) {
            const uint32_t v = cb + j;
            uint64_t kv = ~0ull;
            if (j < n) {
                if (overflow)
                    kv = ((uint64_t)recs[cam.rec_base + v].depth_lb24 << 32) | (cam.rec_base + v);
                else
                    kv = v < n_own ? list[b + v] : gk[v - n_own];
            }
            keys[j] = kv;
        }
        __syncthreads();
        lds_bitonic_sort(keys, n2);
        for (uint32_t sb = 0; sb < n; sb += VIS_STAGE) {
            const uint32_t ns = min(VIS_STAGE, n - sb);
            for (uint32_t q = threadIdx.x; q < ns * 20u; q += blockDim.x) {
                const uint32_t e = q / 20u, f = q - e * 20u;
                ((uint32_t*)&s_rec[e])[f] = ((const uint32_t*)&recs[(uint32_t)keys[sb + e]])[f];
            }
            if (threadIdx.x == 0) s_more = 0u;
            __syncthreads();
            for (uint32_t q = threadIdx.x; q < ns * 16u; q += blockDim.x)
                s_bits[q >> 4][q & 15u] = bits[(size_t)s_rec[q >> 4].cidx * 16u + (q & 15u)];
            __syncthreads();
            bool stop = false;
            for (uint32_t e = 0; e < ns; e++) {
                const uint32_t lb = (uint32_t)(keys[sb + e] >> 32);
                /* Sorted bounds: once no lane can improve, none of the rest can. */
                if (!__any(valid && (uint32_t)(best >> 40) >= lb)) { stop = true; break; }
                const vis_cluster& c = s_rec[e];
                const uint32_t lo = c.rect_lo, hi = c.rect_hi;
                const bool act = valid && px >= (lo & 0xFFFFu) && px <= (hi & 0xFFFFu) && py >= (lo >> 16) &&
                                 py <= (hi >> 16) && (uint32_t)(best >> 40) >= lb;
                if (act) {
                    const uint64_t word = pair_word(c, s_bits[e], dwx, dwy, dwz, far_plane);
                    best = word < best ? word : best;
                }
            }
            /* the next stage only while some wave may still improve */
            if (!stop && lane == 0u) s_more = 1u;
            __syncthreads();
            const bool more = s_more != 0u;
            __syncthreads(); /* s_rec, s_bits and s_more are rewritten by the next stage */
            if (!more) break;
        }
    }
    if (valid) out[(size_t)ci * W * H + (size_t)py * W + px] = best;
}

/* ------------------------------------------------------------------------ */
/* Host side                                                                */
/* ------------------------------------------------------------------------ */

static bool make_cam(const tgx_camera& c, vis_cam& v)
{
    memset(&v, 0, sizeof(v));
    for (int i = 0; i < 3; i++) {
        v.pos[i] = c.camera[i];
        v.bl[i] = c.ray_bl[i]; v.br[i] = c.ray_br[i]; v.tr[i] = c.ray_tr[i]; v.tl[i] = c.ray_tl[i];
    }
    v.far_plane = c.far_plane;
    double m[3][3], quad = 0.0;
    for (int i = 0; i < 3; i++) {
        m[i][0] = c.ray_bl[i];
        m[i][1] = (double)c.ray_br[i] - (double)c.ray_bl[i];
        m[i][2] = (double)c.ray_tl[i] - (double)c.ray_bl[i];
        const double q = (double)c.ray_tr[i] - (double)c.ray_br[i] - (double)c.ray_tl[i] + (double)c.ray_bl[i];
        quad += q * q;
    }
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (!(sqrt(quad) <= 1e-5) || !(fabs(det) > 1e-12)) {
        v.full_screen = 1;
        return true;
    }
    const double inv[9] = {
        (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det,
        (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det, (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det,
        (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det, (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det,
        (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det, (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det,
        (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det};
    for (int i = 0; i < 9; i++) v.inv[i] = (float)inv[i];
    return true;
}

static uint32_t scan_groups(uint64_t n) { return (uint32_t)((n + VIS_SCAN_BLOCK - 1u) / VIS_SCAN_BLOCK); }

/* Exclusive scan of in[0..n) into out (and copy, if not null); sums holds
 * scan_groups(n) u32. The total goes to *total; with stats, the frame
 * statistics are accumulated as well (see vis_scan_top_k). */
static int scan_launch(tgx_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums,
                       uint32_t* copy, uint32_t* total, const uint32_t* rec_count = nullptr, uint32_t n_cams = 0,
                       unsigned long long* stats = nullptr)
{
    const uint32_t nb = scan_groups(n);
    if (nb) hipLaunchKernelGGL(vis_scan_local_k, dim3(nb), dim3(1024), 0, ctx->stream, in, out, n, sums);
    hipLaunchKernelGGL(vis_scan_top_k, dim3(1), dim3(1024), 0, ctx->stream, sums, nb, total, rec_count, n_cams, stats);
    if (nb) hipLaunchKernelGGL(vis_scan_add_k, dim3(nb), dim3(1024), 0, ctx->stream, out, n, sums, copy);
    TGX_LAUNCH_CHECK(ctx, "vis_scan");
    return TGX_OK;
}

/* Device buffers of the occupied-pointer lists: flags, scan and list
 * (n_cptr u32 each), per-object (first, count), scan sums
 * (scan_groups(n_cptr) + 1 u32, the last the number of occupied pointers). */
struct vis_occ_bufs {
    uint32_t* flag;
    uint32_t* pos;
    uint32_t* list;
    uint2*    obj_occ;
    uint32_t* sums;
};

static int occ_launch(tgx_context* ctx, const uint8_t* solid, const vis_occ_bufs& o)
{
    const uint32_t n = ctx->n_cluster_pointers;
    uint32_t* total = o.sums + scan_groups(n);
    hipLaunchKernelGGL(vis_occ_flag_k, dim3((n + 255u) / 256u), dim3(256), 0, ctx->stream,
                       (const uint32_t*)ctx->cluster_pointers.ptr, solid, n, o.flag);
    TGX_LAUNCH_CHECK(ctx, "vis_occ_flag_k");
    int rc;
    if ((rc = scan_launch(ctx, o.flag, o.pos, n, o.sums, nullptr, total))) return rc;
    hipLaunchKernelGGL(vis_occ_list_k, dim3((n + 255u) / 256u), dim3(256), 0, ctx->stream, (const uint32_t*)o.flag,
                       (const uint32_t*)o.pos, n, o.list);
    hipLaunchKernelGGL(vis_occ_obj_k, dim3((ctx->n_objects + 255u) / 256u), dim3(256), 0, ctx->stream,
                       (const tgx_object*)ctx->objects.ptr, ctx->n_objects, (const uint32_t*)o.pos,
                       (const uint32_t*)total, n, o.obj_occ);
    TGX_LAUNCH_CHECK(ctx, "vis_occ_obj_k");
    return TGX_OK;
}

/* Solid flags per cluster and the scene's occupied cluster pointers, once
 * per scene version; the count of occupied pointers comes back to the host
 * (it sizes the per-camera record slots). */
int tgx_vis_solid_refresh(tgx_context* ctx)
{
    if (ctx->vis_solid_version == ctx->scene_version) return TGX_OK;
    const uint32_t n = ctx->n_cluster_pointers;
    if (ctx->n_clusters == 0 || n == 0) {
        ctx->vis_n_occ = 0;
        ctx->vis_solid_version = ctx->scene_version;
        return TGX_OK;
    }
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_solid, ctx->n_clusters, "cluster solid flags"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_occ, 12ull * n, "occupied cluster pointers"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_obj_occ, 8ull * (ctx->n_objects ? ctx->n_objects : 1u),
                                 "occupied pointers per object")))
        return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_scan_sums, 4ull * (scan_groups(n) + 1u), "scan sums"))) return rc;
    hipLaunchKernelGGL(vis_solid_k, dim3((ctx->n_clusters + 255u) / 256u), dim3(256), 0, ctx->stream,
                       (const uint4*)ctx->bits.ptr, ctx->n_clusters, (uint8_t*)ctx->vis_solid.ptr);
    TGX_LAUNCH_CHECK(ctx, "vis_solid_k");
    uint32_t* occ = (uint32_t*)ctx->vis_occ.ptr;
    const vis_occ_bufs o = {occ, occ + n, occ + 2ull * n, (uint2*)ctx->vis_obj_occ.ptr, (uint32_t*)ctx->vis_scan_sums.ptr};
    if ((rc = occ_launch(ctx, (const uint8_t*)ctx->vis_solid.ptr, o))) return rc;
    uint32_t n_occ = 0;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(&n_occ, o.sums + scan_groups(n), 4, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->vis_n_occ = n_occ;
    ctx->vis_solid_version = ctx->scene_version;
    return TGX_OK;
}

/* Record slots per camera: one per occupied cluster pointer (at least one). */
static uint32_t rec_slots(const tgx_context* ctx) { return ctx->vis_n_occ ? ctx->vis_n_occ : 1u; }

/* Tile edge of the visibility passes: 1 << shift pixels. */
static uint32_t vis_tile_shift(const tgx_context* ctx)
{
    return ctx->tuning.vis_tile == 8u ? 3u : ctx->tuning.vis_tile == 32u ? 5u : 4u;
}

static uint32_t tiles_of(uint32_t n, uint32_t shift) { return (n + (1u << shift) - 1u) >> shift; }

/* Device buffers of one visibility batch. */
struct vis_bufs {
    const vis_cam*      cams;
    const uint32_t*     occ_list;   /* occupied cluster pointers            */
    const uint2*        obj_occ;    /* per object: first occupied entry, count */
    uint32_t            slots;      /* record slots per camera              */
    uint32_t*           scan_sums;  /* scan_groups(tiles of the batch) u32  */
    uint32_t*           objcnt;     /* 3 u32 per (camera, object): counts, kept ids, kept offsets */
    vis_cluster*        recs;       /* slots per camera                     */
    uint32_t*           rec_flags;
    uint64_t*           gkeys;
    uint32_t*           counters;   /* 4 n_cams + 8                        */
    uint32_t*           tile_count;
    uint32_t*           tile_off;
    uint64_t*           list;
    uint32_t            capacity;   /* tile-list entries                   */
    unsigned long long* stats;      /* 8 u32, cleared by the caller        */
    bool                timed;      /* profiling scopes allowed            */
};

/* The passes of one batch of n_cams cameras over buffers that hold them. */
static int visibility_launch(tgx_context* ctx, const vis_bufs& b, uint32_t n_cams, uint32_t W, uint32_t H,
                             uint64_t* out_dev)
{
    const uint32_t ts = vis_tile_shift(ctx);
    const uint32_t tiles_x = tiles_of(W, ts), tiles_y = tiles_of(H, ts);
    const uint32_t tiles_per_cam = tiles_x * tiles_y;
    const uint32_t n_tiles = tiles_per_cam * n_cams;
    const uint32_t n_cptr = ctx->n_cluster_pointers;
    /* counters: [0, n_cams) records, [n_cams] list total, [n_cams + 8, 2 n_cams + 8) global counts,
     * [2 n_cams + 8, 3 n_cams + 8) cluster pointers of kept objects, [3 n_cams + 8, 4 n_cams + 8) kept
     * objects */
    uint32_t* d_counters = b.counters;
    uint32_t* d_global_count = d_counters + n_cams + 8;
    uint32_t* d_kept = d_counters + 2 * n_cams + 8;
    uint32_t* d_nkept = d_counters + 3 * n_cams + 8;
    {
        tgx_launch_scope s(ctx, "vis_fill_u32", b.timed);
        hipLaunchKernelGGL(vis_fill_u32, dim3((n_tiles + 255) / 256), dim3(256), 0, ctx->stream, b.tile_count, n_tiles, 0u);
        hipLaunchKernelGGL(vis_fill_u32, dim3((4 * n_cams + 8 + 255) / 256), dim3(256), 0, ctx->stream, d_counters,
                           4 * n_cams + 8, 0u);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_fill_u32");
    {
        tgx_launch_scope s(ctx, "vis_objcull_k", b.timed);
        hipLaunchKernelGGL(vis_objcull_k, dim3((ctx->n_objects + 255u) / 256u, n_cams), dim3(256), 0, ctx->stream,
                           (const tgx_object*)ctx->objects.ptr, ctx->n_objects, b.obj_occ, b.cams, W, H, b.objcnt);
        hipLaunchKernelGGL(vis_objscan_k, dim3(n_cams), dim3(1024), 0, ctx->stream, b.objcnt, ctx->n_objects, d_kept,
                           d_nkept);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_objcull_k");
    {
        tgx_launch_scope s(ctx, "vis_setup", b.timed);
        const uint32_t cap = ctx->tuning.vis_setup_blocks;
        const uint32_t blocks = (b.slots + 255) / 256 < cap ? (b.slots + 255) / 256 : cap;
        hipLaunchKernelGGL(vis_setup, dim3(blocks, n_cams), dim3(256), 0, ctx->stream,
                           (const tgx_object*)ctx->objects.ptr, (const uint32_t*)ctx->cluster_pointers.ptr, b.occ_list,
                           b.obj_occ,
                           (const uint32_t*)b.objcnt, ctx->n_objects, (const uint32_t*)d_nkept, (const uint32_t*)d_kept,
                           b.cams, W, H, tiles_x, ts,
                           b.recs, b.rec_flags, d_counters, d_global_count, b.gkeys, b.tile_count);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_setup");
    {
        /* list offsets; tile_count becomes the scatter cursor */
        tgx_launch_scope s(ctx, "vis_scan", b.timed);
        int rc;
        if ((rc = scan_launch(ctx, b.tile_count, b.tile_off, n_tiles, b.scan_sums, b.tile_count, d_counters + n_cams,
                              d_counters, n_cams, b.stats)))
            return rc;
    }
    {
        tgx_launch_scope s(ctx, "vis_scatter", b.timed);
        const uint64_t threads = (uint64_t)b.slots * VIS_SCATTER_LANES;
        const uint64_t want = (threads + 255) / 256;
        const uint64_t cap = ctx->tuning.vis_setup_blocks;
        hipLaunchKernelGGL(vis_scatter, dim3((uint32_t)(want < cap ? want : cap), n_cams),
                           dim3(256), 0, ctx->stream,
                           b.recs, b.rec_flags, d_counters, b.cams, tiles_x, ts, b.tile_count, b.list, b.capacity);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_scatter");
    {
        /* After the scatter, tile_count[t] is the end of tile t's list. */
        tgx_launch_scope s(ctx, "vis_traverse", b.timed);
        auto* k = ts == 3u ? vis_traverse<3> : ts == 5u ? vis_traverse<5> : vis_traverse<4>;
        hipLaunchKernelGGL(k, dim3(n_tiles), dim3(1u << (2u * ts)), 0, ctx->stream, b.cams, W, H, tiles_x,
                           tiles_per_cam, b.tile_off, b.tile_count, b.list, b.capacity, d_counters, d_global_count,
                           b.gkeys, b.recs, (const uint32_t*)ctx->bits.ptr, out_dev);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_traverse");
    ctx->vis_stats.clusters_total += (uint64_t)n_cptr * n_cams;
    ctx->vis_stats.tiles = tiles_per_cam;
    ctx->vis_stats.tile_size = 1u << ts;
    return TGX_OK;
}

/* Writes every visibility word of a batch of n_cams cameras as empty (a
 * scene without cluster pointers). */
static int visibility_empty(tgx_context* ctx, uint32_t n_cams, uint32_t W, uint32_t H, uint64_t* out_dev, bool timed)
{
    const uint64_t words = 2ull * n_cams * W * H;
    if (words >= (1ull << 32)) return tgx_fail(ctx, TGX_ERR_CAPACITY, "visibility batch too large");
    {
        tgx_launch_scope s(ctx, "vis_fill_u32", timed);
        hipLaunchKernelGGL(vis_fill_u32, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, ctx->stream,
                           (uint32_t*)out_dev, (uint32_t)words, 0xFFFFFFFFu);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_fill_u32");
    return TGX_OK;
}

static bool batch_fits_grid(uint32_t n_cams, uint32_t W, uint32_t H, uint32_t slots, uint32_t ts)
{
    const uint64_t tiles = (uint64_t)tiles_of(W, ts) * tiles_of(H, ts) * n_cams;
    return tiles < (1ull << 31) && (uint64_t)slots * n_cams < (1ull << 32);
}

static int visibility_batch(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t W,
                            uint32_t H, uint64_t* out_dev, const vis_cam* resident)
{
    const uint32_t ts = vis_tile_shift(ctx);
    const uint32_t tiles_per_cam = tiles_of(W, ts) * tiles_of(H, ts);
    const uint32_t n_cptr = ctx->n_cluster_pointers;
    if (n_cptr == 0) return visibility_empty(ctx, n_cams, W, H, out_dev, true);
    int rc;
    if ((rc = tgx_vis_solid_refresh(ctx))) return rc;
    const uint32_t slots = rec_slots(ctx);
    if (!batch_fits_grid(n_cams, W, H, slots, ts)) return tgx_fail(ctx, TGX_ERR_CAPACITY, "visibility batch too large");
    const uint64_t n_slots = (uint64_t)slots * n_cams;
    const uint32_t n_tiles = tiles_per_cam * n_cams;

    std::vector<vis_cam> vc;
    if (!resident) {
        vc.resize(n_cams);
        for (uint32_t i = 0; i < n_cams; i++) {
            make_cam(cams[i], vc[i]);
            vc[i].rec_base = i * slots;
            vc[i].tile_base = i * tiles_per_cam;
        }
        if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_cams, sizeof(vis_cam) * n_cams, "vis cams"))) return rc;
    }
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_clusters, sizeof(vis_cluster) * n_slots, "vis records"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_rec_flags, 4ull * n_slots, "vis record flags"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_global_keys, 8ull * n_slots, "vis global keys"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_counters, 4ull * (4 * n_cams + 8), "vis counters"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_tile_count, 4ull * n_tiles, "tile counts"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_tile_off, 4ull * n_tiles, "tile offsets"))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_objcnt, 12ull * ctx->n_objects * n_cams, "object cull lists")))
        return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_scan_sums, 4ull * (scan_groups(n_tiles) + 1u), "scan sums"))) return rc;
    if (!resident)
        TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->vis_cams.ptr, vc.data(), sizeof(vis_cam) * n_cams, hipMemcpyHostToDevice,
                                          ctx->stream));
    vis_bufs b;
    b.cams = resident ? resident : (const vis_cam*)ctx->vis_cams.ptr;
    b.occ_list = (const uint32_t*)ctx->vis_occ.ptr + 2ull * n_cptr;
    b.obj_occ = (const uint2*)ctx->vis_obj_occ.ptr;
    b.slots = slots;
    b.scan_sums = (uint32_t*)ctx->vis_scan_sums.ptr;
    b.objcnt = (uint32_t*)ctx->vis_objcnt.ptr;
    b.recs = (vis_cluster*)ctx->vis_clusters.ptr;
    b.rec_flags = (uint32_t*)ctx->vis_rec_flags.ptr;
    b.gkeys = (uint64_t*)ctx->vis_global_keys.ptr;
    b.counters = (uint32_t*)ctx->vis_counters.ptr;
    b.tile_count = (uint32_t*)ctx->vis_tile_count.ptr;
    b.tile_off = (uint32_t*)ctx->vis_tile_off.ptr;
    /* The tile lists go into a buffer of fixed capacity (grown between
     * frames from the largest total seen); no host round trip mid-frame. */
    b.list = (uint64_t*)ctx->vis_tile_list.ptr;
    b.capacity = ctx->vis_list_capacity;
    b.stats = (unsigned long long*)ctx->vis_stats_dev.ptr;
    b.timed = true;
    return visibility_launch(ctx, b, n_cams, W, H, out_dev);
}

static int check_args(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene uploaded");
    if (!cams || n_cams == 0 || w == 0 || h == 0 || w > 65535 || h > 65535)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "invalid cameras or resolution (%u cams, %ux%u)", n_cams, w, h);
    for (uint32_t i = 0; i < n_cams; i++)
        if (!(cams[i].far_plane > 0.0f))
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "camera %u: far plane must be > 0", i);
    return TGX_OK;
}

/* Sizes the tile-list buffer. A frame whose lists exceed the capacity is
 * still exact (its overflowing tiles walk every record); the largest total
 * of the previous call, once its readback has landed, grows the buffer. */
static int list_capacity_update(tgx_context* ctx, uint32_t w, uint32_t h)
{
    int rc;
    if (!ctx->vis_stats_event) TGX_HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->vis_stats_event, hipEventDisableTiming));
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_stats_dev, 64, "vis stats"))) return rc;
    uint64_t want = ctx->vis_list_capacity;
    const uint32_t ts = vis_tile_shift(ctx);
    const uint64_t tiles = (uint64_t)tiles_of(w, ts) * tiles_of(h, ts);
    if (want == 0) {
        /* first call: room for 64 entries per tile, two per cluster pointer
         * and at least 2^22, so that large scenes rarely start overflowed */
        want = tiles * 64u > (1u << 22) ? tiles * 64u : (1u << 22);
        if (2ull * ctx->n_cluster_pointers > want) want = 2ull * ctx->n_cluster_pointers;
    }
    if (ctx->vis_list_capacity_fixed) want = ctx->vis_list_capacity_fixed;
    else if (ctx->vis_stats_pending && hipEventQuery(ctx->vis_stats_event) == hipSuccess) {
        ctx->vis_stats_pending = false;
        const uint64_t max_total = ((const unsigned long long*)ctx->host_pinned)[2];
        if (max_total > want) want = max_total + max_total / 4u + 1024u;
    }
    if (want > 0xFFFFFFFFull) want = 0xFFFFFFFFull;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_tile_list, 8ull * want, "tile list"))) return rc;
    ctx->vis_list_capacity = (uint32_t)want;
    return TGX_OK;
}

/* Cameras per batch: bounded so that per-camera record slots stay modest. */
static uint32_t batch_cams(const tgx_context* ctx, uint32_t n_cams)
{
    const uint64_t budget = 1ull << 32; /* 4 GiB of records, flags and keys */
    uint64_t per = (uint64_t)rec_slots(ctx) * (sizeof(vis_cluster) + 12u);
    uint64_t b = per ? budget / per : n_cams;
    if (b < 1) b = 1;
    if (b > 1024) b = 1024; /* grid.y of the per-camera passes */
    return (uint32_t)(b < n_cams ? b : n_cams);
}

int tgx_vis_cams_upload(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                        void** dev_cams)
{
    int rc = check_args(ctx, cams, n_cams, w, h);
    if (rc) return rc;
    if ((rc = tgx_vis_solid_refresh(ctx))) return rc;
    const uint32_t per = batch_cams(ctx, n_cams);
    const uint32_t ts = vis_tile_shift(ctx);
    const uint32_t tiles_per_cam = tiles_of(w, ts) * tiles_of(h, ts);
    std::vector<vis_cam> vc(n_cams);
    for (uint32_t i = 0; i < n_cams; i++) {
        const uint32_t j = i % per; /* index within its batch */
        make_cam(cams[i], vc[i]);
        vc[i].rec_base = j * rec_slots(ctx);
        vc[i].tile_base = j * tiles_per_cam;
    }
    void* d = nullptr;
    TGX_HIP_CHECK(ctx, hipMalloc(&d, sizeof(vis_cam) * n_cams));
    const hipError_t e = hipMemcpy(d, vc.data(), sizeof(vis_cam) * n_cams, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return tgx_fail(ctx, TGX_ERR_HIP, "camera upload: %s", hipGetErrorString(e));
    }
    *dev_cams = d;
    return TGX_OK;
}

int tgx_visibility_record(tgx_context* ctx, const void* dev_cams, const tgx_camera* cams, uint32_t n_cams, uint32_t w,
                          uint32_t h, uint64_t* out_dev)
{
    int rc = check_args(ctx, cams, n_cams, w, h);
    if (rc) return rc;
    if (!ctx->vis_tile_list.ptr || !ctx->vis_stats_dev.ptr)
        return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "record a frame only after one tgx_visibility_device call");
    hipLaunchKernelGGL(vis_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)ctx->vis_stats_dev.ptr, 8u, 0u);
    TGX_LAUNCH_CHECK(ctx, "vis_fill_u32");
    if ((rc = tgx_vis_solid_refresh(ctx))) return rc;
    const uint32_t per = batch_cams(ctx, n_cams);
    const vis_cam* rc_cams = (const vis_cam*)dev_cams;
    for (uint32_t c0 = 0; c0 < n_cams; c0 += per) {
        const uint32_t n = (n_cams - c0) < per ? (n_cams - c0) : per;
        if ((rc = visibility_batch(ctx, cams + c0, n, w, h, out_dev + (size_t)c0 * w * h, rc_cams + c0))) return rc;
    }
    return TGX_OK;
}

extern "C" {

int tgx_visibility_device(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w,
                          uint32_t h, uint64_t* out_dev)
{
    int rc = check_args(ctx, cams, n_cams, w, h);
    if (rc) return rc;
    if (!out_dev) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "out_dev is NULL");
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    memset(&ctx->vis_stats, 0, sizeof(ctx->vis_stats));
    if ((rc = list_capacity_update(ctx, w, h))) return rc;
    {
        tgx_launch_scope s(ctx, "vis_fill_u32");
        hipLaunchKernelGGL(vis_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)ctx->vis_stats_dev.ptr, 8u, 0u);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_fill_u32");
    if ((rc = tgx_vis_solid_refresh(ctx))) return rc;
    const uint32_t per = batch_cams(ctx, n_cams);
    for (uint32_t c0 = 0; c0 < n_cams; c0 += per) {
        const uint32_t n = (n_cams - c0) < per ? (n_cams - c0) : per;
        rc = visibility_batch(ctx, cams + c0, n, w, h, out_dev + (size_t)c0 * w * h, nullptr);
        if (rc) return rc;
    }
    /* Frame statistics travel back asynchronously; the next call (or
     * tgx_visibility_last_stats) reads them. */
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->host_pinned, ctx->vis_stats_dev.ptr, 32, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipEventRecord(ctx->vis_stats_event, ctx->stream));
    ctx->vis_stats_pending = true;
    return TGX_OK;
}

int tgx_visibility(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                   uint64_t* out_host)
{
    int rc = check_args(ctx, cams, n_cams, w, h);
    if (rc) return rc;
    if (!out_host) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "out_host is NULL");
    const size_t bytes = (size_t)n_cams * w * h * sizeof(uint64_t);
    if ((rc = tgx_buffer_reserve(ctx, &ctx->vis_out, bytes, "visibility output"))) return rc;
    if ((rc = tgx_visibility_device(ctx, cams, n_cams, w, h, (uint64_t*)ctx->vis_out.ptr))) return rc;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(out_host, ctx->vis_out.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_visibility_clear_device(tgx_context* ctx, uint64_t* vis_dev, uint64_t n_pixels)
{
    if (!ctx || !vis_dev) return TGX_ERR_INVALID_ARG;
    if (n_pixels == 0) return TGX_OK;
    if (2 * n_pixels >= (1ull << 32)) return tgx_fail(ctx, TGX_ERR_CAPACITY, "clear of %llu pixels", (unsigned long long)n_pixels);
    const uint32_t n = (uint32_t)(2 * n_pixels);
    {
        tgx_launch_scope s(ctx, "vis_fill_u32");
        hipLaunchKernelGGL(vis_fill_u32, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, (uint32_t*)vis_dev, n, 0xFFFFFFFFu);
    }
    TGX_LAUNCH_CHECK(ctx, "vis_fill_u32");
    return TGX_OK;
}

int tgx_visibility_set_list_capacity(tgx_context* ctx, uint32_t entries)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->vis_list_capacity_fixed = entries;
    if (entries == 0) ctx->vis_list_capacity = 0;
    return TGX_OK;
}

int tgx_visibility_last_stats(tgx_context* ctx, tgx_visibility_stats* out)
{
    if (!ctx || !out) return TGX_ERR_INVALID_ARG;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const unsigned long long* hs = (const unsigned long long*)ctx->host_pinned;
    *out = ctx->vis_stats;
    out->clusters_binned = hs[0];
    out->tile_entries = hs[1];
    return TGX_OK;
}

} /* extern "C" */

/* ------------------------------------------------------------------------ */
/* Visibility without allocation                                            */
/* ------------------------------------------------------------------------ */

/* Every work buffer of tgx_visibility_device_noalloc lives in this array
 * (one copy per GPU, held until the process ends). Host code never passes
 * the array's name to a kernel: its device address comes from
 * hipGetSymbolAddress, once per device, into vis_scratch_addr. */
static __device__ __align__(256) unsigned char scratch[TGX_VIS_NOALLOC_SCRATCH_BYTES];

#define VIS_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void* vis_scratch_addr[VIS_MAX_DEVICES];

int tgx_vis_scratch_init(tgx_context* ctx)
{
    if (ctx->device < 0 || ctx->device >= VIS_MAX_DEVICES)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "device %d beyond the scratch table", ctx->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (vis_scratch_addr[ctx->device]) return TGX_OK;
    void* p = nullptr;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    TGX_HIP_CHECK(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return tgx_fail(ctx, TGX_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    vis_scratch_addr[ctx->device] = p;
    return TGX_OK;
}

static size_t align256(size_t x) { return (x + 255u) & ~(size_t)255u; }

/* Scratch layout of one batch of n_cams cameras; returns the bytes before
 * the tile list. */
static uint32_t noalloc_sums(const tgx_context* ctx, uint32_t n_cams, uint32_t tiles_per_cam)
{
    const uint32_t a = scan_groups(ctx->n_cluster_pointers) + 1u, t = scan_groups((uint64_t)tiles_per_cam * n_cams) + 1u;
    return a > t ? a : t;
}

static size_t noalloc_fixed_bytes(const tgx_context* ctx, uint32_t n_cams, uint32_t tiles_per_cam)
{
    const uint64_t slots = (uint64_t)ctx->n_cluster_pointers * n_cams;
    return align256(sizeof(vis_cam) * n_cams) + align256(ctx->n_clusters) +
           align256(12ull * ctx->n_cluster_pointers) + align256(8ull * ctx->n_objects) +
           align256(4ull * noalloc_sums(ctx, n_cams, tiles_per_cam)) +
           align256(12ull * ctx->n_objects * n_cams) + align256(sizeof(vis_cluster) * slots) + align256(4 * slots) +
           align256(8 * slots) + align256(4ull * (4 * n_cams + 8)) + 2 * align256(4ull * tiles_per_cam * n_cams) +
           align256(64);
}

extern "C" {

int tgx_visibility_device_noalloc(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w,
                                  uint32_t h, uint64_t* out_dev)
{
    int rc = check_args(ctx, cams, n_cams, w, h);
    if (rc) return rc;
    if (!out_dev) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "out_dev is NULL");
    if (ctx->device < 0 || ctx->device >= VIS_MAX_DEVICES || !vis_scratch_addr[ctx->device])
        return tgx_fail(ctx, TGX_ERR_HIP, "no scratch address for device %d", ctx->device);
    unsigned char* base = (unsigned char*)vis_scratch_addr[ctx->device];
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    memset(&ctx->vis_stats, 0, sizeof(ctx->vis_stats));
    const uint32_t ts = vis_tile_shift(ctx);
    const uint32_t tiles_per_cam = tiles_of(w, ts) * tiles_of(h, ts);
    const uint32_t n_cptr = ctx->n_cluster_pointers;
    /* Cameras per batch: as many as fit with at least 64 tile-list entries
     * per tile (a frame whose lists overflow stays exact, only slower). */
    uint32_t per = n_cams < VIS_ARG_CAMS ? n_cams : VIS_ARG_CAMS;
    while (per > 0 && noalloc_fixed_bytes(ctx, per, tiles_per_cam) + 8ull * 64u * tiles_per_cam * per >
                          (size_t)TGX_VIS_NOALLOC_SCRATCH_BYTES)
        per--;
    if (per == 0)
        return tgx_fail(ctx, TGX_ERR_CAPACITY,
                        "one camera at %ux%u over %u cluster pointers needs %zu bytes of scratch; the maximum is %llu",
                        w, h, n_cptr, noalloc_fixed_bytes(ctx, 1, tiles_per_cam) + 8ull * 64u * tiles_per_cam,
                        (unsigned long long)TGX_VIS_NOALLOC_SCRATCH_BYTES);
    if (n_cptr && !batch_fits_grid(per, w, h, n_cptr, ts))
        return tgx_fail(ctx, TGX_ERR_CAPACITY, "visibility batch too large");
    for (uint32_t c0 = 0; c0 < n_cams; c0 += per) {
        const uint32_t n = (n_cams - c0) < per ? (n_cams - c0) : per;
        uint64_t* out = out_dev + (size_t)c0 * w * h;
        if (n_cptr == 0) {
            if ((rc = visibility_empty(ctx, n, w, h, out, false))) return rc;
            continue;
        }
        const uint64_t slots = (uint64_t)n_cptr * n;
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            unsigned char* p = base + off;
            off += align256(bytes);
            return (void*)p;
        };
        vis_cam_pack pack;
        memset(&pack, 0, sizeof(pack));
        for (uint32_t i = 0; i < n; i++) {
            make_cam(cams[c0 + i], pack.c[i]);
            pack.c[i].rec_base = i * n_cptr;
            pack.c[i].tile_base = i * tiles_per_cam;
        }
        vis_bufs b;
        vis_cam* d_cams = (vis_cam*)carve(sizeof(vis_cam) * n);
        uint8_t* d_solid = (uint8_t*)carve(ctx->n_clusters);
        uint32_t* d_occ = (uint32_t*)carve(12ull * n_cptr);
        uint2* d_obj_occ = (uint2*)carve(8ull * ctx->n_objects);
        uint32_t* d_sums = (uint32_t*)carve(4ull * noalloc_sums(ctx, n, tiles_per_cam));
        const vis_occ_bufs ob = {d_occ, d_occ + n_cptr, d_occ + 2ull * n_cptr, d_obj_occ, d_sums};
        b.cams = d_cams;
        b.occ_list = d_occ + 2ull * n_cptr;
        b.obj_occ = d_obj_occ;
        b.slots = n_cptr;
        b.scan_sums = d_sums;
        b.objcnt = (uint32_t*)carve(12ull * ctx->n_objects * n);
        b.recs = (vis_cluster*)carve(sizeof(vis_cluster) * slots);
        b.rec_flags = (uint32_t*)carve(4 * slots);
        b.gkeys = (uint64_t*)carve(8 * slots);
        b.counters = (uint32_t*)carve(4ull * (4 * n + 8));
        b.tile_count = (uint32_t*)carve(4ull * tiles_per_cam * n);
        b.tile_off = (uint32_t*)carve(4ull * tiles_per_cam * n);
        b.stats = (unsigned long long*)carve(64);
        const size_t left = (size_t)TGX_VIS_NOALLOC_SCRATCH_BYTES - off;
        const size_t entries = left / 8u;
        b.list = (uint64_t*)(base + off);
        b.capacity = entries > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)entries;
        b.timed = false;
        hipLaunchKernelGGL(vis_cams_k, dim3(1), dim3(64), 0, ctx->stream, pack, n, d_cams);
        TGX_LAUNCH_CHECK(ctx, "vis_cams_k");
        /* The solid flags and the occupied-pointer lists are derived from
         * the scene on every call: the array keeps whatever the last call on
         * this GPU left. */
        hipLaunchKernelGGL(vis_solid_k, dim3((ctx->n_clusters + 255u) / 256u), dim3(256), 0, ctx->stream,
                           (const uint4*)ctx->bits.ptr, ctx->n_clusters, d_solid);
        TGX_LAUNCH_CHECK(ctx, "vis_solid_k");
        if ((rc = occ_launch(ctx, d_solid, ob))) return rc;
        hipLaunchKernelGGL(vis_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)b.stats, 16u, 0u);
        TGX_LAUNCH_CHECK(ctx, "vis_fill_u32");
        if ((rc = visibility_launch(ctx, b, n, w, h, out))) return rc;
    }
    return TGX_OK;
}

} /* extern "C" */
