/*
 * tgx_oracle_svo.c — CPU oracle of tg's sparse voxel octree: build
 * (tg_svo_create and helpers, tg_sparse_voxel_octree.c:23-542), the
 * box/OBB test (tg_physics.c:226-392) and the traversal (tg_svo_traverse,
 * tg_sparse_voxel_octree.c:558-740, with tg_amanatides_woo,
 * tg_amanatides_woo.c:3-114). TEST INFRASTRUCTURE ONLY.
 *
 * Same structure as the reference: depth-first recursion, children of a
 * node allocated contiguously, a cluster-pointer list per node filtered by
 * the OBB test at every level. Differences, each shared with the GPU
 * build and traversal:
 *   - depth is a parameter (the reference fixes 1024^3, :11-16);
 *   - child pointers are absolute u32, cluster lists unbounded (:408-413,
 *     :309 assert at 2^16 nodes / 64 clusters);
 *   - node boxes are taken relative to the root centre during traversal
 *     (the reference subtracts the centre from the origin but keeps world
 *     bounds, :570-594, which agrees only for a centre at 0);
 *   - the leaf AABB uses all eight cluster corners unless
 *     TGX_SVO_FLAG_REFERENCE_CORNER_QUIRK is set (:183,191 use [2],[2]);
 *   - trimmed bounds are clamped into [0, 32] before the (u32)
 *     conversion and the block DDA clamps the entry voxel on both sides
 *     (the reference converts negative floats and indexes voxel -1: UB);
 *   - a per-voxel material grid (first cluster in list order) is added;
 *   - traversal stops after TGX_SVO_MAX_STEPS iterations or once the
 *     distance travelled exceeds the range.
 *
 * Parity pins (tests/test_oracle_svo.py): unrotated grid-aligned objects
 * whose SVO bits must equal their solid voxels exactly; invariants of the
 * tree; hit voxels against a dense-grid 3-D DDA; the quirk flag changing
 * bits only where corner 3 is extremal.
 */
#include "tgx_oracle.h"

#include <omp.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F32_MAX 3.402823466e+38f
#define F32_MIN (-F32_MAX)
#define F32_EPSILON 1.192092896e-07f

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } v4;
typedef struct { float m[16]; } m4;
#define E(mat, r, c) ((mat).m[4 * (c) + (r)])

static float fmin2(float a, float b) { return a < b ? a : b; }
static float fmax2(float a, float b) { return a > b ? a : b; }
static v3 v3_min(v3 a, v3 b) { v3 r = {fmin2(a.x, b.x), fmin2(a.y, b.y), fmin2(a.z, b.z)}; return r; }
static v3 v3_max(v3 a, v3 b) { v3 r = {fmax2(a.x, b.x), fmax2(a.y, b.y), fmax2(a.z, b.z)}; return r; }
static v3 v3_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 v3_add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static v3 v3_mulf(v3 a, float f) { v3 r = {a.x * f, a.y * f, a.z * f}; return r; }
static v3 v3_neg(v3 a) { v3 r = {-a.x, -a.y, -a.z}; return r; }
static float v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float v3_mag(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static v3 v3_normalized(v3 a)
{
    const float mag = v3_mag(a);
    v3 r = {a.x / mag, a.y / mag, a.z / mag};
    return r;
}
static v3 v3_divf(v3 a, float f) { v3 r = {a.x / f, a.y / f, a.z / f}; return r; }
static v3 v3_div_zero_check(v3 a, v3 b, v3 alt)
{
    v3 r;
    r.x = (b.x == 0.0f) ? alt.x : (a.x / b.x);
    r.y = (b.y == 0.0f) ? alt.y : (a.y / b.y);
    r.z = (b.z == 0.0f) ? alt.z : (a.z / b.z);
    return r;
}
static float v3_min_elem(v3 a) { return fmin2(fmin2(a.x, a.y), a.z); }

static m4 m4_identity(void)
{
    m4 r;
    memset(&r, 0, sizeof(r));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
static m4 m4_translate(v3 v)
{
    m4 r = m4_identity();
    E(r, 0, 3) = v.x; E(r, 1, 3) = v.y; E(r, 2, 3) = v.z;
    return r;
}
static m4 m4_mul(m4 a, m4 b)
{
    float o[16];
    tgxo_m4_mul(a.m, b.m, o);
    m4 r;
    memcpy(r.m, o, sizeof(o));
    return r;
}
static m4 m4_inverse(m4 a)
{
    m4 r;
    tgxo_m4_inverse(a.m, r.m);
    return r;
}
static v3 m4_mulv4_xyz(m4 m, v4 v)
{
    v3 r;
    r.x = v.x * E(m,0,0) + v.y * E(m,0,1) + v.z * E(m,0,2) + v.w * E(m,0,3);
    r.y = v.x * E(m,1,0) + v.y * E(m,1,1) + v.z * E(m,1,2) + v.w * E(m,1,3);
    r.z = v.x * E(m,2,0) + v.y * E(m,2,1) + v.z * E(m,2,2) + v.w * E(m,2,3);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Cluster transforms (tg_sparse_voxel_octree.c:23-100)                      */
/* ------------------------------------------------------------------------ */

static void rel_coords(const tgx_object* o, uint32_t rel, uint32_t* rx, uint32_t* ry, uint32_t* rz)
{
    const uint32_t nx = o->n_cluster_pointers_per_dim[0], ny = o->n_cluster_pointers_per_dim[1];
    *rx = rel % nx;
    *ry = (rel / nx) % ny;
    *rz = rel / (nx * ny);
}

static m4 ws2ms_of(const tgx_object* o, uint32_t rel)
{
    m4 r;
    tgxo_world_to_cluster(o, rel, r.m);
    return r;
}

/* T(t) * R * T(-half_object) * T(offset), left-associated (:92-97). */
static m4 ms2ws_of(const tgx_object* o, uint32_t rel)
{
    uint32_t rx, ry, rz;
    rel_coords(o, rel, &rx, &ry, &rz);
    const v3 off = {(float)(rx * 8u), (float)(ry * 8u), (float)(rz * 8u)};
    const float h = 4.0f;
    const v3 half = {(float)o->n_cluster_pointers_per_dim[0] * h, (float)o->n_cluster_pointers_per_dim[1] * h,
                     (float)o->n_cluster_pointers_per_dim[2] * h};
    m4 rot;
    memcpy(rot.m, o->rotation, sizeof(rot.m));
    const v3 t = {o->translation[0], o->translation[1], o->translation[2]};
    const m4 m1 = m4_translate(off);
    const m4 m2 = m4_translate(v3_neg(half));
    const m4 m3 = rot;
    const m4 m4_ = m4_translate(t);
    return m4_mul(m4_mul(m4_mul(m4_, m3), m2), m1);
}

/* ------------------------------------------------------------------------ */
/* tg_intersect_aabb_obb_ignore_contact (tg_physics.c:226-392)               */
/* ------------------------------------------------------------------------ */

int tgxo_aabb_obb_ignore_contact(const float mn[3], const float mx[3], const float corners[24])
{
    v3 c[8];
    for (int i = 0; i < 8; i++) { c[i].x = corners[3 * i]; c[i].y = corners[3 * i + 1]; c[i].z = corners[3 * i + 2]; }
    int sep;
    sep = 1; for (int i = 0; i < 8 && sep; i++) sep = c[i].x <= mn[0]; if (sep) return 0;
    sep = 1; for (int i = 0; i < 8 && sep; i++) sep = c[i].x >= mx[0]; if (sep) return 0;
    sep = 1; for (int i = 0; i < 8 && sep; i++) sep = c[i].y <= mn[1]; if (sep) return 0;
    sep = 1; for (int i = 0; i < 8 && sep; i++) sep = c[i].y >= mx[1]; if (sep) return 0;
    sep = 1; for (int i = 0; i < 8 && sep; i++) sep = c[i].z <= mn[2]; if (sep) return 0;
    sep = 1; for (int i = 0; i < 8 && sep; i++) sep = c[i].z >= mx[2]; if (sep) return 0;

    const v3 nxp = v3_normalized(v3_sub(c[1], c[0]));
    const v3 nyp = v3_normalized(v3_sub(c[2], c[0]));
    const v3 nzp = v3_normalized(v3_sub(c[4], c[0]));
    v3 nrm[6];
    nrm[0] = v3_neg(nxp); nrm[1] = nxp;
    nrm[2] = v3_neg(nyp); nrm[3] = nyp;
    nrm[4] = v3_neg(nzp); nrm[5] = nzp;
    const float l0 = v3_mag(c[0]);
    const float l1 = v3_mag(c[7]);
    const v3 zero = {0.0f, 0.0f, 0.0f};
    const v3 n0 = l0 == 0.0f ? zero : v3_divf(c[0], l0);
    const v3 n1 = l1 == 0.0f ? zero : v3_divf(c[7], l1);
    float dist[6];
    dist[0] = l0 * v3_dot(n0, nrm[0]);
    dist[1] = l1 * v3_dot(n1, nrm[1]);
    dist[2] = l0 * v3_dot(n0, nrm[2]);
    dist[3] = l1 * v3_dot(n1, nrm[3]);
    dist[4] = l0 * v3_dot(n0, nrm[4]);
    dist[5] = l1 * v3_dot(n1, nrm[5]);
    v3 pts[8];
    for (int i = 0; i < 8; i++) {
        pts[i].x = (i & 1) ? mx[0] : mn[0];
        pts[i].y = (i & 2) ? mx[1] : mn[1];
        pts[i].z = (i & 4) ? mx[2] : mn[2];
    }
    for (int p = 0; p < 6; p++) {
        sep = 1;
        for (int i = 0; i < 8 && sep; i++) sep = (v3_dot(nrm[p], pts[i]) - dist[p]) >= 0.0f;
        if (sep) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Build                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct onode {
    uint32_t first_child;
    uint32_t valid, leaf;
    uint32_t level, coord[3];
    uint32_t leaf_idx;
} onode;

struct tgxo_svo {
    tgx_svo_info info;
    onode*    nodes;
    uint32_t  n_nodes, cap_nodes;
    uint32_t* bits;       /* 1024 per leaf  */
    uint16_t* mats;       /* 32768 per leaf */
    uint32_t* leaf_n;
    uint32_t* leaf_node;  /* node index of each leaf */
    uint32_t  n_leaves, cap_leaves;
    uint32_t* list_off;   /* start of each leaf's list (n_leaves + 1 at export) */
    uint32_t* list;
    uint32_t  n_list, cap_list;
    const tgxo_scene* scene;
    uint32_t  flags;
    int       failed;
};

/* Line recorder of the algorithmic-bytes count (tgxo_gi_alg_bytes): one
 * bit per 64-B line of each array the GI frame reads, in the reference's
 * layouts. NULL in every parity path. */
typedef struct line_rec {
    uint64_t* node;   /* nodes, u32 each in build order (tg_sparse_voxel_octree.h:6-11) */
    uint64_t* leafd;  /* leaf data, 260 B each: n and the cluster list (:13-17) */
    uint64_t* bits;   /* voxel pool, 4096 B per leaf (:121-126) */
    uint64_t* mats;   /* leaf materials, 2 B per voxel */
    uint64_t* luts;   /* LUT entries of lit bounce hits, 16 B */
    uint64_t* lutsp;  /* LUT entries of primary surfaces */
    uint64_t* cptr;   /* cluster pointers, 4 B */
    uint64_t* c2o;    /* cluster -> object, 4 B */
    uint64_t* obj;    /* object records, 96 B */
    uint64_t* smat;   /* scene materials, 1 B per voxel */
} line_rec;

static void mark_line(uint64_t* bm, uint64_t line)
{
    if (!bm) return;
    const uint64_t bit = 1ull << (line & 63u);
    /* read first: most lines are marked already, and a shared read keeps
     * the cache line from bouncing between threads */
    if (!(__atomic_load_n(&bm[line >> 6], __ATOMIC_RELAXED) & bit))
        __atomic_fetch_or(&bm[line >> 6], bit, __ATOMIC_RELAXED);
}

static int grow(void** p, uint32_t* cap, uint32_t need, size_t elem)
{
    if (need <= *cap) return 1;
    uint32_t nc = *cap ? *cap : 64;
    while (nc < need) nc *= 2;
    void* q = realloc(*p, (size_t)nc * elem);
    if (!q) return 0;
    *p = q;
    *cap = nc;
    return 1;
}

static uint32_t new_nodes(tgxo_svo* s, uint32_t n)
{
    if (!grow((void**)&s->nodes, &s->cap_nodes, s->n_nodes + n, sizeof(onode))) { s->failed = 1; return 0; }
    const uint32_t first = s->n_nodes;
    memset(&s->nodes[first], 0, sizeof(onode) * n);
    for (uint32_t i = 0; i < n; i++) { s->nodes[first + i].first_child = 0xFFFFFFFFu; s->nodes[first + i].leaf_idx = 0xFFFFFFFFu; }
    s->n_nodes += n;
    return first;
}

static void construct_leaf(tgxo_svo* s, uint32_t node, v3 pmin, v3 pmax, uint32_t n_cp, const uint32_t* cps)
{
    const tgxo_scene* sc = s->scene;
    if (!grow((void**)&s->leaf_n, &s->cap_leaves, s->n_leaves + 1, sizeof(uint32_t))) { s->failed = 1; return; }
    uint32_t cap = s->cap_leaves;
    /* keep the parallel per-leaf arrays at the same capacity */
    s->bits = (uint32_t*)realloc(s->bits, (size_t)cap * TGX_SVO_BLOCK_WORDS * sizeof(uint32_t));
    s->mats = (uint16_t*)realloc(s->mats, (size_t)cap * TGX_SVO_BLOCK_VOXELS * sizeof(uint16_t));
    s->leaf_node = (uint32_t*)realloc(s->leaf_node, (size_t)cap * sizeof(uint32_t));
    s->list_off = (uint32_t*)realloc(s->list_off, (size_t)(cap + 1) * sizeof(uint32_t));
    if (!s->bits || !s->mats || !s->leaf_node || !s->list_off) { s->failed = 1; return; }
    const uint32_t li = s->n_leaves++;
    s->nodes[node].leaf_idx = li;
    s->leaf_node[li] = node;
    s->leaf_n[li] = 0;
    s->list_off[li] = s->n_list;
    uint32_t* blk = &s->bits[(size_t)li * TGX_SVO_BLOCK_WORDS];
    uint16_t* mat = &s->mats[(size_t)li * TGX_SVO_BLOCK_VOXELS];
    memset(blk, 0, TGX_SVO_BLOCK_WORDS * sizeof(uint32_t));
    memset(mat, 0, TGX_SVO_BLOCK_VOXELS * sizeof(uint16_t));

    const v3 pext = v3_sub(pmax, pmin);
    const float e8 = 8.0f;
    const v4 ccs[8] = {{0, 0, 0, 1}, {e8, 0, 0, 1}, {0, e8, 0, 1}, {e8, e8, 0, 1},
                       {0, 0, e8, 1}, {e8, 0, e8, 1}, {0, e8, e8, 1}, {e8, e8, e8, 1}};
    v4 bws[8];
    for (int i = 0; i < 8; i++) {
        bws[i].x = (i & 1) ? pmax.x : pmin.x;
        bws[i].y = (i & 2) ? pmax.y : pmin.y;
        bws[i].z = (i & 4) ? pmax.z : pmin.z;
        bws[i].w = 1.0f;
    }
    uint32_t last_cidx = 0xFFFFFFFFu;
    for (uint32_t k = 0; k < n_cp; k++) {
        const uint32_t cptr = cps[k];
        const uint32_t cidx = sc->cluster_pointers[cptr];
        const uint32_t obj = sc->cluster_idx_to_object[cidx];
        const tgx_object* o = &sc->objects[obj];
        const uint32_t rel = cptr - o->first_cluster_pointer;
        const m4 cs2ws = ms2ws_of(o, rel);
        v3 cw[8];
        for (int i = 0; i < 8; i++) cw[i] = m4_mulv4_xyz(cs2ws, ccs[i]);
        const int c3 = (s->flags & TGX_SVO_FLAG_REFERENCE_CORNER_QUIRK) ? 2 : 3;
        const v3 min_c = v3_min(v3_min(v3_min(cw[0], cw[1]), v3_min(cw[2], cw[c3])),
                                v3_min(v3_min(cw[4], cw[6]), v3_min(cw[5], cw[7])));
        const v3 max_c = v3_max(v3_max(v3_max(cw[0], cw[1]), v3_max(cw[2], cw[c3])),
                                v3_max(v3_max(cw[4], cw[6]), v3_max(cw[5], cw[7])));
        const v3 fmin_c = {floorf(min_c.x), floorf(min_c.y), floorf(min_c.z)};
        const v3 cmax_c = {ceilf(max_c.x), ceilf(max_c.y), ceilf(max_c.z)};
        const m4 ws2cs = m4_inverse(cs2ws);
        v3 bc[8];
        for (int i = 0; i < 8; i++) bc[i] = m4_mulv4_xyz(ws2cs, bws[i]);
        const v3 tmin = v3_max(pmin, fmin_c);
        const v3 tmax = v3_min(pmax, cmax_c);
        float lo[3] = {tmin.x - pmin.x, tmin.y - pmin.y, tmin.z - pmin.z};
        float hi[3] = {tmax.x - pmin.x, tmax.y - pmin.y, tmax.z - pmin.z};
        uint32_t ilo[3], ihi[3];
        for (int a = 0; a < 3; a++) {
            ilo[a] = (uint32_t)fmin2(32.0f, fmax2(0.0f, lo[a]));
            ihi[a] = (uint32_t)fmin2(32.0f, fmax2(0.0f, hi[a]));
        }
        const uint64_t coff = (uint64_t)cidx * 512u;
        const uint16_t lut = (uint16_t)((o->lut_idx & 0xFFu) << 8);
        int any = 0;
        for (uint32_t z = ilo[2]; z < ihi[2]; z++) {
            const float tz = ((float)z + 0.5f) / pext.z;
            const float omtz = 1.0f - tz;
            const float pz0x = omtz * bc[0].x + tz * bc[4].x, pz0y = omtz * bc[0].y + tz * bc[4].y, pz0z = omtz * bc[0].z + tz * bc[4].z;
            const float pz1x = omtz * bc[1].x + tz * bc[5].x, pz1y = omtz * bc[1].y + tz * bc[5].y, pz1z = omtz * bc[1].z + tz * bc[5].z;
            const float pz2x = omtz * bc[2].x + tz * bc[6].x, pz2y = omtz * bc[2].y + tz * bc[6].y, pz2z = omtz * bc[2].z + tz * bc[6].z;
            const float pz3x = omtz * bc[3].x + tz * bc[7].x, pz3y = omtz * bc[3].y + tz * bc[7].y, pz3z = omtz * bc[3].z + tz * bc[7].z;
            for (uint32_t y = ilo[1]; y < ihi[1]; y++) {
                const float ty = ((float)y + 0.5f) / pext.y;
                const float omty = 1.0f - ty;
                const float py0x = omty * pz0x + ty * pz2x, py0y = omty * pz0y + ty * pz2y, py0z = omty * pz0z + ty * pz2z;
                const float py1x = omty * pz1x + ty * pz3x, py1y = omty * pz1y + ty * pz3y, py1z = omty * pz1z + ty * pz3z;
                for (uint32_t x = ilo[0]; x < ihi[0]; x++) {
                    const float tx = ((float)x + 0.5f) / pext.x;
                    const float omtx = 1.0f - tx;
                    const float cx = omtx * py0x + tx * py1x;
                    if (cx < 0.0f || cx >= 8.0f) continue;
                    const float cy = omtx * py0y + tx * py1y;
                    if (cy < 0.0f || cy >= 8.0f) continue;
                    const float cz = omtx * py0z + tx * py1z;
                    if (cz < 0.0f || cz >= 8.0f) continue;
                    const uint32_t v = 64u * (uint32_t)cz + 8u * (uint32_t)cy + (uint32_t)cx;
                    const uint64_t cv = coff + v;
                    if (!(sc->bits[cv / 32u] & (1u << (uint32_t)(cv % 32u)))) continue;
                    const uint32_t bv = 1024u * z + 32u * y + x;
                    const uint32_t m = 1u << (bv & 31u);
                    if (!(blk[bv >> 5] & m)) {
                        blk[bv >> 5] |= m;
                        mat[bv] = (uint16_t)(lut | (sc->materials ? sc->materials[cv] : 0u));
                    }
                    any = 1;
                }
            }
        }
        if (any && last_cidx != cidx) {
            if (!grow((void**)&s->list, &s->cap_list, s->n_list + 1, sizeof(uint32_t))) { s->failed = 1; return; }
            s->list[s->n_list++] = cidx;
            s->leaf_n[li]++;
            last_cidx = cidx;
        }
    }
}

static void construct_inner(tgxo_svo* s, uint32_t node, v3 pmin, v3 pmax, uint32_t n_cp, const uint32_t* cps)
{
    const tgxo_scene* sc = s->scene;
    uint32_t per_child[8] = {0};
    uint32_t* lists = (uint32_t*)malloc(sizeof(uint32_t) * 8 * (size_t)(n_cp ? n_cp : 1));
    if (!lists) { s->failed = 1; return; }
    const v3 ce = v3_mulf(v3_sub(pmax, pmin), 0.5f);
    uint32_t valid = 0;
    const float cmn[3] = {0.0f, 0.0f, 0.0f}, cmx[3] = {8.0f, 8.0f, 8.0f};
    for (uint32_t k = 0; k < n_cp; k++) {
        const uint32_t cptr = cps[k];
        const uint32_t cidx = sc->cluster_pointers[cptr];
        const tgx_object* o = &sc->objects[sc->cluster_idx_to_object[cidx]];
        const m4 ws2ms = ws2ms_of(o, cptr - o->first_cluster_pointer);
        for (uint32_t c = 0; c < 8; c++) {
            const v3 d = {(float)(c % 2) * ce.x, (float)((c / 2) % 2) * ce.y, (float)((c / 4) % 2) * ce.z};
            const v3 cmin = v3_add(pmin, d);
            const v3 cmax = v3_add(cmin, ce);
            float corners[24];
            for (int i = 0; i < 8; i++) {
                const v4 p = {(i & 1) ? cmax.x : cmin.x, (i & 2) ? cmax.y : cmin.y, (i & 4) ? cmax.z : cmin.z, 1.0f};
                const v3 q = m4_mulv4_xyz(ws2ms, p);
                corners[3 * i] = q.x; corners[3 * i + 1] = q.y; corners[3 * i + 2] = q.z;
            }
            if (tgxo_aabb_obb_ignore_contact(cmn, cmx, corners)) {
                valid |= 1u << c;
                uint32_t* l = &lists[(size_t)c * n_cp];
                if (per_child[c] == 0 || l[per_child[c] - 1] != cptr) l[per_child[c]++] = cptr;
            }
        }
    }
    if (valid) {
        s->nodes[node].valid = valid;
        const uint32_t cnt = (uint32_t)__builtin_popcount(valid);
        const uint32_t first = new_nodes(s, cnt);
        if (s->failed) { free(lists); return; }
        s->nodes[node].first_child = first;
        const uint32_t child_vox = (uint32_t)ce.x * (uint32_t)ce.y * (uint32_t)ce.z;
        const int as_leaves = child_vox == TGX_SVO_BLOCK_VOXELS;
        uint32_t off = 0;
        for (uint32_t c = 0; c < 8; c++) {
            if (!(valid & (1u << c))) continue;
            const v3 d = {(float)(c % 2) * ce.x, (float)((c / 2) % 2) * ce.y, (float)((c / 4) % 2) * ce.z};
            const v3 cmin = v3_add(pmin, d);
            const v3 cmax = v3_add(cmin, ce);
            const uint32_t child = first + off;
            onode* cn = &s->nodes[child];
            cn->level = s->nodes[node].level + 1;
            cn->coord[0] = 2u * s->nodes[node].coord[0] + (c & 1u);
            cn->coord[1] = 2u * s->nodes[node].coord[1] + ((c >> 1) & 1u);
            cn->coord[2] = 2u * s->nodes[node].coord[2] + ((c >> 2) & 1u);
            if (!as_leaves) {
                construct_inner(s, child, cmin, cmax, per_child[c], &lists[(size_t)c * n_cp]);
            } else {
                s->nodes[node].leaf |= 1u << c;
                construct_leaf(s, child, cmin, cmax, per_child[c], &lists[(size_t)c * n_cp]);
            }
            if (s->failed) { free(lists); return; }
            off++;
        }
    }
    free(lists);
}

tgxo_svo* tgxo_svo_build(const tgxo_scene* sc, const float mn[3], uint32_t depth, uint32_t flags)
{
    if (depth < TGX_SVO_MIN_DEPTH || depth > TGX_SVO_MAX_DEPTH) return NULL;
    tgxo_svo* s = (tgxo_svo*)calloc(1, sizeof(tgxo_svo));
    s->scene = sc;
    s->flags = flags;
    s->info.min[0] = mn[0]; s->info.min[1] = mn[1]; s->info.min[2] = mn[2];
    s->info.depth = depth;
    s->info.flags = flags;
    (void)new_nodes(s, 1); /* root */
    /* Non-empty clusters, object by object (:496-537). */
    uint32_t* cps = (uint32_t*)malloc(sizeof(uint32_t) * (sc->n_cluster_pointers + 1u));
    uint32_t n = 0;
    uint32_t cur = 0;
    while (cur < sc->n_cluster_pointers) {
        const tgx_object* o = &sc->objects[sc->cluster_idx_to_object[sc->cluster_pointers[cur]]];
        const uint32_t cnt = o->n_cluster_pointers_per_dim[0] * o->n_cluster_pointers_per_dim[1] * o->n_cluster_pointers_per_dim[2];
        cur += cnt;
        for (uint32_t r = 0; r < cnt; r++) {
            const uint32_t cptr = o->first_cluster_pointer + r;
            const uint32_t cidx = sc->cluster_pointers[cptr];
            const uint32_t* w = &sc->bits[(size_t)cidx * 16u];
            int any = 0;
            for (int i = 0; i < 16; i++) any |= w[i] != 0;
            if (any) cps[n++] = cptr;
        }
    }
    const float side = (float)(1u << depth);
    const v3 pmin = {mn[0], mn[1], mn[2]};
    const v3 pmax = {mn[0] + side, mn[1] + side, mn[2] + side};
    construct_inner(s, 0, pmin, pmax, n, cps);
    free(cps);
    if (s->failed) { tgxo_svo_free(s); return NULL; }
    if (s->list_off) s->list_off[s->n_leaves] = s->n_list;
    s->info.n_nodes = s->n_nodes;
    s->info.n_leaves = s->n_leaves;
    s->info.n_list_entries = s->n_list;
    return s;
}

void tgxo_svo_free(tgxo_svo* s)
{
    if (!s) return;
    free(s->nodes); free(s->bits); free(s->mats); free(s->leaf_n); free(s->leaf_node);
    free(s->list_off); free(s->list);
    free(s);
}

void tgxo_svo_get_info(const tgxo_svo* s, tgx_svo_info* out) { *out = s->info; }

/* ------------------------------------------------------------------------ */
/* Canonical export: nodes by (level, Morton(coord))                         */
/* ------------------------------------------------------------------------ */

static uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1FFFFFu;
    x = (x | x << 32) & 0x1F00000000FFFFull;
    x = (x | x << 16) & 0x1F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

uint64_t tgxo_svo_node_key(uint32_t level, const uint32_t c[3])
{
    return ((uint64_t)level << 58) | spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
}

static const tgxo_svo* g_sort_svo;
static int cmp_nodes(const void* a, const void* b)
{
    const onode* na = &g_sort_svo->nodes[*(const uint32_t*)a];
    const onode* nb = &g_sort_svo->nodes[*(const uint32_t*)b];
    const uint64_t ka = tgxo_svo_node_key(na->level, na->coord), kb = tgxo_svo_node_key(nb->level, nb->coord);
    return ka < kb ? -1 : ka > kb;
}

int tgxo_svo_export(const tgxo_svo* s, tgx_svo_node* nodes, uint32_t* leaf_bits, uint16_t* leaf_mats,
                    uint32_t* leaf_n, uint32_t* list_offsets, uint32_t* list)
{
    const uint32_t nn = s->n_nodes;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * nn);
    uint32_t* rank = (uint32_t*)malloc(sizeof(uint32_t) * nn);
    for (uint32_t i = 0; i < nn; i++) order[i] = i;
    g_sort_svo = s;
    qsort(order, nn, sizeof(uint32_t), cmp_nodes);
    for (uint32_t i = 0; i < nn; i++) rank[order[i]] = i;
    /* leaves in canonical order */
    uint32_t* leaf_rank = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_leaves + 1u));
    uint32_t* leaf_of = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_leaves + 1u));
    uint32_t nl = 0;
    for (uint32_t i = 0; i < nn; i++) {
        const onode* n = &s->nodes[order[i]];
        if (n->leaf_idx != 0xFFFFFFFFu) { leaf_rank[n->leaf_idx] = nl; leaf_of[nl] = n->leaf_idx; nl++; }
    }
    for (uint32_t i = 0; nodes && i < nn; i++) {
        const onode* n = &s->nodes[order[i]];
        tgx_svo_node* o = &nodes[i];
        o->level = n->level;
        o->coord[0] = n->coord[0]; o->coord[1] = n->coord[1]; o->coord[2] = n->coord[2];
        o->valid_mask = n->valid;
        o->leaf_mask = n->leaf;
        o->first_child = n->first_child == 0xFFFFFFFFu ? 0xFFFFFFFFu : rank[n->first_child];
        o->leaf_idx = n->leaf_idx == 0xFFFFFFFFu ? 0xFFFFFFFFu : leaf_rank[n->leaf_idx];
    }
    uint32_t pos = 0;
    for (uint32_t j = 0; j < nl; j++) {
        const uint32_t li = leaf_of[j];
        if (leaf_bits) memcpy(&leaf_bits[(size_t)j * TGX_SVO_BLOCK_WORDS], &s->bits[(size_t)li * TGX_SVO_BLOCK_WORDS], TGX_SVO_BLOCK_WORDS * 4);
        if (leaf_mats) memcpy(&leaf_mats[(size_t)j * TGX_SVO_BLOCK_VOXELS], &s->mats[(size_t)li * TGX_SVO_BLOCK_VOXELS], TGX_SVO_BLOCK_VOXELS * 2);
        if (leaf_n) leaf_n[j] = s->leaf_n[li];
        const uint32_t b = s->list_off[li], e = s->list_off[li + 1];
        if (list_offsets) list_offsets[j] = pos;
        for (uint32_t k = b; k < e; k++) { if (list) list[pos] = s->list[k]; pos++; }
    }
    if (list_offsets) list_offsets[nl] = pos;
    free(order); free(rank); free(leaf_rank); free(leaf_of);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Traversal (tg_sparse_voxel_octree.c:558-740)                              */
/* ------------------------------------------------------------------------ */

static int intersect_ray_aabb(v3 o, v3 d, v3 mn, v3 mx, float* enter, float* exit_)
{
    const v3 vmin = {F32_MIN, F32_MIN, F32_MIN}, vmax = {F32_MAX, F32_MAX, F32_MAX};
    const v3 a = v3_div_zero_check(v3_sub(mn, o), d, vmin);
    const v3 b = v3_div_zero_check(v3_sub(mx, o), d, vmax);
    const v3 n = v3_min(a, b), f = v3_max(a, b);
    *enter = fmax2(fmax2(n.x, n.y), n.z);
    *exit_ = fmin2(fmin2(f.x, f.y), f.z);
    return *exit_ > 0.0f && *enter <= *exit_;
}

/* tg_amanatides_woo over a 32^3 block; the entry voxel is clamped into the
 * block on both sides. Returns 1 with the voxel coordinates on a hit. */
static int block_dda(v3 h, v3 d, const uint32_t* grid, int* ox, int* oy, int* oz, uint64_t* rec_bits,
                     uint64_t rec_base)
{
    const float e = 32.0f;
    const float em1 = e - 1.0f;
    const float fx = fmin2(fmax2(floorf(h.x), 0.0f), em1);
    const float fy = fmin2(fmax2(floorf(h.y), 0.0f), em1);
    const float fz = fmin2(fmax2(floorf(h.z), 0.0f), em1);
    int x = (int)fx, y = (int)fy, z = (int)fz;
    int sx = 0, sy = 0, sz = 0;
    float tmx = F32_MAX, tmy = F32_MAX, tmz = F32_MAX;
    float tdx = F32_MAX, tdy = F32_MAX, tdz = F32_MAX;
    if (d.x > 0.0f) { sx = 1; tmx = ((float)(x + 1) - h.x) / d.x; tdx = 1.0f / d.x; }
    else if (d.x < 0.0f) { sx = -1; tmx = (h.x - (float)x) / -d.x; tdx = 1.0f / -d.x; }
    if (d.y > 0.0f) { sy = 1; tmy = ((float)(y + 1) - h.y) / d.y; tdy = 1.0f / d.y; }
    else if (d.y < 0.0f) { sy = -1; tmy = (h.y - (float)y) / -d.y; tdy = 1.0f / -d.y; }
    if (d.z > 0.0f) { sz = 1; tmz = ((float)(z + 1) - h.z) / d.z; tdz = 1.0f / d.z; }
    else if (d.z < 0.0f) { sz = -1; tmz = (h.z - (float)z) / -d.z; tdz = 1.0f / -d.z; }
    for (uint32_t it = 0; it < 4u * TGX_DDA_MAX_ITERATIONS; it++) {
        const uint32_t v = 1024u * (uint32_t)z + 32u * (uint32_t)y + (uint32_t)x;
        mark_line(rec_bits, rec_base + (v >> 9)); /* 16 words per line */
        if (grid[v >> 5] & (1u << (v & 31u))) { *ox = x; *oy = y; *oz = z; return 1; }
        if (tmx < tmy) {
            if (tmx < tmz) { tmx += tdx; x += sx; if (x < 0 || x >= 32) break; }
            else           { tmz += tdz; z += sz; if (z < 0 || z >= 32) break; }
        } else {
            if (tmy < tmz) { tmy += tdy; y += sy; if (y < 0 || y >= 32) break; }
            else           { tmz += tdz; z += sz; if (z < 0 || z >= 32) break; }
        }
    }
    return 0;
}

/* Entry face of the voxel: axis of the largest near-plane distance, in the
 * max(max(nx, ny), nz) evaluation order, pointing against the ray. */
static void entry_normal(v3 o, v3 d, v3 mn, v3 mx, int32_t nrm[3])
{
    const v3 vmin = {F32_MIN, F32_MIN, F32_MIN}, vmax = {F32_MAX, F32_MAX, F32_MAX};
    const v3 a = v3_div_zero_check(v3_sub(mn, o), d, vmin);
    const v3 b = v3_div_zero_check(v3_sub(mx, o), d, vmax);
    const v3 n = v3_min(a, b);
    int ax = n.x > n.y ? 0 : 1;
    const float nax = ax == 0 ? n.x : n.y;
    if (!(nax > n.z)) ax = 2;
    nrm[0] = nrm[1] = nrm[2] = 0;
    const float dc = ax == 0 ? d.x : (ax == 1 ? d.y : d.z);
    nrm[ax] = dc > 0.0f ? -1 : 1;
}

static void traverse_rec(const tgxo_svo* s, const float origin[3], const float dir[3], float max_distance,
                         tgx_svo_hit* out, const line_rec* rec);

void tgxo_svo_traverse(const tgxo_svo* s, const float origin[3], const float dir[3], float max_distance,
                       tgx_svo_hit* out)
{
    traverse_rec(s, origin, dir, max_distance, out, NULL);
}

static void traverse_rec(const tgxo_svo* s, const float origin[3], const float dir[3], float max_distance,
                         tgx_svo_hit* out, const line_rec* rec)
{
    memset(out, 0, sizeof(*out));
    out->distance = F32_MAX;
    const float side = (float)(1u << s->info.depth);
    const float half = side * 0.5f;
    /* origin relative to the root centre; root box [-half, half]^3 */
    const v3 centre = {s->info.min[0] + half, s->info.min[1] + half, s->info.min[2] + half};
    const v3 o = {origin[0] - centre.x, origin[1] - centre.y, origin[2] - centre.z};
    const v3 d = {dir[0], dir[1], dir[2]};
    const v3 rmin = {-half, -half, -half}, rmax = {half, half, half};
    const v3 vmin = {F32_MIN, F32_MIN, F32_MIN}, vmax = {F32_MAX, F32_MAX, F32_MAX};

    uint32_t st_idx[TGX_SVO_MAX_DEPTH];
    v3 st_min[TGX_SVO_MAX_DEPTH], st_max[TGX_SVO_MAX_DEPTH];
    float enter, exit_;
    if (!intersect_ray_aabb(o, d, rmin, rmax, &enter, &exit_)) return;
    v3 pos = enter > 0.0f ? v3_add(o, v3_mulf(d, enter)) : o;
    float dist = enter > 0.0f ? enter : 0.0f;
    uint32_t ss = 1;
    st_idx[0] = 0;
    st_min[0] = rmin;
    st_max[0] = rmax;
    uint32_t steps = 0;
    while (ss > 0) {
        if (steps >= TGX_SVO_MAX_STEPS || dist > max_distance) break;
        steps++;
        const uint32_t pi = st_idx[ss - 1];
        const v3 pmn = st_min[ss - 1], pmx = st_max[ss - 1];
        const onode* pn = &s->nodes[pi];
        if (rec) mark_line(rec->node, (uint64_t)pi * 4u / 64u);
        const v3 ce = v3_mulf(v3_sub(pmx, pmn), 0.5f);
        uint32_t rc = 0;
        v3 cmin = pmn;
        v3 cmax = v3_add(cmin, ce);
        if (cmax.x < pos.x || (pos.x == cmax.x && d.x > 0.0f)) { rc += 1; cmin.x += ce.x; cmax.x += ce.x; }
        if (cmax.y < pos.y || (pos.y == cmax.y && d.y > 0.0f)) { rc += 2; cmin.y += ce.y; cmax.y += ce.y; }
        if (cmax.z < pos.z || (pos.z == cmax.z && d.z > 0.0f)) { rc += 4; cmin.z += ce.z; cmax.z += ce.z; }
        int advance = 1;
        if (pn->valid & (1u << rc)) {
            const uint32_t off = (uint32_t)__builtin_popcount(pn->valid & ((1u << rc) - 1u));
            const uint32_t ci = pn->first_child + off;
            if (pn->leaf & (1u << rc)) {
                const uint32_t li = s->nodes[ci].leaf_idx;
                if (rec) {
                    mark_line(rec->node, (uint64_t)ci * 4u / 64u);
                    mark_line(rec->leafd, (uint64_t)li * 260u / 64u);
                }
                if (s->leaf_n[li] != 0) {
                    const v3 hg = v3_sub(pos, cmin);
                    int vx, vy, vz;
                    if (block_dda(hg, d, &s->bits[(size_t)li * TGX_SVO_BLOCK_WORDS], &vx, &vy, &vz,
                                  rec ? rec->bits : NULL, (uint64_t)li * 64u)) {
                        const v3 vmn = v3_add(cmin, (v3){(float)vx, (float)vy, (float)vz});
                        const v3 vmx = v3_add(vmn, (v3){1.0f, 1.0f, 1.0f});
                        (void)intersect_ray_aabb(pos, d, vmn, vmx, &enter, &exit_);
                        out->distance = dist + enter;
                        out->hit = out->distance <= max_distance;
                        const onode* cn = &s->nodes[ci];
                        out->block[0] = cn->coord[0]; out->block[1] = cn->coord[1]; out->block[2] = cn->coord[2];
                        out->voxel = 1024u * (uint32_t)vz + 32u * (uint32_t)vy + (uint32_t)vx;
                        entry_normal(pos, d, vmn, vmx, out->normal);
                        out->steps = steps;
                        return;
                    }
                }
            } else {
                advance = 0;
                st_idx[ss] = ci;
                st_min[ss] = cmin;
                st_max[ss] = cmax;
                ss++;
            }
        }
        if (advance) {
            v3 a = v3_div_zero_check(v3_sub(cmin, pos), d, vmin);
            v3 b = v3_div_zero_check(v3_sub(cmax, pos), d, vmax);
            v3 f = v3_max(a, b);
            exit_ = v3_min_elem(f);
            const float t_adv = exit_ + F32_EPSILON;
            pos = v3_add(pos, v3_mulf(d, t_adv));
            dist += t_adv;
            while (ss > 0) {
                a = v3_div_zero_check(v3_sub(st_min[ss - 1], pos), d, vmin);
                b = v3_div_zero_check(v3_sub(st_max[ss - 1], pos), d, vmax);
                f = v3_max(a, b);
                exit_ = v3_min_elem(f);
                if (exit_ > F32_EPSILON) break;
                ss--;
            }
        }
    }
    out->steps = steps;
}

void tgxo_svo_trace(const tgxo_svo* s, const float* origins, const float* dirs, uint32_t n, float max_distance,
                    tgx_svo_hit* out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; i++) tgxo_svo_traverse(s, &origins[3 * i], &dirs[3 * i], max_distance, &out[i]);
}

/* ------------------------------------------------------------------------ */
/* One-bounce GI (the estimator documented at tgx_gi_params in tgx.h)        */
/* ------------------------------------------------------------------------ */

#define SQRT3 1.73205080757f
#define SHADOW_OFFSET 0.015625f
#define INV_PI 0.318309886f

/* MurmurHash3 finaliser (util.inc:47-56). */
static uint32_t fmix32(uint32_t v)
{
    v ^= v >> 16;
    v *= 0x85ebca6bu;
    v ^= v >> 13;
    v *= 0xc2b2ae35u;
    v ^= v >> 16;
    return v;
}

/* xorshift32 (util.inc:11-19) and its [low, high] float (util.inc:21-31). */
static uint32_t xorshift32(uint32_t* s)
{
    uint32_t r = *s;
    r ^= r << 13;
    r ^= r >> 17;
    r ^= r << 5;
    *s = r;
    return r;
}
static float xorshift_f32_range(uint32_t* s, float lo, float hi)
{
    const float u = (float)xorshift32(s) / (float)0xFFFFFFFFu;
    return u * (hi - lo) + lo;
}

static v3 cosine_direction(v3 n, uint32_t* st)
{
    for (int attempt = 0; attempt < 32; attempt++) {
        const float x = xorshift_f32_range(st, -1.0f, 1.0f);
        const float y = xorshift_f32_range(st, -1.0f, 1.0f);
        const float z = xorshift_f32_range(st, -1.0f, 1.0f);
        const float l2 = x * x + y * y + z * z;
        if (l2 > 1.0f || l2 < 1e-6f) continue;
        const float l = sqrtf(l2);
        const float dx = n.x + x / l, dy = n.y + y / l, dz = n.z + z / l;
        const float dl2 = dx * dx + dy * dy + dz * dz;
        if (dl2 < 1e-6f) continue;
        const float dl = sqrtf(dl2);
        v3 r = {dx / dl, dy / dl, dz / dl};
        return r;
    }
    return n;
}

void tgxo_gi_direction(const float n[3], uint32_t seed, float out[3])
{
    uint32_t st = seed;
    const v3 nv = {n[0], n[1], n[2]};
    const v3 d = cosine_direction(nv, &st);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

uint32_t tgxo_gi_seed(uint32_t pixel, uint32_t sample, uint32_t spp, uint32_t frame)
{
    const uint32_t s = fmix32(pixel * spp + sample + 0x9E3779B9u * (frame + 1u));
    return s ? s : 0x6D2B79F5u;
}

/* Primary surface of a visibility word: world hit point, world normal and
 * albedo (shading.frag:116-229). Returns 0 if the pixel is not shaded. */
int tgxo_gi_surface(const tgxo_scene* sc, const tgx_camera* cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                    uint64_t word, float out_p[3], float out_n[3], float out_albedo[3])
{
    const float depth = (float)(word >> 40) / 16777215.0f;
    if (!(depth < 1.0f)) return 0;
    const uint32_t cptr = (uint32_t)(word >> 9) & 0x7FFFFFFFu;
    const uint32_t v = (uint32_t)word & 511u;
    const uint32_t cidx = sc->cluster_pointers[cptr];
    const tgx_object* o = &sc->objects[sc->cluster_idx_to_object[cidx]];
    const uint32_t rel = cptr - o->first_cluster_pointer;
    const m4 ws2ms = ws2ms_of(o, rel);
    const v4 cpos = {cam->camera[0], cam->camera[1], cam->camera[2], 1.0f};
    const v3 orig = m4_mulv4_xyz(ws2ms, cpos);
    const float fx = ((float)px + 0.5f) / (float)w;
    const float fy = 1.0f - ((float)py + 0.5f) / (float)h;
    const float omy = 1.0f - fy, omx = 1.0f - fx;
    v3 dws;
    {
        const float lx = cam->ray_bl[0] * omy + cam->ray_tl[0] * fy, ly = cam->ray_bl[1] * omy + cam->ray_tl[1] * fy,
                    lz = cam->ray_bl[2] * omy + cam->ray_tl[2] * fy;
        const float rx = cam->ray_br[0] * omy + cam->ray_tr[0] * fy, ry = cam->ray_br[1] * omy + cam->ray_tr[1] * fy,
                    rz = cam->ray_br[2] * omy + cam->ray_tr[2] * fy;
        dws.x = lx * omx + rx * fx; dws.y = ly * omx + ry * fx; dws.z = lz * omx + rz * fx;
    }
    const v4 dv = {dws.x, dws.y, dws.z, 0.0f};
    const v3 dms = v3_normalized(m4_mulv4_xyz(ws2ms, dv));
    const v3 vmn = {(float)(v % 8u), (float)((v / 8u) % 8u), (float)(v / 64u)};
    const v3 vmx = {vmn.x + 1.0f, vmn.y + 1.0f, vmn.z + 1.0f};
    float enter, exit_;
    if (!intersect_ray_aabb(orig, dms, vmn, vmx, &enter, &exit_)) return 0;
    const v3 hit = enter > 0.0f ? v3_add(orig, v3_mulf(dms, enter)) : orig;
    const v3 vc = {vmn.x + 0.5f, vmn.y + 0.5f, vmn.z + 0.5f};
    v3 n = v3_sub(hit, vc);
    const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
#define SIGNF(a) ((a) > 0.0f ? 1.0f : ((a) < 0.0f ? -1.0f : 0.0f))
    if (ax > ay) {
        n.y = 0.0f;
        if (ax > az) { n.x = SIGNF(n.x); n.z = 0.0f; } else { n.z = SIGNF(n.z); n.x = 0.0f; }
    } else {
        n.x = 0.0f;
        if (ay > az) { n.y = SIGNF(n.y); n.z = 0.0f; } else { n.z = SIGNF(n.z); n.y = 0.0f; }
    }
#undef SIGNF
    m4 rot;
    memcpy(rot.m, o->rotation, sizeof(rot.m));
    const v4 nv = {n.x, n.y, n.z, 0.0f};
    const v3 nws = v3_normalized(m4_mulv4_xyz(rot, nv));
    const m4 cs2ws = ms2ws_of(o, rel);
    const v4 hv = {hit.x, hit.y, hit.z, 1.0f};
    const v3 pws = m4_mulv4_xyz(cs2ws, hv);
    const uint32_t mat = sc->materials ? sc->materials[(size_t)cidx * 512u + v] : 0u;
    const tgx_lut_entry* e = &sc->luts[(size_t)o->lut_idx * TGX_LUT_ENTRIES + mat];
    out_p[0] = pws.x; out_p[1] = pws.y; out_p[2] = pws.z;
    out_n[0] = nws.x; out_n[1] = nws.y; out_n[2] = nws.z;
    out_albedo[0] = e->albedo[0]; out_albedo[1] = e->albedo[1]; out_albedo[2] = e->albedo[2];
    return 1;
}

uint64_t tgxo_gi_surface_count(const tgxo_scene* sc, const tgx_camera* cam, uint32_t w, uint32_t h,
                               const uint64_t* vis, int n_threads)
{
    if (n_threads > 0) omp_set_num_threads(n_threads);
    uint64_t n = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : n)
    for (int64_t y = 0; y < (int64_t)h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const uint64_t word = vis[(size_t)y * w + x];
            float p[3], nn[3], a[3];
            if (word != ~0ull && tgxo_gi_surface(sc, cam, w, h, x, (uint32_t)y, word, p, nn, a)) n++;
        }
    return n;
}

/* Leaf index of the hit's block (binary search over the canonical order
 * is not needed: the traversal below records it). */
static void traverse_leaf(const tgxo_svo* s, const float o[3], const float d[3], float maxd, tgx_svo_hit* h,
                          uint32_t* leaf, const line_rec* rec)
{
    traverse_rec(s, o, d, maxd, h, rec);
    *leaf = 0xFFFFFFFFu;
    if (!h->hit) return;
    /* locate the leaf node of the hit block by descending from the root */
    uint32_t ni = 0;
    const uint32_t leaf_level = s->info.depth - 5u;
    for (uint32_t l = 0; l < leaf_level; l++) {
        const uint32_t sh = leaf_level - l - 1u;
        const uint32_t c = ((h->block[0] >> sh) & 1u) | (((h->block[1] >> sh) & 1u) << 1) | (((h->block[2] >> sh) & 1u) << 2);
        const onode* n = &s->nodes[ni];
        ni = n->first_child + (uint32_t)__builtin_popcount(n->valid & ((1u << c) - 1u));
    }
    *leaf = s->nodes[ni].leaf_idx;
}

static void gi_impl(const tgxo_svo* s, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis,
                    const tgx_gi_params* prm, uint32_t row_step, float* radiance, float* irradiance,
                    uint64_t counts[4], const line_rec* rec);

void tgxo_gi(const tgxo_svo* s, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis,
             const tgx_gi_params* prm, uint32_t row_step, float* radiance, float* irradiance, uint64_t counts[4])
{
    gi_impl(s, cam, w, h, vis, prm, row_step, radiance, irradiance, counts, NULL);
}

/* Marks the scene lines tgxo_gi_surface reads for the visibility word of a
 * visible pixel (decode, maps, object record, material, LUT entry). */
static void mark_surface(const tgxo_scene* sc, uint64_t word, const line_rec* rec)
{
    const uint32_t cptr = (uint32_t)(word >> 9) & 0x7FFFFFFFu, v = (uint32_t)word & 511u;
    const uint32_t cidx = sc->cluster_pointers[cptr];
    const uint32_t obj = sc->cluster_idx_to_object[cidx];
    mark_line(rec->cptr, (uint64_t)cptr * 4u / 64u);
    mark_line(rec->c2o, (uint64_t)cidx * 4u / 64u);
    mark_line(rec->obj, (uint64_t)obj * 96u / 64u);
    mark_line(rec->obj, ((uint64_t)obj * 96u + 95u) / 64u);
    const uint32_t mat = sc->materials ? sc->materials[(size_t)cidx * 512u + v] : 0u;
    if (sc->materials) mark_line(rec->smat, ((uint64_t)cidx * 512u + v) / 64u);
    mark_line(rec->lutsp, ((uint64_t)sc->objects[obj].lut_idx * TGX_LUT_ENTRIES + mat) * 16u / 64u);
}

static void gi_impl(const tgxo_svo* s, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis,
                    const tgx_gi_params* prm, uint32_t row_step, float* radiance, float* irradiance,
                    uint64_t counts[4], const line_rec* rec)
{
    const tgxo_scene* sc = s->scene;
    const v3 sun = {prm->sun_dir[0], prm->sun_dir[1], prm->sun_dir[2]};
    uint64_t n_vis = 0, n_b = 0, n_bh = 0, n_sh = 0;
    if (row_step == 0) row_step = 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : n_vis, n_b, n_bh, n_sh)
    for (int64_t yy = 0; yy < (int64_t)h; yy += row_step) {
        const uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < w; x++) {
            const size_t pix = (size_t)y * w + x;
            float p[3], nrm[3], alb[3];
            float acc[3] = {0.0f, 0.0f, 0.0f};
            radiance[3 * pix] = radiance[3 * pix + 1] = radiance[3 * pix + 2] = 0.0f;
            if (irradiance) irradiance[3 * pix] = irradiance[3 * pix + 1] = irradiance[3 * pix + 2] = 0.0f;
            if (!tgxo_gi_surface(sc, cam, w, h, x, y, vis[pix], p, nrm, alb)) continue;
            if (rec) mark_surface(sc, vis[pix], rec);
            n_vis++;
            const v3 n = {nrm[0], nrm[1], nrm[2]};
            for (uint32_t k = 0; k < prm->spp; k++) {
                uint32_t st = tgxo_gi_seed((uint32_t)pix, k, prm->spp, prm->frame);
                const v3 d = cosine_direction(n, &st);
                const float ob[3] = {p[0] + SQRT3 * n.x, p[1] + SQRT3 * n.y, p[2] + SQRT3 * n.z};
                const float dd[3] = {d.x, d.y, d.z};
                tgx_svo_hit hb;
                uint32_t leaf;
                traverse_leaf(s, ob, dd, prm->max_distance, &hb, &leaf, rec);
                n_b++;
                if (!hb.hit) continue;
                n_bh++;
                if (hb.distance < 0.0f) continue; /* started inside solid: occluded */
                int ax = hb.normal[0] != 0 ? 0 : (hb.normal[1] != 0 ? 1 : 2);
                const float cosv = (float)hb.normal[ax] * prm->sun_dir[ax];
                if (!(cosv > 0.0f)) continue;
                float os[3] = {ob[0] + d.x * hb.distance, ob[1] + d.y * hb.distance, ob[2] + d.z * hb.distance};
                os[ax] = os[ax] + (float)hb.normal[ax] * SHADOW_OFFSET;
                const float sd[3] = {sun.x, sun.y, sun.z};
                tgx_svo_hit hs;
                traverse_rec(s, os, sd, prm->max_distance, &hs, rec);
                n_sh++;
                if (hs.hit) continue;
                const uint16_t m = s->mats[(size_t)leaf * TGX_SVO_BLOCK_VOXELS + hb.voxel];
                if (rec) {
                    mark_line(rec->mats, ((uint64_t)leaf * TGX_SVO_BLOCK_VOXELS + hb.voxel) * 2u / 64u);
                    mark_line(rec->luts, ((uint64_t)(m >> 8) * TGX_LUT_ENTRIES + (m & 255u)) * 16u / 64u);
                }
                const tgx_lut_entry* e = &sc->luts[(size_t)(m >> 8) * TGX_LUT_ENTRIES + (m & 255u)];
                const float wgt = cosv * prm->sun_radiance * INV_PI;
                acc[0] = acc[0] + e->albedo[0] * wgt;
                acc[1] = acc[1] + e->albedo[1] * wgt;
                acc[2] = acc[2] + e->albedo[2] * wgt;
            }
            const float inv = 1.0f / (float)prm->spp;
            const float ex = acc[0] * inv, ey = acc[1] * inv, ez = acc[2] * inv;
            radiance[3 * pix] = alb[0] * ex;
            radiance[3 * pix + 1] = alb[1] * ey;
            radiance[3 * pix + 2] = alb[2] * ez;
            if (irradiance) { irradiance[3 * pix] = ex; irradiance[3 * pix + 1] = ey; irradiance[3 * pix + 2] = ez; }
        }
    }
    counts[0] = n_vis; counts[1] = n_b; counts[2] = n_bh; counts[3] = n_sh;
}

/* ------------------------------------------------------------------------ */
/* Algorithmic bytes of a GI frame                                          */
/* ------------------------------------------------------------------------ */

static uint64_t* line_bitmap(uint64_t bytes)
{
    return (uint64_t*)calloc((size_t)((bytes + 63u) / 64u / 64u + 1u), sizeof(uint64_t));
}

static uint64_t count_lines(const uint64_t* bm, uint64_t bytes)
{
    uint64_t n = 0;
    for (uint64_t i = 0; i < (bytes + 63u) / 64u / 64u + 1u; i++) n += (uint64_t)__builtin_popcountll(bm[i]);
    return n;
}

int tgxo_gi_alg_bytes(const tgxo_svo* s, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis,
                      const tgx_gi_params* prm, uint32_t with_irradiance, uint64_t out[16])
{
    const tgxo_scene* sc = s->scene;
    const uint64_t b_node = 4ull * s->n_nodes, b_leafd = 260ull * s->n_leaves;
    const uint64_t b_bits = 4096ull * s->n_leaves, b_mats = 2ull * TGX_SVO_BLOCK_VOXELS * s->n_leaves;
    const uint64_t b_luts = 16ull * TGX_LUT_ENTRIES * (sc->n_luts ? sc->n_luts : 1u);
    const uint64_t b_cptr = 4ull * sc->n_cluster_pointers, b_c2o = 4ull * sc->n_clusters;
    const uint64_t b_obj = 96ull * sc->n_objects, b_smat = 512ull * sc->n_clusters;
    line_rec r = {line_bitmap(b_node), line_bitmap(b_leafd), line_bitmap(b_bits), line_bitmap(b_mats),
                  line_bitmap(b_luts), line_bitmap(b_luts), line_bitmap(b_cptr), line_bitmap(b_c2o),
                  line_bitmap(b_obj), line_bitmap(b_smat)};
    float* rad = (float*)malloc(sizeof(float) * 3u * (size_t)w * h);
    int ok = r.node && r.leafd && r.bits && r.mats && r.luts && r.lutsp && r.cptr && r.c2o && r.obj && r.smat && rad;
    if (ok) {
        uint64_t counts[4];
        gi_impl(s, cam, w, h, vis, prm, 1u, rad, NULL, counts, &r);
        const uint64_t P = (uint64_t)w * h;
        uint64_t* n = out + 4;
        n[0] = count_lines(r.node, b_node);
        n[1] = count_lines(r.leafd, b_leafd);
        n[2] = count_lines(r.bits, b_bits);
        n[3] = count_lines(r.mats, b_mats);
        n[4] = count_lines(r.luts, b_luts);
        n[5] = count_lines(r.cptr, b_cptr);
        n[6] = count_lines(r.c2o, b_c2o);
        n[7] = count_lines(r.obj, b_obj);
        n[8] = count_lines(r.smat, b_smat);
        n[9] = count_lines(r.lutsp, b_luts);
        n[10] = (8u * P + 63u) / 64u; /* every visibility word is read once */
        n[11] = 0;
        out[0] = 64u * (n[10] + n[5] + n[6] + n[7] + n[8] + n[9]);
        out[1] = 64u * (n[0] + n[1] + n[2] + n[3] + n[4]);
        out[2] = (with_irradiance ? 24u : 12u) * P;
        out[3] = out[0] + out[1] + out[2];
    }
    free(r.node); free(r.leafd); free(r.bits); free(r.mats); free(r.luts); free(r.lutsp);
    free(r.cptr); free(r.c2o); free(r.obj); free(r.smat); free(rad);
    return ok ? 0 : -1;
}
