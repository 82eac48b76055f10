/*
 * tgx_gi.hip — SVO ray traversal and one-bounce GI on gfx950.
 *
 * Traversal (tg_svo_traverse, tg_sparse_voxel_octree.c:558-740). The
 * reference's node stack is exactly the path from the root to the current
 * node, so a lane keeps only the current node, its level and its integer
 * coordinates; popping walks parent links, and every node box is recomputed
 * from (level, coordinates) with the values the reference's stack holds
 * (integers below 2^24, exact in fp32). Slab quotient pairs (box min, box
 * max) are reduced to the one quotient that min/max would pick, chosen by
 * the sign of the direction component: subtraction and IEEE division are
 * monotone, so the value is the same and half of the divisions disappear.
 * The pop test (exit distance > FLT_EPSILON) needs no division at all
 * (axis_exit_gt_eps), and the DDA step lengths 1/|d| are computed once per
 * ray.
 *
 * Layout read by the traversal (written by the SVO build):
 *   inner[n_inner]   uint2 per inner node: first child, parent | valid << 24.
 *                    Held in LDS with the brick masks when both fit.
 *                    Leaves follow the inner nodes in node order, so leaf
 *                    index = node - n_inner,
 *                    and a child is a leaf exactly at the leaf level.
 *   bricks[n_leaves] u64 per leaf: one bit per 8^3 brick holding a solid
 *                    voxel. The 32^3 DDA tests the brick bit before it
 *                    loads a bit word, and keeps the last word, so most DDA
 *                    iterations read no memory. Voxels tested and their
 *                    order are the reference's; only loads are skipped.
 *
 * gi_trace_k is persistent: each lane runs one ray as a state machine
 * whose steps are up to `node_iters` node iterations of the reference's
 * loop while the lane descends, then up to `burst` iterations of the leaf
 * DDA, so lanes of a wave stay busy while they are in different phases
 * (both tuning). In the uncounted DDA, once enough of a wave's walking
 * lanes are in 8^3 bricks without solid voxels, those lanes cross their
 * bricks in one exact step (brick_cross). Whenever `refill` (tuning) lanes
 * of a wave have finished, the wave hands that many new rays to them
 * (ballot + lane rank), from a slice of the queue it took with one atomic
 * (`chunk` rays, tuning).
 *
 * GI frame (estimator documented at tgx_gi_params in tgx.h):
 *   gi_gen_k     surface of each visible pixel, spp bounce rays at
 *                slots fixed by the list entry (gi_gen_tail_k then moves
 *                the queue's tail)
 *   gi_trace_k   bounce rays; a lane whose bounce hit faces the sun stores
 *                the sample record (weight, hit node and voxel) and traces
 *                the shadow ray next, clearing the weight if it is occluded
 *   gi_resolve_k per pixel, each sample's contribution (the hit voxel's
 *                albedo times its weight) summed in sample order
 * With sorted shadow rays (tuning trace_sort bit 1) the shadow rays go
 * through a queue instead: gi_shade_k writes them, a second gi_trace_k
 * traces them and gi_lit_k stores the unshadowed sample records. Both forms
 * do the same fp32 operations per sample. Sample records are stored per
 * (pixel, sample) and resolved and summed in sample order, so the fp32
 * result equals a sequential per-pixel loop bit for bit.
 */
#include "tgx_gi_dev.h"
#include "tgx_m4.h"

#define GI_SQRT3 1.73205080757f
#define SVO_INF __builtin_inff()
#define GI_SHADOW_OFFSET 0.015625f
#define GI_INV_PI 0.318309886f

#define TRACE_BLOCK 1024
#define DDA_BURST 8u
#define TRACE_ADV_MIN 16u   /* lanes of a wave that make it run the advances */
#define TRACE_LDS_NODES_MAX 8000u   /* inner nodes + leaf brick masks in LDS (64000 B) */

/* ------------------------------------------------------------------------ */
/* Traversal                                                                */
/* ------------------------------------------------------------------------ */

struct trace_params {
    const uint2*              inner;
    const unsigned long long* bricks;
    const uint32_t*           leaf_bits;
    uint32_t                  n_inner;
    uint32_t                  n_leaves;
    uint32_t                  depth;
    uint32_t                  leaf_level;
    float                     mnx, mny, mnz;
    float                     half;          /* 2^depth / 2 */
    float                     cx, cy, cz;    /* root centre, min + half */
};

/* Ray record of the queues (32 B). */
struct svo_ray {
    float    ox, oy, oz;
    float    dx, dy, dz;
    uint32_t id;
    uint32_t pad;
};

/* Result record (24 B): flags bit 0 hit, bit 1 reached a voxel (node and
 * voxel valid), bits 2-3 entry axis, bit 4 normal sign is positive. */
struct svo_res {
    float    distance;
    uint32_t node;
    uint32_t voxel;
    uint32_t flags;
    uint32_t steps;
    uint32_t pad;
};

/* Near / far plane distance along one axis: the numerator is selected
 * before the one division, so a single IEEE division is issued. */
__device__ static inline float axis_near(float o, float d, float mn, float mx)
{
    const float q = (d > 0.0f ? mn - o : mx - o) / d;
    return d == 0.0f ? -SVO_F32_MAX : q;
}
__device__ static inline float axis_far(float o, float d, float mn, float mx)
{
    const float q = (d > 0.0f ? mx - o : mn - o) / d;
    return d == 0.0f ? SVO_F32_MAX : q;
}

/* exit of tg_sparse_voxel_octree.c:704-707 (min over axes of the far plane) */
__device__ static inline float box_exit(float px, float py, float pz, float dx, float dy, float dz, float mnx,
                                        float mny, float mnz, float side)
{
    return fmin2(fmin2(axis_far(px, dx, mnx, mnx + side), axis_far(py, dy, mny, mny + side)),
                 axis_far(pz, dz, mnz, mnz + side));
}

/* box_exit(...) > SVO_EPSILON for one axis, without dividing. fl(a / d) >
 * 2^-23 holds exactly when a / d > 2^-23 (1 + 2^-24), the midpoint to the
 * next float (a tie rounds to the even 2^-23). For a normal t = d 2^-23 no
 * float lies in (|t|, |t| (1 + 2^-24)], so the test is a > t for d > 0 and
 * a < t for d < 0. Components with |d| < 2^-100 take the division. */
__device__ static inline bool axis_exit_gt_eps(float p, float d, float mn, float mx)
{
    if (d == 0.0f) return true; /* quotient F32_MAX */
    if (fabsf(d) < 0x1p-100f) return axis_far(p, d, mn, mx) > SVO_EPSILON;
    const float t = d * 0x1p-23f;
    return d > 0.0f ? (mx - p) > t : (mn - p) < t;
}

/* tg_amanatides_woo's per-axis quantities: pos / d for d > 0, neg / -d for
 * d < 0, F32_MAX for d == 0; one division. */
__device__ static inline float dda_len(float pos, float neg, float d)
{
    const float q = (d > 0.0f ? pos : neg) / (d > 0.0f ? d : -d);
    return d == 0.0f ? SVO_F32_MAX : q;
}

/* -1, 0 or 1 */
__device__ static inline int dir_step(float d) { return d > 0.0f ? 1 : (d < 0.0f ? -1 : 0); }

/* Per-lane ray state (22 registers). `node` packs the current node's
 * coordinates (7 bits each: inner nodes sit at level <= depth - 6 <= 7), its
 * level, the child slot being visited and the DDA phase flag; `vox` packs
 * the leaf DDA voxel and iteration count. */
/* axis_exit_gt_eps as a 0/1 integer for |d| >= 2^-100 or d == 0: the sign of
 * the exact-sign difference a - t (IEEE subtraction is 0 only for equal
 * operands, and rounding keeps the sign). Callers handle |d| < 2^-100. */
__device__ static inline uint32_t axis_exit_fast(float p, float d, float mn, float mx)
{
    const float a = (d > 0.0f ? mx : mn) - p;
    const float t = d * 0x1p-23f;
    const float u = d > 0.0f ? a - t : t - a;
    return u > 0.0f ? 1u : (d == 0.0f ? 1u : 0u);
}

struct trace_lane {
    float    px, py, pz;   /* position relative to the root centre */
    float    dx, dy, dz;
    float    dist;
    uint32_t steps;
    uint32_t cur;          /* current node, SVO_NONE once popped empty */
    uint32_t node;         /* cx | cy << 7 | cz << 14 | lvl << 21 | rc << 24 | dda << 27 */
    uint32_t leaf;         /* leaf index during the DDA */
    uint32_t vox;          /* x | y << 5 | z << 10 | it << 15 */
    float    tmx, tmy, tmz, tdx, tdy, tdz;
    unsigned long long bm; /* brick mask of the leaf */
    uint32_t wi, wv;       /* last bit word index and value */
};

#define ND_CX(n)  ((n) & 127u)
#define ND_CY(n)  (((n) >> 7) & 127u)
#define ND_CZ(n)  (((n) >> 14) & 127u)
#define ND_LVL(n) (((n) >> 21) & 7u)
#define ND_RC(n)  (((n) >> 24) & 7u)
#define ND_DDA    (1u << 27)
#define ND_ADV    (1u << 28)   /* an advance out of the child slot is pending */
#define ND_SETUP  (1u << 29)   /* entered a leaf: its DDA is set up at the lane's next burst */

__device__ static inline svo_res res_miss(uint32_t steps)
{
    svo_res r = {SVO_F32_MAX, SVO_NONE, 0u, 0u, steps, 0u};
    return r;
}

/* Ray setup (tg_sparse_voxel_octree.c:570-594) for the ray whose world-space
 * origin and direction the lane holds in px, py, pz and dx, dy, dz: false
 * if the root box is missed (the result is then a miss with 0 steps). */
__device__ static inline bool lane_init_self(const trace_params& P, float half, trace_lane& L, bool have_td = false)
{
    const float ox = L.px - P.cx, oy = L.py - P.cy, oz = L.pz - P.cz;
    L.steps = 0;
    L.cur = 0;
    L.node = 0;
    if (!have_td) {
        L.tdx = dda_len(1.0f, 1.0f, L.dx);
        L.tdy = dda_len(1.0f, 1.0f, L.dy);
        L.tdz = dda_len(1.0f, 1.0f, L.dz);
    }
    if (__all(fabsf(ox) < half && fabsf(oy) < half && fabsf(oz) < half)) {
        /* origins strictly inside the root box: every near-plane quotient is
         * <= 0 and the far one > 0, so the slab test below would keep the
         * origin with distance 0 */
        L.px = ox; L.py = oy; L.pz = oz;
        L.dist = 0.0f;
        return true;
    }
    const float enter = fmax2(fmax2(axis_near(ox, L.dx, -half, half), axis_near(oy, L.dy, -half, half)),
                              axis_near(oz, L.dz, -half, half));
    const float exit_ = fmin2(fmin2(axis_far(ox, L.dx, -half, half), axis_far(oy, L.dy, -half, half)),
                              axis_far(oz, L.dz, -half, half));
    if (!(exit_ > 0.0f && enter <= exit_)) return false;
    L.px = ox; L.py = oy; L.pz = oz;
    if (enter > 0.0f) { L.px = ox + L.dx * enter; L.py = oy + L.dy * enter; L.pz = oz + L.dz * enter; }
    L.dist = enter > 0.0f ? enter : 0.0f;
    return true;
}

template <bool kLds>
__device__ static inline uint2 inner_at(const uint2* __restrict__ g, const uint2* s, uint32_t i)
{
    if constexpr (kLds) return s[i];
    else return g[i];
}

/* Brick masks follow the inner nodes in LDS. */
template <bool kLds>
__device__ static inline unsigned long long bricks_at(const unsigned long long* __restrict__ g, const uint2* s,
                                                      uint32_t n_inner, uint32_t leaf)
{
    if constexpr (kLds) {
        const uint2 v = s[n_inner + leaf];
        return ((unsigned long long)v.y << 32) | v.x;
    } else {
        return g[leaf];
    }
}

__device__ static inline uint32_t lane_rank(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

/* Minimum corner of child slot rc of the node packed in n. */
__device__ static inline void child_min(uint32_t n, uint32_t rc, float half, float cs, float& x, float& y, float& z)
{
    x = -half + (float)(2u * ND_CX(n) + (rc & 1u)) * cs;
    y = -half + (float)(2u * ND_CY(n) + ((rc >> 1) & 1u)) * cs;
    z = -half + (float)(2u * ND_CZ(n) + (rc >> 2)) * cs;
}

/* Pops every node the position has left: a wave-uniform loop over the
 * lanes in `adv`, walking parent links; cur / n are the current node and
 * its packed coordinates. kDivide uses the dividing exit test (needed for
 * direction components below 2^-100). */
template <bool kLds, bool kDivide>
__device__ static inline void pop_walk(const trace_params& P, const uint2* s_inner, float half, const trace_lane& L,
                                       bool adv, uint32_t& cur, uint32_t& n)
{
    uint32_t popping = adv ? 1u : 0u;
    while (true) {
        const uint32_t live = popping & (cur != SVO_NONE ? 1u : 0u);
        if (!__any(live)) break;
        const uint32_t lvl = ND_LVL(n);
        const float ns = (float)(1u << (P.depth - lvl));
        const float nmx = -half + (float)ND_CX(n) * ns, nmy = -half + (float)ND_CY(n) * ns,
                    nmz = -half + (float)ND_CZ(n) * ns;
        uint32_t inside;
        if constexpr (kDivide)
            inside = (axis_exit_gt_eps(L.px, L.dx, nmx, nmx + ns) && axis_exit_gt_eps(L.py, L.dy, nmy, nmy + ns) &&
                      axis_exit_gt_eps(L.pz, L.dz, nmz, nmz + ns)) ? 1u : 0u;
        else
            inside = axis_exit_fast(L.px, L.dx, nmx, nmx + ns) & axis_exit_fast(L.py, L.dy, nmy, nmy + ns) &
                     axis_exit_fast(L.pz, L.dz, nmz, nmz + ns);
        const uint32_t up = live & (inside ^ 1u);
        const uint32_t par = inner_at<kLds>(P.inner, s_inner, live ? cur : 0u).y & 0xFFFFFFu;
        cur = up ? (par == 0xFFFFFFu ? SVO_NONE : par) : cur;
        n = up ? ((ND_CX(n) >> 1) | ((ND_CY(n) >> 1) << 7) | ((ND_CZ(n) >> 1) << 14) | (((lvl - 1u) & 7u) << 21)) : n;
        popping = up;
    }
}

/* Exact skip of an 8^3 brick holding no solid voxel: the state of
 * tg_amanatides_woo's 32^3 walk (tg_amanatides_woo.c:3-114) when its voxel
 * first leaves the brick of (x, y, z), with the reference's additions in
 * the reference's order. The walk steps the axis with the smallest t_max,
 * ties going to z, then y (tmx < tmy ? (tmx < tmz ? x : z) : (tmy < tmz ?
 * y : z)); so the steps inside the brick are the merge of the three axes'
 * t_max sequences V_a(j) = fl(V_a(j-1) + t_delta_a) ordered by (value,
 * z < y < x). Axis a leaves the brick at its N_a-th step, taken at
 * V_a(N_a - 1); the brick is left by the axis e whose exit event comes
 * first, and every other axis b has taken the steps whose events precede
 * it: those with V_b(j) < V_e(N_e - 1), or equal where b wins the tie. No
 * voxel of the brick is solid, so no step between stops the walk. The
 * caller guarantees |t_delta| < 2^100 on moving axes, so no t_max
 * overflows and a non-moving axis (t_max = F32_MAX) never steps. `moves`
 * gets the steps that stay inside the 32^3 block. */
__device__ static inline void brick_skip(int& x, int& y, int& z, float& tmx, float& tmy, float& tmz, float tdx,
                                         float tdy, float tdz, int sx, int sy, int sz, uint32_t& moves)
{
    const int nx = sx > 0 ? 8 - (x & 7) : (x & 7) + 1;
    const int ny = sy > 0 ? 8 - (y & 7) : (y & 7) + 1;
    const int nz = sz > 0 ? 8 - (z & 7) : (z & 7) + 1;
    float ex_x = tmx, ex_y = tmy, ex_z = tmz;
#pragma unroll
    for (int j = 1; j < 8; j++) {
        ex_x = j < nx ? ex_x + tdx : ex_x;
        ex_y = j < ny ? ex_y + tdy : ex_y;
        ex_z = j < nz ? ex_z + tdz : ex_z;
    }
    ex_x = sx ? ex_x : SVO_INF;
    ex_y = sy ? ex_y : SVO_INF;
    ex_z = sz ? ex_z : SVO_INF;
    const bool ez = ex_z <= ex_y && ex_z <= ex_x;
    const bool ey = !ez && ex_y <= ex_x;
    const bool ex = !ez && !ey;
    const float te = ez ? ex_z : (ey ? ex_y : ex_x);
    /* events of the other axes before the exit event; x wins no tie, y
     * wins a tie against x, z against x and y */
    float ux = tmx, uy = tmy, uz = tmz;
    uint32_t kx = 0, ky = 0, kz = 0;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const bool px = ux < te;
        const bool py = uy < te || (ex && uy == te);
        const bool pz = uz < te || (!ez && uz == te);
        ux = px ? ux + tdx : ux;
        uy = py ? uy + tdy : uy;
        uz = pz ? uz + tdz : uz;
        kx += px ? 1u : 0u;
        ky += py ? 1u : 0u;
        kz += pz ? 1u : 0u;
    }
    kx = ex ? (uint32_t)nx : kx;
    ky = ey ? (uint32_t)ny : ky;
    kz = ez ? (uint32_t)nz : kz;
    tmx = ex ? te + tdx : ux;
    tmy = ey ? te + tdy : uy;
    tmz = ez ? te + tdz : uz;
    x += sx * (int)kx;
    y += sy * (int)ky;
    z += sz * (int)kz;
    const uint32_t outside = (((uint32_t)x | (uint32_t)y | (uint32_t)z) >> 5) != 0u ? 1u : 0u;
    moves = kx + ky + kz - outside;
}

/* Crossing of an 8^3 brick without solid voxels for the uncounted DDA
 * (every moving axis' |t_delta| below 2^100), cheaper than brick_skip:
 * the exit axis comes from the approximate exit times
 * A_a = t_max_a + (N_a - 1) t_delta_a (one fma each). The sequential sums
 * V_a(N_a - 1) differ from them by at most 8 ulps, so when the runner-up
 * is later by more than 2^-19 of its value (and at least 2^-100) the axis
 * is certain; otherwise the lane does not cross (returns 0) and walks the
 * brick voxel by voxel. The exit axis' event time V_e(N_e - 1) is then
 * summed exactly, and each other axis b takes the steps whose events
 * precede it: V_b(j) < V_e, or <= where b wins the tie (z over y over x),
 * written as < the next float above V_e. */
__device__ static inline uint32_t brick_cross(int& x, int& y, int& z, float& tmx, float& tmy, float& tmz,
                                              float tdx, float tdy, float tdz, int sx, int sy, int sz)
{
    /* voxels left to the brick's far face on each axis, minus one:
     * 7 - (x & 7) moving up, x & 7 moving down */
    const uint32_t mx = ((uint32_t)~sx) >> 29, my = ((uint32_t)~sy) >> 29, mz = ((uint32_t)~sz) >> 29;
    const uint32_t rx = ((uint32_t)x ^ mx) & 7u, ry = ((uint32_t)y ^ my) & 7u, rz = ((uint32_t)z ^ mz) & 7u;
    const float ax = sx ? fmaf((float)rx, tdx, tmx) : SVO_INF;
    const float ay = sy ? fmaf((float)ry, tdy, tmy) : SVO_INF;
    const float az = sz ? fmaf((float)rz, tdz, tmz) : SVO_INF;
    const uint32_t ez = az <= ay && az <= ax ? 1u : 0u;
    const uint32_t ey = (ez ^ 1u) & (ay <= ax ? 1u : 0u);
    const uint32_t ex = (ez | ey) ^ 1u;
    const float ae = ez ? az : (ey ? ay : ax);
    /* the runner-up is the median of the three */
    const float second = __builtin_amdgcn_fmed3f(ax, ay, az);
    const uint32_t sure = (second == SVO_INF ? 1u : 0u) |
                          ((second >= 0x1p-100f ? 1u : 0u) & (second - ae > second * 0x1p-19f ? 1u : 0u));
    if (!sure) return 0u;
    const float tde = ez ? tdz : (ey ? tdy : tdx);
    const uint32_t re = ez ? rz : (ey ? ry : rx);
    float te = ez ? tmz : (ey ? tmy : tmx);
#pragma unroll
    for (uint32_t j = 0; j < 7u; j++) te = j < re ? te + tde : te;
    const float up = __uint_as_float(__float_as_uint(te) + 1u);
    /* the two other axes p < q (in x, y, z order) take the steps whose
     * events come first: below te, or at te where they win the tie (z over
     * y over x), written as below the next float above te */
    const float tdp = ex ? tdy : tdx, tdq = ez ? tdy : tdz;
    const float thp = ex ? up : te, thq = ez ? te : up;
    float up_ = ex ? tmy : tmx, uq = ez ? tmy : tmz;
    int kp = 0, kq = 0;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const bool cp = up_ < thp, cq = uq < thq;
        up_ = cp ? up_ + tdp : up_;
        uq = cq ? uq + tdq : uq;
        kp += cp ? 1 : 0;
        kq += cq ? 1 : 0;
    }
    tmx = ex ? te + tdx : up_;
    tmy = ey ? te + tdy : (ex ? up_ : uq);
    tmz = ez ? te + tdz : uq;
    x += sx * (ex ? (int)rx + 1 : kp);
    y += sy * (ey ? (int)ry + 1 : (ex ? kp : kq));
    z += sz * (ez ? (int)rz + 1 : kq);
    return 1u;
}

/* One step of a lane: up to node_iters iterations of the reference's node
 * loop while the lane descends, then up to DDA_BURST iterations of the leaf
 * DDA. Returns true when the ray is done; res is then its result. */
template <bool kLds, bool kSkip = false, uint32_t kDda = 0u>
__device__ static inline bool lane_step(const trace_params& P, const uint2* s_inner, float half, float maxd,
                                        trace_lane& L, svo_res& res, bool active, bool allow_node, bool allow_dda,
                                        bool defer_leave = false, uint32_t cross = 0u, uint32_t node_iters = 1u,
                                        uint32_t burst = DDA_BURST)
{
    bool adv = false, done = false;
    /* a lane with a pending advance waits for the wave's next advance */
    const bool pend = active && (L.node & ND_ADV);
    allow_node = allow_node && active && !pend;
    allow_dda = allow_dda && active && !pend;
    /* a lane that descended to an inner node may take the next node
     * iteration in the same step (wave-uniform loop) */
    bool go = allow_node && !(L.node & ND_DDA);
#pragma unroll 1
    for (uint32_t ni = 0; ni < node_iters; ni++) {
        if (ni != 0u && !__any(go)) break;
        if (go && (L.cur == SVO_NONE || L.steps >= TGX_SVO_MAX_STEPS || L.dist > maxd)) {
            res = res_miss(L.steps);
            done = true;
            go = false;
            allow_dda = false;
        }
        if (go) {
            /* one iteration of the reference's node loop */
            L.steps++;
            const uint2 nd = inner_at<kLds>(P.inner, s_inner, L.cur);
            const uint32_t n = L.node;
            const uint32_t lvl = ND_LVL(n);
            const float ps = (float)(1u << (P.depth - lvl));
            const float ce = ps * 0.5f;
            const float pmx = -half + (float)ND_CX(n) * ps, pmy = -half + (float)ND_CY(n) * ps,
                        pmz = -half + (float)ND_CZ(n) * ps;
            const float hx = pmx + ce, hy = pmy + ce, hz = pmz + ce;
            uint32_t rc = 0;
            if (hx < L.px || (L.px == hx && L.dx > 0.0f)) rc |= 1u;
            if (hy < L.py || (L.py == hy && L.dy > 0.0f)) rc |= 2u;
            if (hz < L.pz || (L.pz == hz && L.dz > 0.0f)) rc |= 4u;
            L.node = (n & ~(7u << 24)) | (rc << 24);
            adv = true;
            const uint32_t valid = nd.y >> 24;
            if (valid & (1u << rc)) {
                const uint32_t ci = nd.x + (uint32_t)__popc(valid & ((1u << rc) - 1u));
                if (lvl + 1u < P.leaf_level) {
                    L.cur = ci;
                    L.node = (2u * ND_CX(n) + (rc & 1u)) | ((2u * ND_CY(n) + ((rc >> 1) & 1u)) << 7) |
                             ((2u * ND_CZ(n) + (rc >> 2)) << 14) | ((lvl + 1u) << 21);
                    adv = false;
                } else {
                    const uint32_t leaf = ci - P.n_inner;
                    const unsigned long long bm = bricks_at<kLds>(P.bricks, s_inner, P.n_inner, leaf);
                    if (bm != 0ull) {
                        /* the DDA over the 32^3 block starts at the lane's
                         * next burst (its set-up runs once per step for all
                         * lanes that entered a leaf) */
                        L.bm = bm;
                        L.leaf = leaf;
                        L.node |= ND_DDA | ND_SETUP;
                        adv = false;
                    }
                }
            }
            go = !adv && !(L.node & ND_DDA);
        }
    }
    /* Up to DDA_BURST iterations of the leaf DDA. The loop is wave-uniform
     * (one __any exit) and lanes that stop are predicated off, which keeps
     * the scalar unit free of per-lane exit bookkeeping. */
    const bool in_dda = (L.node & ND_DDA) && allow_dda;
    if (__any(in_dda && (L.node & ND_SETUP))) {
        if (in_dda && (L.node & ND_SETUP)) {
            /* tg_amanatides_woo setup over the 32^3 block (the child slot's
             * box), entry voxel clamped */
            float cmx, cmy, cmz;
            child_min(L.node, ND_RC(L.node), half, 32.0f, cmx, cmy, cmz);
            const float gx = L.px - cmx, gy = L.py - cmy, gz = L.pz - cmz;
            const int x0 = (int)fmin2(fmax2(floorf(gx), 0.0f), 31.0f);
            const int y0 = (int)fmin2(fmax2(floorf(gy), 0.0f), 31.0f);
            const int z0 = (int)fmin2(fmax2(floorf(gz), 0.0f), 31.0f);
            L.tmx = dda_len((float)(x0 + 1) - gx, gx - (float)x0, L.dx);
            L.tmy = dda_len((float)(y0 + 1) - gy, gy - (float)y0, L.dy);
            L.tmz = dda_len((float)(z0 + 1) - gz, gz - (float)z0, L.dz);
            L.vox = (uint32_t)x0 | ((uint32_t)y0 << 5) | ((uint32_t)z0 << 10);
            L.wi = 0xFFFFFFFFu;
            L.node &= ~ND_SETUP;
        }
    }
    bool hit = false, leave = false;
    int x = (int)(L.vox & 31u), y = (int)((L.vox >> 5) & 31u), z = (int)((L.vox >> 10) & 31u);
    uint32_t it = L.vox >> 15;
    if (__any(in_dda)) {
        const uint32_t* grid = P.leaf_bits + (size_t)L.leaf * TGX_SVO_BLOCK_WORDS;
        const int sx = dir_step(L.dx), sy = dir_step(L.dy), sz = dir_step(L.dz);
        /* Lane conditions are kept as 0/1 integers in vector registers:
         * booleans combined with && / || would become 64-bit lane masks
         * combined on the CU's shared scalar unit, which then limits the
         * loop. Only single comparisons feed selects. */
        uint32_t live = in_dda ? 1u : 0u, hit_i = 0u, leave_i = 0u;
        /* The loop works on register copies of the lane state: a select
         * between two fields of L would be folded into a load through a
         * selected address, which forces L into scratch memory. */
        float tmx = L.tmx, tmy = L.tmy, tmz = L.tmz;
        const float tdx = L.tdx, tdy = L.tdy, tdz = L.tdz;
        const unsigned long long bm = L.bm;
        uint32_t wi_last = L.wi, wv = L.wv;
        /* brick skips need every moving axis' t_delta below 2^100 */
        const uint32_t skip_ok = (!kSkip || ((sx == 0 || tdx < 0x1p100f) && (sy == 0 || tdy < 0x1p100f) &&
                                             (sz == 0 || tdz < 0x1p100f))) ? 1u : 0u;
        /* With every moving axis' t_delta below 2^100 no t_max overflows,
         * so each step moves a coordinate and the walk leaves the 32^3
         * block within 94 steps: the iteration cap cannot bind, and a wave
         * whose DDA lanes all have such directions runs the walk without
         * counting iterations, two per exit test. A lane's count is then
         * left behind, which only lanes that never take this path read. */
        const bool small_td = (sx == 0 || tdx < 0x1p100f) && (sy == 0 || tdy < 0x1p100f) &&
                              (sz == 0 || tdz < 0x1p100f);
        if (kDda != 0u && !kSkip && __all(small_td || !in_dda)) {
            /* the count is not kept on this path (see above) */
            it = 0u;
#pragma unroll 1
            for (uint32_t k = 0; k < burst; k += 2) {
                if constexpr (kDda == 2u) {
                    /* once at least cross/64 of the wave's walking lanes
                     * are in bricks without solid voxels, those lanes cross
                     * their bricks (brick_cross) before the two steps */
                    const uint32_t b = ((uint32_t)x >> 3) | (((uint32_t)y >> 3) << 2) | (((uint32_t)z >> 3) << 4);
                    const uint32_t emp = live & ((uint32_t)(bm >> b) ^ 1u) & 1u;
                    const uint32_t n_emp = (uint32_t)__popcll(__ballot(emp != 0u));
                    const uint32_t n_live = (uint32_t)__popcll(__ballot(live != 0u));
                    if (64u * n_emp >= cross * n_live && n_emp != 0u) {
                        if (emp && brick_cross(x, y, z, tmx, tmy, tmz, tdx, tdy, tdz, sx, sy, sz)) {
                            const uint32_t outside =
                                (((uint32_t)x | (uint32_t)y | (uint32_t)z) >> 5) != 0u ? 1u : 0u;
                            leave_i |= outside;
                            live &= outside ^ 1u;
                        }
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < 2u; u++) {
                    const uint32_t b = ((uint32_t)x >> 3) | (((uint32_t)y >> 3) << 2) | (((uint32_t)z >> 3) << 4);
                    const uint32_t inb = live & (uint32_t)(bm >> b);
                    const uint32_t wi = 32u * (uint32_t)z + (uint32_t)y;
                    if ((inb & 1u) && wi != wi_last) { wi_last = wi; wv = grid[wi]; }
                    const uint32_t h = inb & (wv >> (uint32_t)x) & 1u;
                    const float m = fminf(fminf(tmx, tmy), tmz);
                    const uint32_t sz_ = tmz == m ? 1u : 0u;
                    const uint32_t sy_ = (tmy == m ? 1u : 0u) & (sz_ ^ 1u);
                    const uint32_t sx_ = (sz_ | sy_) ^ 1u;
                    const uint32_t move = live & (h ^ 1u);
                    const uint32_t mx = move & sx_, my = move & sy_, mz = move & sz_;
                    /* t_max += t_delta on the stepping axis as fma(1, td, t)
                     * (the same rounded sum) and fma(0, td, t) elsewhere (t
                     * itself, or +0 for a t of -0, which compares and adds
                     * alike) */
                    tmx = fmaf((float)mx, tdx, tmx);
                    tmy = fmaf((float)my, tdy, tmy);
                    tmz = fmaf((float)mz, tdz, tmz);
                    /* 24-bit multiply-adds (mx..mz are 0 or 1) */
                    x = x + __mul24((int)mx, sx);
                    y = y + __mul24((int)my, sy);
                    z = z + __mul24((int)mz, sz);
                    const uint32_t outside = (((uint32_t)x | (uint32_t)y | (uint32_t)z) >> 5) != 0u ? 1u : 0u;
                    const uint32_t lv = move & outside;
                    hit_i |= h;
                    leave_i |= lv;
                    live &= (h | lv) ^ 1u;
                }
                if (!__any(live != 0u)) break;
            }
        } else
#pragma unroll 1
        for (uint32_t k = 0; k < DDA_BURST; k++) {
            const uint32_t cap = it >> 8;
            uint32_t b = ((uint32_t)x >> 3) | (((uint32_t)y >> 3) << 2) | (((uint32_t)z >> 3) << 4);
            if constexpr (kSkip) {
                /* a lane in a brick without solid voxels crosses it in one
                 * go; it continues in the next brick with the step below */
                const uint32_t emp = live & (cap ^ 1u) & skip_ok & ((uint32_t)(bm >> b) ^ 1u) & 1u;
                if (emp) {
                    uint32_t mv;
                    brick_skip(x, y, z, tmx, tmy, tmz, tdx, tdy, tdz, sx, sy, sz, mv);
                    it += mv;
                    const uint32_t outside = (((uint32_t)x | (uint32_t)y | (uint32_t)z) >> 5) != 0u ? 1u : 0u;
                    leave_i |= outside;
                    live &= outside ^ 1u;
                    b = ((uint32_t)x >> 3) | (((uint32_t)y >> 3) << 2) | (((uint32_t)z >> 3) << 4);
                }
            }
            const uint32_t inb = live & (cap ^ 1u) & (uint32_t)(bm >> b);
            const uint32_t wi = 32u * (uint32_t)z + (uint32_t)y;
            if ((inb & 1u) && wi != wi_last) { wi_last = wi; wv = grid[wi]; }
            const uint32_t h = inb & (wv >> (uint32_t)x) & 1u;
            const float m = fminf(fminf(tmx, tmy), tmz);
            const uint32_t sz_ = tmz == m ? 1u : 0u;
            const uint32_t sy_ = (tmy == m ? 1u : 0u) & (sz_ ^ 1u);
            const uint32_t sx_ = (sz_ | sy_) ^ 1u;
            const uint32_t move = live & (cap ^ 1u) & (h ^ 1u);
            const uint32_t mx = move & sx_, my = move & sy_, mz = move & sz_;
            tmx = mx ? tmx + tdx : tmx;
            tmy = my ? tmy + tdy : tmy;
            tmz = mz ? tmz + tdz : tmz;
            x = mx ? x + sx : x;
            y = my ? y + sy : y;
            z = mz ? z + sz : z;
            const uint32_t outside = (((uint32_t)x | (uint32_t)y | (uint32_t)z) >> 5) != 0u ? 1u : 0u;
            const uint32_t lv = live & (cap | (move & outside));
            it += move & (outside ^ 1u);
            hit_i |= h;
            leave_i |= lv;
            live &= (h | lv) ^ 1u;
            if (!__any(live != 0u)) break;
        }
        L.tmx = tmx; L.tmy = tmy; L.tmz = tmz;
        L.wi = wi_last; L.wv = wv;
        hit = hit_i != 0u;
        leave = leave_i != 0u;
        if (in_dda && !hit && !leave) L.vox = (uint32_t)x | ((uint32_t)y << 5) | ((uint32_t)z << 10) | (it << 15);
        if (leave) {
            L.node &= ~ND_DDA;
            adv = true;
        }
        /* a lane that left its leaf may leave the advance to a later step
         * (ND_ADV), taken with other lanes in the same phase */
        if (defer_leave && leave) {
            L.node |= ND_ADV;
            adv = false;
        }
    }
    if (hit) {
        /* distance to the voxel's entry plane and the entry face */
        float vmx, vmy, vmz;
        child_min(L.node, ND_RC(L.node), half, 32.0f, vmx, vmy, vmz);
        vmx = vmx + (float)x; vmy = vmy + (float)y; vmz = vmz + (float)z;
        const float nx = axis_near(L.px, L.dx, vmx, vmx + 1.0f);
        const float ny = axis_near(L.py, L.dy, vmy, vmy + 1.0f);
        const float nz = axis_near(L.pz, L.dz, vmz, vmz + 1.0f);
        const float enter = fmax2(fmax2(nx, ny), nz);
        res.distance = L.dist + enter;
        res.node = P.n_inner + L.leaf;
        res.voxel = 1024u * (uint32_t)z + 32u * (uint32_t)y + (uint32_t)x;
        uint32_t ax = nx > ny ? 0u : 1u;
        const float nax = ax == 0u ? nx : ny;
        if (!(nax > nz)) ax = 2u;
        const uint32_t neg_x = L.dx > 0.0f ? 0u : 16u, neg_y = L.dy > 0.0f ? 0u : 16u, neg_z = L.dz > 0.0f ? 0u : 16u;
        res.flags = (res.distance <= maxd ? 1u : 0u) | 2u | (ax << 2) | (ax == 0u ? neg_x : (ax == 1u ? neg_y : neg_z));
        res.steps = L.steps;
        res.pad = 0u;
        done = true;
    }
    /* Advances run for many lanes at once: a lane's advance waits (ND_ADV)
     * until enough of the wave's active lanes want one. */
    const bool want = adv || pend;
    const uint32_t n_want = (uint32_t)__popcll(__ballot(want));
    const uint32_t n_act = (uint32_t)__popcll(__ballot(active));
    const bool run_adv = n_want >= TRACE_ADV_MIN || 2u * n_want >= n_act;
    if (!run_adv && adv) L.node |= ND_ADV;
    adv = run_adv && want;
    if (__any(adv)) {
        /* leave the child box, then pop every node the position has left
         * (wave-uniform loop, lanes predicated) */
        uint32_t n = L.node & ~ND_ADV;
        uint32_t cur = L.cur;
        if (adv) {
            const float cs = (float)(1u << (P.depth - ND_LVL(n) - 1u));
            float cmx, cmy, cmz;
            child_min(n, ND_RC(n), half, cs, cmx, cmy, cmz);
            const float t_adv = box_exit(L.px, L.py, L.pz, L.dx, L.dy, L.dz, cmx, cmy, cmz, cs) + SVO_EPSILON;
            L.px = L.px + L.dx * t_adv;
            L.py = L.py + L.dy * t_adv;
            L.pz = L.pz + L.dz * t_adv;
            L.dist += t_adv;
        }
        /* a wave with a direction component below 2^-100 in an advancing
         * lane takes the dividing pop test; all others the exact
         * division-free one */
        const bool tiny = adv && (fabsf(L.dx) < 0x1p-100f || fabsf(L.dy) < 0x1p-100f || fabsf(L.dz) < 0x1p-100f);
        if (__any(tiny))
            pop_walk<kLds, true>(P, s_inner, half, L, adv, cur, n);
        else
            pop_walk<kLds, false>(P, s_inner, half, L, adv, cur, n);
        if (adv) {
            L.cur = cur;
            L.node = n;
        }
    }
    return done;
}

/* Fused GI tracing (gi_trace_k<*, true>): the bounce queue traced with its
 * shadow rays in the same launch. When a lane's bounce ray hits a face that
 * faces the sun, the lane writes the sample record (weight, hit node and
 * voxel) and continues with the shadow ray itself; an occluded shadow ray
 * clears the weight again. Same arithmetic as gi_shade_k + gi_lit_k. The
 * records carry the flags gi_resolve_k counts bounce hits, shadow rays and
 * shadow hits from. */
/* Leaf materials are stored for solid voxels only, leaf by leaf in bit
 * order: solid voxel v of leaf l is entry mat_off[l] + (solid voxels of the
 * leaf's earlier words, wpre) + (solid voxels below v in its word). */
struct leaf_mat_view {
    const uint16_t* pool;
    const uint32_t* mat_off;
    const uint16_t* wpre;
    const uint32_t* bits;
};

__device__ static inline uint16_t leaf_material(const leaf_mat_view& M, uint32_t leaf, uint32_t voxel)
{
    const size_t w = (size_t)leaf * TGX_SVO_BLOCK_WORDS + (voxel >> 5);
    const uint32_t below = M.bits[w] & ((1u << (voxel & 31u)) - 1u);
    return M.pool[M.mat_off[leaf] + M.wpre[w] + (uint32_t)__popc(below)];
}

struct trace_fuse {
    float*               contrib;     /* 3 floats per sample: weight, hit node, voxel | flags */
    float                sun0, sun1, sun2;
    float                sun_td0, sun_td1, sun_td2;   /* 1 / |sun component| (F32_MAX for 0) */
    float                sun_radiance;
};

#define TRACE_SHADOW 0x80000000u   /* slot of a lane tracing a shadow ray: id | TRACE_SHADOW */
#define TRACE_SEGS 8u              /* queue segments of the lane-machine launch (one per XCD) */

/* Sample records of the GI trace: (weight, hit node, hit voxel | flags).
 * A bounce miss leaves (0, 0, 0); a hit sets SAMPLE_HIT; a sun-facing hit
 * writes its weight, node and voxel with SAMPLE_HIT | SAMPLE_LIT (its
 * shadow ray is traced); an occluded shadow ray clears the weight and sets
 * SAMPLE_OCC. The resolve pass of the fused trace counts bounce hits,
 * shadow rays and shadow hits from the flags. */
#define SAMPLE_HIT 0x80000000u
#define SAMPLE_LIT 0x40000000u
#define SAMPLE_OCC 0x20000000u
#define SAMPLE_VOXEL 0x1FFFFFFFu

/* Persistent traversal of queue rays[0 .. *n_rays) up to distance maxd.
 * Plain tracing (kFused false): the result of ray i goes to out[i] and hits
 * (flag bit 0) are added to *hit_counter when it is not null. Fused GI
 * (kFused true): see trace_fuse; out and hit_counter are not used. `cursor`
 * holds TRACE_SEGS segment cursors 16 words apart, 0 at launch. */
template <bool kLds, bool kFused, uint32_t kDda>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void gi_trace_k(trace_params P, const svo_ray* __restrict__ rays,
                                                          const uint32_t* __restrict__ n_rays_ptr, float maxd,
                                                          uint32_t* __restrict__ cursor, svo_res* __restrict__ out,
                                                          uint32_t* __restrict__ hit_counter, uint32_t refill,
                                                          uint32_t vote, uint32_t cross, uint32_t node_iters,
                                                          uint32_t burst, uint32_t chunk, uint32_t small_rays,
                                                          uint32_t small_blocks, const trace_fuse F)
{
    extern __shared__ uint2 s_inner[];
    /* a queue of fewer than small_rays rays is traced by the first
     * small_blocks workgroups (one per CU) alone: with few rays per lane the
     * launch is as long as its slowest rays, which run faster with fewer
     * waves per SIMD */
    const uint32_t n_rays = *n_rays_ptr;
    const bool small = n_rays < small_rays;
    if (small && blockIdx.x >= small_blocks) return;
    if constexpr (kLds) {
        for (uint32_t i = threadIdx.x; i < P.n_inner; i += blockDim.x) s_inner[i] = P.inner[i];
        const uint2* gb = (const uint2*)P.bricks;
        for (uint32_t i = threadIdx.x; i < P.n_leaves; i += blockDim.x) s_inner[P.n_inner + i] = gb[i];
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & 63u;
    /* fused GI: each lane's bounce ray origin and id, behind the nodes; the
     * wave's part of the stash is a uniform base indexed by lane */
    float4* s_ray = (float4*)(s_inner + (kLds ? ((P.n_inner + P.n_leaves + 1u) & ~1u) : 0u)) +
                    __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    const float half = P.half;
    trace_lane L = {};
    uint32_t slot = 0;
    /* the wave's bounce hits (plain trace) */
    uint32_t n_hit0 = 0;
    /* pending: the lane holds a ray (world origin in px..pz, direction in
     * dx..dz) that is set up at the next iteration's start */
    bool active = false, exhausted = false, pending = false;
    /* The wave's slice of the queue, rays [q_next, q_end) taken from the
     * cursor and not handed out yet: one atomic takes `chunk` rays for
     * several refills (same-address atomics of all waves serialise). Near
     * the queue's end (less than two chunks per wave left, as of the last
     * atomic) a wave takes only what it hands out, so no wave holds back
     * rays that idle waves could trace. */
    uint32_t q_next = 0, q_end = 0, seen = 0;
    bool drained = false;
    const uint32_t n_waves = (small && small_blocks < gridDim.x ? small_blocks : gridDim.x) * (TRACE_BLOCK / 64u);
    /* The queue is split into TRACE_SEGS segments with a cursor each (16
     * words apart). A wave starts in segment blockIdx.x mod TRACE_SEGS
     * (workgroups are dealt to the XCDs in turn, so an XCD's waves share a
     * segment and its L2 holds that part of the queue's data) and moves to
     * the next segment when its own runs dry. */
    uint32_t seg = blockIdx.x % TRACE_SEGS, seg_left = TRACE_SEGS;
    /* segment k is rays [n k / TRACE_SEGS, n (k + 1) / TRACE_SEGS) */
    auto seg_start = [n_rays](uint32_t k) { return (uint32_t)((uint64_t)n_rays * k / TRACE_SEGS); };
    seen = seg_start(seg);
    const uint32_t seg_waves = (n_waves + TRACE_SEGS - 1u) / TRACE_SEGS;
    while (true) {
        /* wave-uniform refill once `refill` lanes are idle */
        const unsigned long long idle = __ballot(!active && !pending);
        const uint32_t want = (uint32_t)__popcll(idle);
        if (!exhausted && want >= refill) {
            const uint32_t avail = q_end - q_next;
            uint32_t nb = 0, nend = 0;
            if (avail < want && !drained) {
                const uint32_t seg_lo = seg_start(seg), seg_hi = seg_start(seg + 1u);
                const uint32_t need = want - avail;
                const bool tail = (uint64_t)(seg_hi - (seen < seg_hi ? seen : seg_hi)) <
                                  2ull * seg_waves * chunk;
                const uint32_t take = tail || chunk < need ? need : chunk;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(cursor + 16u * seg, take);
                /* the whole wave runs this block, so lane 0 is the first active */
                base = __builtin_amdgcn_readfirstlane(base) + seg_lo;
                seen = base;
                if (base < seg_hi) {
                    nb = base;
                    nend = (uint64_t)base + take < seg_hi ? base + take : seg_hi;
                }
                if ((uint64_t)base + take >= seg_hi) {
                    /* this segment has run dry: on to the next one */
                    if (--seg_left == 0u) {
                        drained = true;
                    } else {
                        seg = seg + 1u == TRACE_SEGS ? 0u : seg + 1u;
                        seen = seg_start(seg);
                    }
                }
            }
            const uint32_t rk = lane_rank(idle);
            if (!active && !pending) {
                const uint32_t r = rk < avail ? q_next + rk : nb + (rk - avail);
                if (rk < avail || rk - avail < nend - nb) {
                    slot = r;
                    const svo_ray q = rays[r];
                    L.px = q.ox; L.py = q.oy; L.pz = q.oz;
                    L.dx = q.dx; L.dy = q.dy; L.dz = q.dz;
                    if constexpr (kFused) s_ray[lane] = make_float4(q.ox, q.oy, q.oz, __uint_as_float(q.id));
                    pending = true;
                }
            }
            if (avail >= want) {
                q_next += want;
            } else {
                const uint32_t used = want - avail;
                q_next = nend - nb > used ? nb + used : nend;
                q_end = nend;
            }
            exhausted = drained && q_next == q_end;
        }
        if (pending) {
            pending = false;
            /* a shadow ray's step lengths are the sun's, set with its ray */
            active = lane_init_self(P, half, L, kFused && (slot & TRACE_SHADOW));
            if (!kFused && !active) out[slot] = res_miss(0u);
        }
        if (!__any(active || pending)) {
            if (exhausted) break;
            continue;
        }
        /* phase vote (tuning): a wave runs only DDA bursts in an iteration
         * once at least vote/64 of its active lanes are in a leaf DDA, and
         * only node steps otherwise */
        bool allow_node = true, allow_dda = true;
        if (vote) {
            const uint32_t n_dda = (uint32_t)__popcll(__ballot(active && (L.node & ND_DDA)));
            const uint32_t n_all = (uint32_t)__popcll(__ballot(active));
            allow_dda = 64u * n_dda >= vote * n_all;
            allow_node = !allow_dda;
        }
        svo_res res;
        /* per-lane events of this iteration, counted below at a uniform point:
         * 1 bounce hit, 2 shadow ray started, 4 shadow hit; << 3 for the
         * second list */
        uint32_t ev = 0u;
        if (lane_step<kLds, false, kDda>(P, s_inner, half, maxd, L, res, active, allow_node, allow_dda, false,
                                         cross, node_iters, burst) &&
            active) {
            active = false;
            if constexpr (!kFused) {
                out[slot] = res;
                ev = res.flags & 1u;
            } else if (slot & TRACE_SHADOW) {
                const uint32_t id = slot & ~TRACE_SHADOW;
                if (res.flags & 1u) {
                    /* occluded: the sample contributes nothing (weight 0) */
                    F.contrib[3ull * id] = 0.0f;
                    F.contrib[3ull * id + 2] = __uint_as_float(SAMPLE_HIT | SAMPLE_LIT | SAMPLE_OCC);
                }
            } else {
                /* the ray's origin and id from LDS; its direction is L.dx..dz */
                const float4 o = s_ray[lane];
                const uint32_t id = __float_as_uint(o.w);
                /* cosine of the hit face with the sun (gi_shade_k) */
                const uint32_t ax = (res.flags >> 2) & 3u;
                const float sign = (res.flags & 16u) ? 1.0f : -1.0f;
                const float sc = ax == 0u ? F.sun0 : (ax == 1u ? F.sun1 : F.sun2);
                const float cosv = ((res.flags & 1u) && !(res.distance < 0.0f)) ? sign * sc : 0.0f;
                if (cosv > 0.0f) {
                    /* the shadow ray goes into the lane's own (dead) ray
                     * registers and is set up with the next pending rays; a
                     * shadow ray that misses the root box is unoccluded */
                    L.px = o.x + L.dx * res.distance;
                    L.py = o.y + L.dy * res.distance;
                    L.pz = o.z + L.dz * res.distance;
                    const float off = sign * GI_SHADOW_OFFSET;
                    if (ax == 0u) L.px = L.px + off; else if (ax == 1u) L.py = L.py + off; else L.pz = L.pz + off;
                    L.dx = F.sun0; L.dy = F.sun1; L.dz = F.sun2;
                    L.tdx = F.sun_td0; L.tdy = F.sun_td1; L.tdz = F.sun_td2;
                    /* the sample record: weight, hit node and voxel (the
                     * albedo is applied when the samples are resolved) */
                    const float wgt = cosv * F.sun_radiance * GI_INV_PI;
                    float* c = F.contrib + 3ull * id;
                    c[0] = wgt; c[1] = __uint_as_float(res.node); c[2] = __uint_as_float(res.voxel | SAMPLE_HIT | SAMPLE_LIT);
                    slot = id | TRACE_SHADOW;
                    pending = true;
                } else if (res.flags & 1u) {
                    F.contrib[3ull * id + 2] = __uint_as_float(SAMPLE_HIT);
                }
            }
        }
        if constexpr (!kFused) {
            if (__any(ev != 0u)) n_hit0 += (uint32_t)__popcll(__ballot(ev & 1u));
        }
    }
    if constexpr (!kFused) {
        if (lane == 0 && hit_counter && n_hit0) atomicAdd(hit_counter, n_hit0);
    }
}

/* ------------------------------------------------------------------------ */
/* Regrouped fused GI tracing                                               */
/* ------------------------------------------------------------------------ */

/* gi_trace_rg_k runs the fused GI launch of gi_trace_k<*, true> with the
 * same per-ray arithmetic, but keeps the waves of a workgroup in one phase:
 * after every step the workgroup sorts its rays by phase (leaf DDA first,
 * then node walk, then idle lanes) through LDS, so a wave runs DDA bursts
 * for (nearly) all of its lanes or node steps and advances for (nearly) all
 * of them, instead of both at partial width. A lane leaving its leaf leaves
 * the advance to the node phase (ND_ADV). Idle lanes collect in the last
 * waves, which refill from the queue 64 rays per atomic. One 1024-lane
 * workgroup per CU; LDS holds the traversal layout (when it fits) and six
 * float4 planes of lane state. */
#define RG_BLOCK 1024u
#define RG_WAVES (RG_BLOCK / 64u)
#define RG_PLANES 6u

struct rg_ctl {
    uint32_t cnt[RG_WAVES][2];   /* per wave: DDA lanes, node-walk lanes */
    uint32_t exhausted;
    uint32_t pad[3];
};

/* Result of a lane's finished ray in the fused GI trace (same operations as
 * gi_trace_k<*, true>): a sun-facing bounce hit stores its sample record and
 * turns the lane into its shadow ray (pending set-up); another hit stores
 * flags SAMPLE_HIT; an occluded shadow ray clears its weight and sets SAMPLE_OCC. */
__device__ static inline void fused_result(const trace_params& P, const trace_fuse& F, trace_lane& L,
                                               const svo_res& res, uint32_t& slot, float ox, float oy, float oz,
                                               uint32_t sid, bool& pending)
{
    if (slot & TRACE_SHADOW) {
        const uint32_t id = slot & ~TRACE_SHADOW;
        if (res.flags & 1u) {
            F.contrib[3ull * id] = 0.0f;
            F.contrib[3ull * id + 2] = __uint_as_float(SAMPLE_HIT | SAMPLE_LIT | SAMPLE_OCC);
        }
        return;
    }
    const uint32_t ax = (res.flags >> 2) & 3u;
    const float sign = (res.flags & 16u) ? 1.0f : -1.0f;
    const float sc = ax == 0u ? F.sun0 : (ax == 1u ? F.sun1 : F.sun2);
    const float cosv = ((res.flags & 1u) && !(res.distance < 0.0f)) ? sign * sc : 0.0f;
    if (cosv > 0.0f) {
        L.px = ox + L.dx * res.distance;
        L.py = oy + L.dy * res.distance;
        L.pz = oz + L.dz * res.distance;
        const float off = sign * GI_SHADOW_OFFSET;
        if (ax == 0u) L.px = L.px + off; else if (ax == 1u) L.py = L.py + off; else L.pz = L.pz + off;
        L.dx = F.sun0; L.dy = F.sun1; L.dz = F.sun2;
        const float wgt = cosv * F.sun_radiance * GI_INV_PI;
        float* c = F.contrib + 3ull * sid;
        c[0] = wgt; c[1] = __uint_as_float(res.node); c[2] = __uint_as_float(res.voxel | SAMPLE_HIT | SAMPLE_LIT);
        slot = sid | TRACE_SHADOW;
        pending = true;
    } else if (res.flags & 1u) {
        F.contrib[3ull * sid + 2] = __uint_as_float(SAMPLE_HIT);
    }
}

template <bool kLds, bool kSkip>
__global__ __launch_bounds__(RG_BLOCK, 1) void gi_trace_rg_k(trace_params P, const svo_ray* __restrict__ rays,
                                                             const uint32_t* __restrict__ n_rays_ptr, float maxd,
                                                             uint32_t* __restrict__ cursor, uint32_t refill,
                                                             uint32_t node_iters, const trace_fuse F)
{
    extern __shared__ uint2 s_inner[];
    const uint32_t lds_nodes = kLds ? ((P.n_inner + P.n_leaves + 1u) & ~1u) : 0u;
    float4* planes = (float4*)(s_inner + lds_nodes);
    rg_ctl* ctl = (rg_ctl*)(planes + RG_PLANES * RG_BLOCK);
    if constexpr (kLds) {
        for (uint32_t i = threadIdx.x; i < P.n_inner; i += blockDim.x) s_inner[i] = P.inner[i];
        const uint2* gb = (const uint2*)P.bricks;
        for (uint32_t i = threadIdx.x; i < P.n_leaves; i += blockDim.x) s_inner[P.n_inner + i] = gb[i];
    }
    if (threadIdx.x == 0) ctl->exhausted = 0u;
    __syncthreads();
    const uint32_t n_rays = *n_rays_ptr;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const float half = P.half;
    trace_lane L = {};
    uint32_t slot = 0, sid = 0;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    bool active = false, pending = false;
    while (true) {
        /* refill: idle lanes sit in the last waves after a regroup */
        const unsigned long long idle = __ballot(!active && !pending);
        const uint32_t want = (uint32_t)__popcll(idle);
        if (want >= refill && *(volatile uint32_t*)&ctl->exhausted == 0u) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(cursor, want);
            base = __builtin_amdgcn_readfirstlane(base);
            if (base + want >= n_rays && lane == 0) *(volatile uint32_t*)&ctl->exhausted = 1u;
            if (!active && !pending) {
                const uint32_t r = base + lane_rank(idle);
                if (r < n_rays) {
                    slot = r;
                    const svo_ray q = rays[r];
                    L.px = q.ox; L.py = q.oy; L.pz = q.oz;
                    L.dx = q.dx; L.dy = q.dy; L.dz = q.dz;
                    ox = q.ox; oy = q.oy; oz = q.oz; sid = q.id;
                    pending = true;
                }
            }
        }
        if (pending) {
            pending = false;
            active = lane_init_self(P, half, L);
        }
        /* one step of each lane's phase; node-walk lanes take up to
         * node_iters steps while the wave has node work */
        const bool was_dda = active && (L.node & ND_DDA);
        for (uint32_t k = 0; k < node_iters; k++) {
            const bool step_dda = k == 0 && was_dda;
            const bool step_node = active && !was_dda && !(L.node & ND_DDA);
            if (!__any(step_dda || step_node)) break;
            svo_res res;
            if (lane_step<kLds, kSkip>(P, s_inner, half, maxd, L, res, active && (step_dda || step_node), step_node,
                                       step_dda, true) &&
                active && (step_dda || step_node)) {
                active = false;
                fused_result(P, F, L, res, slot, ox, oy, oz, sid, pending);
            }
        }
        /* regroup: DDA lanes first, then node-walk lanes (pending shadow
         * rays included), then idle lanes */
        const bool key_d = active && (L.node & ND_DDA);
        const bool key_n = (active || pending) && !key_d;
        const unsigned long long bd = __ballot(key_d), bn = __ballot(key_n);
        if (lane == 0) {
            ctl->cnt[wave][0] = (uint32_t)__popcll(bd);
            ctl->cnt[wave][1] = (uint32_t)__popcll(bn);
        }
        __syncthreads();
        uint32_t td = 0, tn = 0, pd = 0, pn = 0;
        for (uint32_t w = 0; w < RG_WAVES; w++) {
            const uint32_t cd = ctl->cnt[w][0], cn = ctl->cnt[w][1];
            pd += w < wave ? cd : 0u;
            pn += w < wave ? cn : 0u;
            td += cd;
            tn += cn;
        }
        const bool done_all = ctl->exhausted != 0u && td + tn == 0u;
        if (done_all) break;
        if (key_d || key_n) {
            const uint32_t dst = key_d ? pd + lane_rank(bd) : td + pn + lane_rank(bn);
            const uint32_t flags = (active ? 1u : 0u) | (pending ? 2u : 0u);
            planes[0 * RG_BLOCK + dst] = make_float4(L.px, L.py, L.pz, L.dist);
            planes[1 * RG_BLOCK + dst] = make_float4(L.dx, L.dy, L.dz, __uint_as_float(L.steps));
            planes[2 * RG_BLOCK + dst] = make_float4(__uint_as_float(L.cur), __uint_as_float(L.node),
                                                     __uint_as_float(L.leaf), __uint_as_float(L.vox));
            planes[3 * RG_BLOCK + dst] = make_float4(L.tmx, L.tmy, L.tmz, __uint_as_float(slot));
            planes[4 * RG_BLOCK + dst] = make_float4(L.tdx, L.tdy, L.tdz, __uint_as_float(flags));
            planes[5 * RG_BLOCK + dst] = make_float4(ox, oy, oz, __uint_as_float(sid));
        }
        __syncthreads();
        const uint32_t me = threadIdx.x;
        if (me < td + tn) {
            const float4 a = planes[0 * RG_BLOCK + me], b = planes[1 * RG_BLOCK + me];
            const float4 c = planes[2 * RG_BLOCK + me], d = planes[3 * RG_BLOCK + me];
            const float4 e = planes[4 * RG_BLOCK + me], f = planes[5 * RG_BLOCK + me];
            L.px = a.x; L.py = a.y; L.pz = a.z; L.dist = a.w;
            L.dx = b.x; L.dy = b.y; L.dz = b.z; L.steps = __float_as_uint(b.w);
            L.cur = __float_as_uint(c.x); L.node = __float_as_uint(c.y);
            L.leaf = __float_as_uint(c.z); L.vox = __float_as_uint(c.w);
            L.tmx = d.x; L.tmy = d.y; L.tmz = d.z; slot = __float_as_uint(d.w);
            L.tdx = e.x; L.tdy = e.y; L.tdz = e.z;
            const uint32_t flags = __float_as_uint(e.w);
            active = (flags & 1u) != 0u;
            pending = (flags & 2u) != 0u;
            ox = f.x; oy = f.y; oz = f.z; sid = __float_as_uint(f.w);
            /* the leaf's brick mask and an empty bit-word cache */
            if (active && (L.node & ND_DDA)) L.bm = bricks_at<kLds>(P.bricks, s_inner, P.n_inner, L.leaf);
            L.wi = 0xFFFFFFFFu;
        } else {
            active = false;
            pending = false;
        }
        /* the planes and counts are written again only after the next
         * round's first barrier */
    }
}

/* ------------------------------------------------------------------------ */
/* One-bounce GI                                                            */
/* ------------------------------------------------------------------------ */

struct gi_cam {
    float pos[3];
    float bl[3], br[3], tr[3], tl[3];
};

__global__ void gi_fill_k(uint32_t* p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* xorshift32 and its [lo, hi] float (util.inc:11-31) */
__device__ static inline float gi_rand(uint32_t& s, float lo, float hi)
{
    uint32_t r = s;
    r ^= r << 13; r ^= r >> 17; r ^= r << 5;
    s = r;
    const float u = (float)r / (float)0xFFFFFFFFu;
    return u * (hi - lo) + lo;
}
__device__ static inline f3 gi_cosine_direction(f3 n, uint32_t& st)
{
    /* Up to 32 attempts, each drawing a point of the cube and keeping it if
     * it lies in the unit ball (and n plus its direction is not ~0). The
     * draws run in an inner loop until a lane's point is kept, so the
     * normalisation runs once per kept point instead of in every pass of a
     * loop whose lanes accept at different passes; the draws and results
     * are the same. */
    int attempt = 0;
    while (attempt < 32) {
        float x, y, z, l2;
        do {
            x = gi_rand(st, -1.0f, 1.0f);
            y = gi_rand(st, -1.0f, 1.0f);
            z = gi_rand(st, -1.0f, 1.0f);
            l2 = x * x + y * y + z * z;
            attempt++;
        } while ((l2 > 1.0f || l2 < 1e-6f) && attempt < 32);
        if (l2 > 1.0f || l2 < 1e-6f) break;
        const float l = sqrtf(l2);
        const float dx = n.x + x / l, dy = n.y + y / l, dz = n.z + z / l;
        const float dl2 = dx * dx + dy * dy + dz * dz;
        if (dl2 < 1e-6f) continue;
        const float dl = sqrtf(dl2);
        f3 r = {dx / dl, dy / dl, dz / dl};
        return r;
    }
    return n;
}

__device__ static inline float gi_sign(float a) { return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f); }

__device__ static inline bool gi_slab(f3 o, f3 d, f3 mn, f3 mx, float& enter)
{
    const float nx = axis_near(o.x, d.x, mn.x, mx.x), ny = axis_near(o.y, d.y, mn.y, mx.y), nz = axis_near(o.z, d.z, mn.z, mx.z);
    const float fx = axis_far(o.x, d.x, mn.x, mx.x), fy = axis_far(o.y, d.y, mn.y, mx.y), fz = axis_far(o.z, d.z, mn.z, mx.z);
    enter = fmax2(fmax2(nx, ny), nz);
    const float exit_ = fmin2(fmin2(fx, fy), fz);
    return exit_ > 0.0f && enter <= exit_;
}

/* Primary surface of a visibility word (shading.frag:116-229). */
__device__ static bool gi_surface(const tgx_object* __restrict__ objects, const uint32_t* __restrict__ cptrs,
                                  const uint32_t* __restrict__ c2o, const uint8_t* __restrict__ materials,
                                  const tgx_lut_entry* __restrict__ luts, const gi_cam& cam, uint32_t W, uint32_t H,
                                  uint32_t px, uint32_t py, uint64_t word, f3& p, f3& n, f3& albedo)
{
    const float depth = (float)(word >> 40) / 16777215.0f;
    if (!(depth < 1.0f)) return false;
    const uint32_t cptr = (uint32_t)(word >> 9) & 0x7FFFFFFFu;
    const uint32_t v = (uint32_t)word & 511u;
    const uint32_t cidx = cptrs[cptr];
    const tgx_object ob = objects[c2o[cidx]];
    const uint32_t rel = cptr - ob.first_cluster_pointer;
    const uint32_t nx = ob.n_cluster_pointers_per_dim[0], ny = ob.n_cluster_pointers_per_dim[1];
    const uint32_t rx = rel % nx, ry = (rel / nx) % ny, rz = rel / (nx * ny);
    const tgx_m4 w2c = tgx_world_to_cluster(ob.rotation, ob.translation, ob.n_cluster_pointers_per_dim, rx, ry, rz);
    const tgx_v4 o4 = tgx_m4_mulv4(w2c, cam.pos[0], cam.pos[1], cam.pos[2], 1.0f);
    const float fx = ((float)px + 0.5f) / (float)W;
    const float fy = 1.0f - ((float)py + 0.5f) / (float)H;
    const float omy = 1.0f - fy, omx = 1.0f - fx;
    const float lx = cam.bl[0] * omy + cam.tl[0] * fy, ly = cam.bl[1] * omy + cam.tl[1] * fy, lz = cam.bl[2] * omy + cam.tl[2] * fy;
    const float qx = cam.br[0] * omy + cam.tr[0] * fy, qy = cam.br[1] * omy + cam.tr[1] * fy, qz = cam.br[2] * omy + cam.tr[2] * fy;
    const float dwx = lx * omx + qx * fx, dwy = ly * omx + qy * fx, dwz = lz * omx + qz * fx;
    const tgx_v4 d4 = tgx_m4_mulv4(w2c, dwx, dwy, dwz, 0.0f);
    const float dm = sqrtf(d4.x * d4.x + d4.y * d4.y + d4.z * d4.z);
    const f3 dms = {d4.x / dm, d4.y / dm, d4.z / dm};
    const f3 orig = {o4.x, o4.y, o4.z};
    const f3 vmn = {(float)(v % 8u), (float)((v / 8u) % 8u), (float)(v / 64u)};
    const f3 vmx = {vmn.x + 1.0f, vmn.y + 1.0f, vmn.z + 1.0f};
    float enter;
    if (!gi_slab(orig, dms, vmn, vmx, enter)) return false;
    f3 hit = orig;
    if (enter > 0.0f) { hit.x = orig.x + dms.x * enter; hit.y = orig.y + dms.y * enter; hit.z = orig.z + dms.z * enter; }
    f3 nn = {hit.x - (vmn.x + 0.5f), hit.y - (vmn.y + 0.5f), hit.z - (vmn.z + 0.5f)};
    const float ax = fabsf(nn.x), ay = fabsf(nn.y), az = fabsf(nn.z);
    if (ax > ay) {
        nn.y = 0.0f;
        if (ax > az) { nn.x = gi_sign(nn.x); nn.z = 0.0f; } else { nn.z = gi_sign(nn.z); nn.x = 0.0f; }
    } else {
        nn.x = 0.0f;
        if (ay > az) { nn.y = gi_sign(nn.y); nn.z = 0.0f; } else { nn.z = gi_sign(nn.z); nn.y = 0.0f; }
    }
    tgx_m4 rot;
    for (int i = 0; i < 16; i++) rot.m[i] = ob.rotation[i];
    const tgx_v4 n4 = tgx_m4_mulv4(rot, nn.x, nn.y, nn.z, 0.0f);
    const float nm = sqrtf(n4.x * n4.x + n4.y * n4.y + n4.z * n4.z);
    n.x = n4.x / nm; n.y = n4.y / nm; n.z = n4.z / nm;
    const tgx_m4 c2w = tgx_cluster_to_world(ob.rotation, ob.translation, ob.n_cluster_pointers_per_dim, rx, ry, rz);
    const tgx_v4 p4 = tgx_m4_mulv4(c2w, hit.x, hit.y, hit.z, 1.0f);
    p.x = p4.x; p.y = p4.y; p.z = p4.z;
    const uint32_t mat = materials ? (uint32_t)materials[(size_t)cidx * 512u + v] : 0u;
    const tgx_lut_entry e = luts[(size_t)ob.lut_idx * TGX_LUT_ENTRIES + mat];
    albedo.x = e.albedo[0]; albedo.y = e.albedo[1]; albedo.z = e.albedo[2];
    return true;
}

#define GI_BLOCK 256u
#define SHADE_ITEMS 4u

/* Surface pass: the primary hit of every pixel (shading.frag:116-229) and
 * its visibility flag. Outputs of pixels without a hit are cleared. */
__global__ __launch_bounds__(GI_BLOCK) void gi_surface_k(const tgx_object* __restrict__ objects,
                                                         const uint32_t* __restrict__ cptrs,
                                                         const uint32_t* __restrict__ c2o,
                                                         const uint8_t* __restrict__ materials,
                                                         const tgx_lut_entry* __restrict__ luts, gi_cam cam,
                                                         uint32_t W, uint32_t H, const uint64_t* __restrict__ vis,
                                                         gi_surf_rec* __restrict__ surf, uint32_t* __restrict__ flag,
                                                         float* __restrict__ radiance, float* __restrict__ irradiance)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= W * H) return;
    f3 p = {0, 0, 0}, n = {0, 0, 0}, alb = {0, 0, 0};
    const bool ok = gi_surface(objects, cptrs, c2o, materials, luts, cam, W, H, pix % W, pix / W, vis[pix], p, n, alb);
    gi_surf_rec r;
    r.p[0] = p.x; r.p[1] = p.y; r.p[2] = p.z;
    r.n[0] = n.x; r.n[1] = n.y; r.n[2] = n.z;
    r.albedo[0] = alb.x; r.albedo[1] = alb.y; r.albedo[2] = alb.z;
    r.valid = ok ? 1u : 0u;
    surf[pix] = r;
    flag[pix] = ok ? 1u : 0u;
    if (!ok) {
        if (radiance) { radiance[3ull * pix] = 0.0f; radiance[3ull * pix + 1] = 0.0f; radiance[3ull * pix + 2] = 0.0f; }
        if (irradiance) { irradiance[3ull * pix] = 0.0f; irradiance[3ull * pix + 1] = 0.0f; irradiance[3ull * pix + 2] = 0.0f; }
    }
}

/* Visible pixels in ascending order from the flags and their exclusive scan. */
__global__ void gi_list_k(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off, uint32_t n,
                          uint32_t* __restrict__ list, uint32_t* __restrict__ count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) list[off[i]] = i;
    if (i == n - 1) *count = off[i] + flag[i];
}

/* spp cosine bounce rays per listed pixel, from p + sqrt(3) n, one thread
 * per (entry, sample); ray id = id_base + list entry * spp + sample
 * (id_base separates the lists that share a queue). Entry i's rays take the
 * queue slots after the tail counters[1] had at launch, so the launch needs
 * no atomics and a wave writes 64 consecutive rays; gi_gen_tail_k then
 * moves the tail and the list's ray count. */
__global__ __launch_bounds__(GI_BLOCK) void gi_gen_k(const gi_surf_rec* __restrict__ surf,
                                                     const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ count_ptr, tgx_gi_params prm,
                                                     const uint32_t* __restrict__ frame_salt,
                                                     svo_ray* __restrict__ rays, float* __restrict__ contrib,
                                                     const uint32_t* __restrict__ counters, uint32_t id_base,
                                                     uint32_t salt_add, uint32_t max_count)
{
    /* a device counter can advance the frame of replayed work */
    if (frame_salt) prm.frame += 0x632BE5ABu * (*frame_salt + salt_add + 1u);
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = j / prm.spp, k = j - i * prm.spp;
    if (i >= *count_ptr || i >= max_count) return;
    /* entry i's samples take slots [tail + i spp, + spp) after the queue's
     * tail at launch; gi_gen_tail_k advances the tail afterwards */
    const uint32_t pix = list[i];
    const gi_surf_rec sr = surf[pix];
    const f3 n = {sr.n[0], sr.n[1], sr.n[2]};
    const f3 ob = {sr.p[0] + GI_SQRT3 * n.x, sr.p[1] + GI_SQRT3 * n.y, sr.p[2] + GI_SQRT3 * n.z};
    uint32_t st = gi_fmix32(pix * prm.spp + k + 0x9E3779B9u * (prm.frame + 1u));
    if (st == 0u) st = 0x6D2B79F5u;
    const f3 d = gi_cosine_direction(n, st);
    const svo_ray r = {ob.x, ob.y, ob.z, d.x, d.y, d.z, id_base + j, 0u};
    rays[counters[1] + j] = r;
    const size_t c = 3ull * ((size_t)id_base + (size_t)j);
    contrib[c] = 0.0f; contrib[c + 1] = 0.0f; contrib[c + 2] = 0.0f;
}

/* Advances the bounce queue's tail (and the list's generated-ray count)
 * past one gi_gen_k launch: min(count, max_count) entries of spp rays. */
__global__ void gi_gen_tail_k(const uint32_t* __restrict__ count_ptr, uint32_t max_count, uint32_t spp,
                              uint32_t* __restrict__ counters, uint32_t* __restrict__ stats)
{
    const uint32_t c = *count_ptr;
    const uint32_t n = (c < max_count ? c : max_count) * spp;
    counters[1] += n;
    if (stats) stats[1] += n;
}

/* Cosine of a bounce hit's entry face with the sun; > 0 when the hit takes
 * part in the estimate (hits at negative distance started inside a solid
 * voxel and count as occluded). */
__device__ static inline float shade_cos(const svo_res& r, const tgx_gi_params& prm)
{
    if (!(r.flags & 1u) || r.distance < 0.0f) return 0.0f;
    const uint32_t ax = (r.flags >> 2) & 3u;
    const float sign = (r.flags & 16u) ? 1.0f : -1.0f;
    const float sc = ax == 0u ? prm.sun_dir[0] : (ax == 1u ? prm.sun_dir[1] : prm.sun_dir[2]);
    return sign * sc;
}

/* Bounce hits that face the sun: the sample record (weight, hit node and
 * voxel) and a shadow ray (started 1/64 voxel off the hit face). SHADE_ITEMS rays per thread.
 * With two lists in the queue (stats0 and stats1 not null), each list's
 * bounce hits and shadow rays are counted apart; ids below split belong to
 * the first. */
__global__ __launch_bounds__(GI_BLOCK) void gi_shade_k(uint32_t n_inner, leaf_mat_view mats,
                                                       const tgx_lut_entry* __restrict__ luts, tgx_gi_params prm,
                                                       const svo_ray* __restrict__ rays,
                                                       const uint32_t* __restrict__ n_rays_ptr,
                                                       const svo_res* __restrict__ res, svo_ray* __restrict__ shadow,
                                                       float* __restrict__ weight, uint32_t* __restrict__ counters,
                                                       uint32_t split, uint32_t* __restrict__ stats0,
                                                       uint32_t* __restrict__ stats1)
{
    const uint32_t n_rays = *n_rays_ptr;
    const uint32_t first = blockIdx.x * GI_BLOCK * SHADE_ITEMS + threadIdx.x;
    uint32_t emit = 0, count = 0, hits1 = 0, hits0 = 0, emit1 = 0;
    for (uint32_t k = 0; k < SHADE_ITEMS; k++) {
        const uint32_t i = first + k * GI_BLOCK;
        if (i >= n_rays) continue;
        const svo_res r = res[i];
        const bool second = stats0 && rays[i].id >= split;
        if (r.flags & 1u) { if (second) hits1++; else hits0++; }
        if (shade_cos(r, prm) > 0.0f) {
            emit |= 1u << k;
            count++;
            if (second) emit1++;
        }
    }
    uint32_t slot = block_append<GI_BLOCK>(&counters[3], count);
    if (stats0) {
        (void)block_append<GI_BLOCK>(&stats0[2], hits0);
        (void)block_append<GI_BLOCK>(&stats0[3], count - emit1);
        (void)block_append<GI_BLOCK>(&stats1[2], hits1);
        (void)block_append<GI_BLOCK>(&stats1[3], emit1);
    }
    for (uint32_t k = 0; k < SHADE_ITEMS; k++) {
        if (!(emit & (1u << k))) continue;
        const uint32_t i = first + k * GI_BLOCK;
        const svo_res r = res[i];
        const float cosv = shade_cos(r, prm);
        const uint32_t ax = (r.flags >> 2) & 3u;
        const float sign = (r.flags & 16u) ? 1.0f : -1.0f;
        const svo_ray b = rays[i];
        float ox = b.ox + b.dx * r.distance, oy = b.oy + b.dy * r.distance, oz = b.oz + b.dz * r.distance;
        const float off = sign * GI_SHADOW_OFFSET;
        if (ax == 0u) ox = ox + off; else if (ax == 1u) oy = oy + off; else oz = oz + off;
        const svo_ray sr = {ox, oy, oz, prm.sun_dir[0], prm.sun_dir[1], prm.sun_dir[2], b.id, slot};
        shadow[slot] = sr;
        const float wgt = cosv * prm.sun_radiance * GI_INV_PI;
        weight[3ull * slot] = wgt;
        weight[3ull * slot + 1] = __uint_as_float(r.node);
        weight[3ull * slot + 2] = __uint_as_float(r.voxel);
        slot++;
    }
}

/* Unshadowed samples store their sample record at (entry, sample). */
__global__ __launch_bounds__(256) void gi_lit_k(const svo_ray* __restrict__ shadow, const uint32_t* __restrict__ n_shadow,
                                                const svo_res* __restrict__ res, const float* __restrict__ weight,
                                                float* __restrict__ contrib)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= *n_shadow || (res[j].flags & 1u)) return;
    const svo_ray r = shadow[j];
    const size_t c = 3ull * r.id, w = 3ull * r.pad; /* pad = the ray's weight slot */
    contrib[c] = weight[w]; contrib[c + 1] = weight[w + 1]; contrib[c + 2] = weight[w + 2];
}

/* Zero stand-ins for the material look-ups of samples with weight 0, so
 * the look-ups of a chunk of samples are issued without branches and
 * overlap in flight. */
static __device__ uint32_t g_zero_u32[1];
static __device__ uint16_t g_zero_u16[1];
static __device__ tgx_lut_entry g_zero_lut;

/* The contributions of samples k0 .. k0 + RESOLVE_CHUNK - 1 of an entry
 * whose records start at c: the albedo of the hit voxel's material times
 * the weight, 0 for weight 0 and for samples at or past spp. Adds the
 * record flags of the samples below spp to n_hit, n_sh and n_occ. */
#define RESOLVE_CHUNK 4u
__device__ static inline void sample_values(const float* __restrict__ c, uint32_t k0, uint32_t spp, uint32_t n_inner,
                                            const leaf_mat_view& M, const tgx_lut_entry* __restrict__ luts,
                                            float v[RESOLVE_CHUNK][3], uint32_t& n_hit, uint32_t& n_sh,
                                            uint32_t& n_occ)
{
    float w[RESOLVE_CHUNK];
    uint32_t leaf[RESOLVE_CHUNK], vox[RESOLVE_CHUNK];
#pragma unroll
    for (uint32_t j = 0; j < RESOLVE_CHUNK; j++) {
        const bool in = k0 + j < spp;
        const uint32_t k = in ? k0 + j : spp - 1u;
        const float wr = c[3 * k];
        const uint32_t f = __float_as_uint(c[3 * k + 2]);
        w[j] = in ? wr : 0.0f;
        leaf[j] = __float_as_uint(c[3 * k + 1]) - n_inner;
        vox[j] = f & SAMPLE_VOXEL;
        n_hit += in ? f >> 31 : 0u;
        n_sh += in ? (f >> 30) & 1u : 0u;
        n_occ += in ? (f >> 29) & 1u : 0u;
    }
    uint32_t bw[RESOLVE_CHUNK], off[RESOLVE_CHUNK], pre[RESOLVE_CHUNK];
#pragma unroll
    for (uint32_t j = 0; j < RESOLVE_CHUNK; j++) {
        const bool live = w[j] != 0.0f;
        const size_t wi = (size_t)leaf[j] * TGX_SVO_BLOCK_WORDS + (vox[j] >> 5);
        bw[j] = *(live ? M.bits + wi : g_zero_u32);
        off[j] = *(live ? M.mat_off + leaf[j] : g_zero_u32);
        pre[j] = *(live ? M.wpre + wi : g_zero_u16);
    }
    uint32_t mat[RESOLVE_CHUNK];
#pragma unroll
    for (uint32_t j = 0; j < RESOLVE_CHUNK; j++) {
        const bool live = w[j] != 0.0f;
        const uint32_t below = bw[j] & ((1u << (vox[j] & 31u)) - 1u);
        mat[j] = *(live ? M.pool + off[j] + pre[j] + (uint32_t)__popc(below) : g_zero_u16);
    }
#pragma unroll
    for (uint32_t j = 0; j < RESOLVE_CHUNK; j++) {
        const bool live = w[j] != 0.0f;
        const tgx_lut_entry* e =
            live ? luts + ((size_t)(mat[j] >> 8) * TGX_LUT_ENTRIES + (mat[j] & 255u)) : &g_zero_lut;
        const float al0 = e->albedo[0], al1 = e->albedo[1], al2 = e->albedo[2];
        v[j][0] = live ? al0 * w[j] : 0.0f;
        v[j][1] = live ? al1 * w[j] : 0.0f;
        v[j][2] = live ? al2 * w[j] : 0.0f;
    }
}

/* One list entry's estimate from its spp sample records, summed in sample
 * order. */
__device__ static inline void resolve_entry(const gi_surf_rec* __restrict__ surf, const uint32_t* __restrict__ list,
                                            uint32_t i, const float* __restrict__ contrib, uint32_t spp,
                                            uint32_t n_inner, const leaf_mat_view& mats,
                                            const tgx_lut_entry* __restrict__ luts, float* __restrict__ e_entry,
                                            float* __restrict__ e_var, float* __restrict__ radiance,
                                            float* __restrict__ irradiance, uint32_t& n_hit, uint32_t& n_sh,
                                            uint32_t& n_occ)
{
    const uint32_t pix = list[i];
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    const float* c = contrib + 3ull * (size_t)i * spp;
    for (uint32_t k0 = 0; k0 < spp; k0 += RESOLVE_CHUNK) {
        float v[RESOLVE_CHUNK][3];
        sample_values(c, k0, spp, n_inner, mats, luts, v, n_hit, n_sh, n_occ);
#pragma unroll
        for (uint32_t j = 0; j < RESOLVE_CHUNK; j++)
            if (k0 + j < spp) { a0 = a0 + v[j][0]; a1 = a1 + v[j][1]; a2 = a2 + v[j][2]; }
    }
    const float inv = 1.0f / (float)spp;
    const float e0 = a0 * inv, e1 = a1 * inv, e2 = a2 * inv;
    if (e_entry) { e_entry[3ull * i] = e0; e_entry[3ull * i + 1] = e1; e_entry[3ull * i + 2] = e2; }
    if (e_var) {
        float q = 0.0f;
        uint32_t d0 = 0, d1 = 0, d2 = 0;
        for (uint32_t k0 = 0; k0 < spp; k0 += RESOLVE_CHUNK) {
            float v[RESOLVE_CHUNK][3];
            sample_values(c, k0, spp, n_inner, mats, luts, v, d0, d1, d2);
#pragma unroll
            for (uint32_t j = 0; j < RESOLVE_CHUNK; j++)
                if (k0 + j < spp) {
                    const float x0 = v[j][0] - e0, x1 = v[j][1] - e1, x2 = v[j][2] - e2;
                    q += x0 * x0 + x1 * x1 + x2 * x2;
                }
        }
        e_var[i] = spp > 1u ? q / ((float)spp * (float)(spp - 1u)) : 0.0f;
    }
    if (radiance) {
        const gi_surf_rec s = surf[pix];
        radiance[3ull * pix] = s.albedo[0] * e0;
        radiance[3ull * pix + 1] = s.albedo[1] * e1;
        radiance[3ull * pix + 2] = s.albedo[2] * e2;
    }
    if (irradiance) { irradiance[3ull * pix] = e0; irradiance[3ull * pix + 1] = e1; irradiance[3ull * pix + 2] = e2; }
}

__device__ static inline uint32_t wave_sum(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}

/* Per list entry: E = (sum of its sample contributions in sample order) /
 * spp, radiance = albedo * E; grid-stride over the entries. With `stats`
 * (records of the fused trace), the list's bounce hits, shadow rays and
 * shadow hits are added to stats[2], stats[3] and stats[7] from the record
 * flags, one atomic per block and counter (same-address atomics from every
 * wave would serialise). */
#define RESOLVE_BLOCK 1024u
#define RESOLVE_MAX_BLOCKS 256u
__global__ __launch_bounds__(RESOLVE_BLOCK) void gi_resolve_k(const gi_surf_rec* __restrict__ surf,
                                                    const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ count_ptr, uint32_t max_count,
                                                    const float* __restrict__ contrib, uint32_t spp,
                                                    uint32_t n_inner, leaf_mat_view mats,
                                                    const tgx_lut_entry* __restrict__ luts,
                                                    float* __restrict__ e_entry, float* __restrict__ e_var,
                                                    float* __restrict__ radiance, float* __restrict__ irradiance,
                                                    uint32_t* __restrict__ stats)
{
    const uint32_t count = *count_ptr;
    const uint32_t n = count < max_count ? count : max_count;
    uint32_t n_hit = 0, n_sh = 0, n_occ = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        resolve_entry(surf, list, i, contrib, spp, n_inner, mats, luts, e_entry, e_var, radiance, irradiance, n_hit,
                      n_sh, n_occ);
    }
    if (stats) {
        __shared__ uint32_t s_cnt[RESOLVE_BLOCK / 64u][3];
        n_hit = wave_sum(n_hit);
        n_sh = wave_sum(n_sh);
        n_occ = wave_sum(n_occ);
        const uint32_t w = threadIdx.x >> 6;
        if ((threadIdx.x & 63u) == 0u) { s_cnt[w][0] = n_hit; s_cnt[w][1] = n_sh; s_cnt[w][2] = n_occ; }
        __syncthreads();
        if (threadIdx.x < 3u) {
            uint32_t t = 0;
            for (uint32_t k = 0; k < RESOLVE_BLOCK / 64u; k++) t += s_cnt[k][threadIdx.x];
            if (t) atomicAdd(&stats[threadIdx.x == 0u ? 2 : threadIdx.x == 1u ? 3 : 7], t);
        }
    }
}

/* Sort key of a queued ray: its direction octant, then the Morton code of
 * the 8-voxel cell of its origin (relative to the SVO root, 256 cells per
 * axis), so rays that leave one region in similar directions end up in the
 * same wave. Slots past the queue get the largest key. */
__device__ static inline uint32_t spread8(uint32_t v)
{
    v &= 0xFFu;
    v = (v | (v << 8)) & 0x0000F00Fu;
    v = (v | (v << 4)) & 0x000C30C3u;
    v = (v | (v << 2)) & 0x00249249u;
    return v;
}

__global__ void gi_ray_keys_k(const svo_ray* __restrict__ rays, const uint32_t* __restrict__ n_ptr, uint32_t cap,
                              float mnx, float mny, float mnz, float cell_scale, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ idx)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    idx[i] = i;
    if (i >= *n_ptr) { keys[i] = 0xFFFFFFFFu; return; }
    const svo_ray r = rays[i];
    const uint32_t oct = (r.dx < 0.0f ? 1u : 0u) | (r.dy < 0.0f ? 2u : 0u) | (r.dz < 0.0f ? 4u : 0u);
    const uint32_t cx = (uint32_t)fminf(fmaxf((r.ox - mnx) * cell_scale, 0.0f), 255.0f);
    const uint32_t cy = (uint32_t)fminf(fmaxf((r.oy - mny) * cell_scale, 0.0f), 255.0f);
    const uint32_t cz = (uint32_t)fminf(fmaxf((r.oz - mnz) * cell_scale, 0.0f), 255.0f);
    keys[i] = (oct << 24) | spread8(cx) | (spread8(cy) << 1) | (spread8(cz) << 2);
}

__global__ void gi_gather_rays_k(const svo_ray* __restrict__ src, const uint32_t* __restrict__ idx,
                                 const uint32_t* __restrict__ n_ptr, svo_ray* __restrict__ dst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *n_ptr) dst[i] = src[idx[i]];
}

/* svo_res -> tgx_svo_hit for the C ABI */
__global__ void svo_hits_k(const svo_gnode* __restrict__ nodes, const svo_res* __restrict__ res, uint32_t n,
                           tgx_svo_hit* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const svo_res r = res[i];
    tgx_svo_hit h;
    h.hit = r.flags & 1u;
    h.distance = r.distance;
    h.block[0] = h.block[1] = h.block[2] = 0u;
    h.voxel = r.voxel;
    h.normal[0] = h.normal[1] = h.normal[2] = 0;
    if (r.flags & 2u) {
        const svo_gnode cn = nodes[r.node];
        h.block[0] = cn.cx; h.block[1] = cn.cy; h.block[2] = cn.cz;
        h.normal[(r.flags >> 2) & 3u] = (r.flags & 16u) ? 1 : -1;
    }
    h.steps = r.steps;
    out[i] = h;
}

/* SVO primary visibility (debug_visibility_svo.frag:27-70): one ray per
 * pixel from the camera along the normalised corner-ray lerp
 * (tg_lerp_corner_ray_directions, common.inc:48-53). */
__global__ void svo_primary_rays_k(gi_cam cam, uint32_t W, uint32_t H, svo_ray* __restrict__ rays)
{
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= W * H) return;
    const uint32_t px = pix % W, py = pix / W;
    const float fx = ((float)px + 0.5f) / (float)W;
    const float fy = 1.0f - ((float)py + 0.5f) / (float)H;
    const float omy = 1.0f - fy, omx = 1.0f - fx;
    const float lx = cam.bl[0] * omy + cam.tl[0] * fy, ly = cam.bl[1] * omy + cam.tl[1] * fy, lz = cam.bl[2] * omy + cam.tl[2] * fy;
    const float qx = cam.br[0] * omy + cam.tr[0] * fy, qy = cam.br[1] * omy + cam.tr[1] * fy, qz = cam.br[2] * omy + cam.tr[2] * fy;
    const float dx = lx * omx + qx * fx, dy = ly * omx + qy * fx, dz = lz * omx + qz * fx;
    const float m = sqrtf(dx * dx + dy * dy + dz * dz);
    svo_ray r = {cam.pos[0], cam.pos[1], cam.pos[2], dx / m, dy / m, dz / m, pix, 0u};
    rays[pix] = r;
}

/* The visibility word of an SVO hit (debug_visibility_svo.frag:59-68):
 * depth24 = u64(d 16777215) for d = distance / far <= 1, then the leaf
 * node's index (canonical order) in place of the cluster pointer and the
 * block voxel index mod 512. */
__global__ void svo_vis_pack_k(const svo_res* __restrict__ res, uint32_t n, float far_plane, uint64_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const svo_res r = res[i];
    uint64_t word = TGX_VISIBILITY_EMPTY;
    if (r.flags & 2u) {
        const float d = r.distance / far_plane;
        if (d <= 1.0f)
            word = ((uint64_t)(d * 16777215.0f) << 40) | ((uint64_t)(r.node & 0x7FFFFFFFu) << 9) | (uint64_t)(r.voxel % 512u);
    }
    out[i] = word;
}

/* ------------------------------------------------------------------------ */
/* Host side                                                                */
/* ------------------------------------------------------------------------ */

static trace_params params_of(tgx_context* ctx)
{
    trace_params P;
    P.inner = (const uint2*)ctx->svo_inner.ptr;
    P.bricks = (const unsigned long long*)ctx->svo_bricks.ptr;
    P.leaf_bits = (const uint32_t*)ctx->svo_leaf_bits.ptr;
    P.n_inner = ctx->svo_n_inner;
    P.n_leaves = ctx->svo_info.n_leaves;
    P.depth = ctx->svo_info.depth;
    P.leaf_level = ctx->svo_info.depth - 5u;
    P.mnx = ctx->svo_info.min[0]; P.mny = ctx->svo_info.min[1]; P.mnz = ctx->svo_info.min[2];
    /* the kernels' former in-kernel expressions, in the same fp32 order */
    P.half = (float)(1u << P.depth) * 0.5f;
    P.cx = P.mnx + P.half; P.cy = P.mny + P.half; P.cz = P.mnz + P.half;
    return P;
}

/* Persistent trace of the queue rays[0 .. *n_dev) (n_max bounds it):
 * 1024-lane workgroups (tuning: per CU), the traversal layout in LDS when it
 * fits; the queue-segment cursors live in ctx->gi_counts. */
static int launch_trace(tgx_context* ctx, const svo_ray* rays, const uint32_t* n_dev, uint32_t n_max, float maxd,
                        svo_res* out, uint32_t* hit_counter, const trace_fuse* fuse = nullptr)
{
    const trace_fuse F = fuse ? *fuse : trace_fuse{};
    if (n_max == 0) return TGX_OK;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    const uint32_t want = (uint32_t)((n_max + TRACE_BLOCK - 1) / TRACE_BLOCK);
    const uint32_t cap = (uint32_t)cus * ctx->tuning.trace_blocks_per_cu;
    const dim3 grid(want < cap ? want : cap);
    const trace_params P = params_of(ctx);
    const uint32_t refill = ctx->tuning.trace_refill;
    const uint32_t vote = ctx->tuning.trace_vote;
    /* with more than one workgroup per CU, queues below trace_small rays
     * per lane of one workgroup per CU run on one workgroup per CU */
    const uint64_t small64 = grid.x > (uint32_t)cus ? (uint64_t)ctx->tuning.trace_small * TRACE_BLOCK * (uint32_t)cus : 0u;
    const uint32_t small_rays = small64 < 0xFFFFFFFFull ? (uint32_t)small64 : 0xFFFFFFFFu;
    {
        tgx_launch_scope s(ctx, "gi_trace_k");
        /* dynamic LDS: nodes and brick masks (when they fit), then for the
         * fused GI trace one 16-B ray stash per lane */
        const bool fused = F.contrib != nullptr;
        const uint32_t stash = fused ? 16u * TRACE_BLOCK : 0u;
        const bool lds = (uint64_t)P.n_inner + P.n_leaves <= TRACE_LDS_NODES_MAX;
        const uint32_t shmem = (lds ? 8u * ((P.n_inner + P.n_leaves + 1u) & ~1u) : 0u) + stash;
        /* DDA form: 0 counted, 1 uncounted, 2 uncounted with brick crossing */
        const uint32_t dda = ctx->tuning.trace_dda == 0u ? 0u : (ctx->tuning.trace_cross ? 2u : 1u);
        decltype(&gi_trace_k<true, true, 0u>) kernel;
        if (lds)
            kernel = fused ? (dda == 2u ? gi_trace_k<true, true, 2u> : dda ? gi_trace_k<true, true, 1u> : gi_trace_k<true, true, 0u>)
                           : (dda == 2u ? gi_trace_k<true, false, 2u> : dda ? gi_trace_k<true, false, 1u> : gi_trace_k<true, false, 0u>);
        else
            kernel = fused ? (dda == 2u ? gi_trace_k<false, true, 2u> : dda ? gi_trace_k<false, true, 1u> : gi_trace_k<false, true, 0u>)
                           : (dda == 2u ? gi_trace_k<false, false, 2u> : dda ? gi_trace_k<false, false, 1u> : gi_trace_k<false, false, 0u>);
        /* the segment cursors start at 0 (a fill kernel, so a captured
         * frame clears them on every replay) */
        uint32_t* seg_cursors = (uint32_t*)ctx->gi_counts.ptr + GI_CNT_SEG;
        hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(128), 0, ctx->stream, seg_cursors, 16ull * TRACE_SEGS, 0u);
        hipLaunchKernelGGL(kernel, grid, dim3(TRACE_BLOCK), shmem, ctx->stream, P, rays, n_dev, maxd, seg_cursors, out,
                           hit_counter, refill, vote, ctx->tuning.trace_cross, ctx->tuning.trace_node,
                           ctx->tuning.trace_burst, ctx->tuning.trace_chunk, small_rays, (uint32_t)cus, F);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_trace_k");
    return TGX_OK;
}

static leaf_mat_view leaf_mats_of(const tgx_context* ctx)
{
    return {(const uint16_t*)ctx->svo_leaf_mats.ptr, (const uint32_t*)ctx->svo_leaf_moff.ptr,
            (const uint16_t*)ctx->svo_leaf_wpre.ptr, (const uint32_t*)ctx->svo_leaf_bits.ptr};
}

/* The fused GI launch as regrouped waves (gi_trace_rg_k): one 1024-lane
 * workgroup per CU, node_iters node steps per round. */
static int launch_trace_rg(tgx_context* ctx, const svo_ray* rays, const uint32_t* n_dev, uint32_t n_max, float maxd,
                           uint32_t* cursor, const trace_fuse& F, uint32_t node_iters)
{
    if (n_max == 0) return TGX_OK;
    int cus = 256, max_lds = 65536;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    (void)hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    const uint32_t want = (uint32_t)((n_max + RG_BLOCK - 1) / RG_BLOCK);
    const dim3 grid(want < (uint32_t)cus ? want : (uint32_t)cus);
    const trace_params P = params_of(ctx);
    const size_t state = sizeof(float4) * RG_PLANES * RG_BLOCK + sizeof(rg_ctl);
    const size_t nodes = 8u * (size_t)((P.n_inner + P.n_leaves + 1u) & ~1u);
    const bool lds = (uint64_t)P.n_inner + P.n_leaves <= TRACE_LDS_NODES_MAX && nodes + state <= (size_t)max_lds;
    if (state > (size_t)max_lds)
        return tgx_fail(ctx, TGX_ERR_CAPACITY, "regrouped trace needs %zu B of LDS per workgroup, the device has %d",
                        state, max_lds);
    const size_t shmem = (lds ? nodes : 0u) + state;
    const bool skip = ctx->tuning.trace_skip != 0u;
    auto kernel = lds ? (skip ? gi_trace_rg_k<true, true> : gi_trace_rg_k<true, false>)
                      : (skip ? gi_trace_rg_k<false, true> : gi_trace_rg_k<false, false>);
    /* the LDS limit is raised once per variant, before any capture */
    const uint32_t variant = 1u << ((lds ? 2u : 0u) + (skip ? 1u : 0u));
    if (!(ctx->rg_lds_set & variant)) {
        TGX_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               max_lds));
        ctx->rg_lds_set |= variant;
    }
    {
        tgx_launch_scope s(ctx, "gi_trace_k");
        hipLaunchKernelGGL(kernel, grid, dim3(RG_BLOCK), shmem, ctx->stream, P, rays, n_dev, maxd, cursor,
                           ctx->tuning.trace_refill, node_iters, F);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_trace_rg_k");
    return TGX_OK;
}

/* Reorders the ray queue rays[0 .. *n_dev) into `sorted` by gi_ray_keys_k
 * (cap bounds the count). */
static int sort_rays(tgx_context* ctx, const svo_ray* rays, const uint32_t* n_dev, uint32_t cap, svo_ray* sorted)
{
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->gi_sort, 16ull * cap, "gi ray sort"))) return rc;
    uint32_t* keys = (uint32_t*)ctx->gi_sort.ptr;
    uint32_t* idx = keys + cap;
    uint32_t* keys2 = idx + cap;
    uint32_t* idx2 = keys2 + cap;
    hipLaunchKernelGGL(gi_ray_keys_k, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, rays, n_dev, cap,
                       ctx->svo_info.min[0], ctx->svo_info.min[1], ctx->svo_info.min[2],
                       256.0f / (float)(1u << ctx->svo_info.depth), keys, idx);
    TGX_LAUNCH_CHECK(ctx, "gi_ray_keys_k");
    if ((rc = tgx_prim_sort_pairs_u32(ctx, keys, keys2, idx, idx2, cap, 27u))) return rc;
    {
        tgx_launch_scope s(ctx, "gi_gather_rays_k");
        hipLaunchKernelGGL(gi_gather_rays_k, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, rays, idx2, n_dev, sorted);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_gather_rays_k");
    return TGX_OK;
}

int tgx_gi_validate(tgx_context* ctx, uint32_t w, uint32_t h, const tgx_gi_params* prm)
{
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene uploaded");
    int rc;
    if ((rc = tgx_svo_check_current(ctx))) return rc;
    if (!ctx->luts.ptr || ctx->n_luts == 0) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "the scene has no LUTs");
    if (prm->spp < 1 || prm->spp > 64 || !(prm->max_distance > 0.0f))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "spp must be in [1, 64] and max_distance > 0");
    if (w == 0 || h == 0 || (uint64_t)w * h * prm->spp >= (1ull << 31))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "bad resolution");
    return TGX_OK;
}

int tgx_gi_surfaces_at(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis_dev,
                       float* radiance_dev, float* irradiance_dev, gi_surf_rec* surf, uint32_t* list_scratch,
                       uint32_t* count)
{
    const uint64_t n_pix = (uint64_t)w * h;
    int rc;
    gi_cam gc;
    for (int i = 0; i < 3; i++) {
        gc.pos[i] = cam->camera[i];
        gc.bl[i] = cam->ray_bl[i]; gc.br[i] = cam->ray_br[i]; gc.tr[i] = cam->ray_tr[i]; gc.tl[i] = cam->ray_tl[i];
    }
    uint32_t* list = list_scratch;
    uint32_t* flag = list + n_pix;
    uint32_t* off = flag + n_pix;
    {
        tgx_launch_scope s(ctx, "gi_surface_k");
        hipLaunchKernelGGL(gi_surface_k, dim3(blocks_for(n_pix)), dim3(GI_BLOCK), 0, ctx->stream,
                           (const tgx_object*)ctx->objects.ptr, (const uint32_t*)ctx->cluster_pointers.ptr,
                           (const uint32_t*)ctx->cluster_idx_to_object.ptr, tgx_scene_materials(ctx),
                           (const tgx_lut_entry*)ctx->luts.ptr, gc, w, h, vis_dev, surf, flag, radiance_dev,
                           irradiance_dev);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_surface_k");
    if ((rc = tgx_prim_exclusive_scan_u32(ctx, flag, off, (uint32_t)n_pix))) return rc;
    hipLaunchKernelGGL(gi_list_k, dim3(blocks_for(n_pix)), dim3(256), 0, ctx->stream, flag, off, (uint32_t)n_pix, list,
                       count);
    TGX_LAUNCH_CHECK(ctx, "gi_list_k");
    return TGX_OK;
}

int tgx_gi_surfaces(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis_dev,
                    float* radiance_dev, float* irradiance_dev)
{
    const uint64_t n_pix = (uint64_t)w * h;
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->gi_counts, 4ull * GI_CNT_TOTAL, "gi counters")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_surf, sizeof(gi_surf_rec) * n_pix, "gi surfaces")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_list, 12ull * n_pix, "gi pixel list")))
        return rc;
    return tgx_gi_surfaces_at(ctx, cam, w, h, vis_dev, radiance_dev, irradiance_dev, (gi_surf_rec*)ctx->gi_surf.ptr,
                              (uint32_t*)ctx->gi_list.ptr, (uint32_t*)ctx->gi_counts.ptr + GI_CNT_VISIBLE);
}

int tgx_gi_trace_sets(tgx_context* ctx, const gi_trace_set* sets, uint32_t n_sets)
{
    if (n_sets == 0 || n_sets > 1024) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "1 to 1024 trace sets");
    uint64_t n_samples = 0;
    std::vector<uint32_t> id_base(n_sets);
    for (uint32_t k = 0; k < n_sets; k++) {
        id_base[k] = (uint32_t)n_samples;
        n_samples += (uint64_t)sets[k].max_count * sets[k].prm.spp;
    }
    if (n_samples >= (1ull << 31)) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%llu samples exceed 2^31",
                                                   (unsigned long long)n_samples);
    const bool fused = !(ctx->tuning.trace_sort & 2u);
    int rc;
    /* result records, shadow queues and weights only for the queued form */
    if ((rc = tgx_buffer_reserve(ctx, &ctx->gi_rays, sizeof(svo_ray) * n_samples, "gi bounce rays")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_contrib, 12ull * n_samples, "gi contributions")))
        return rc;
    if (!fused &&
        ((rc = tgx_buffer_reserve(ctx, &ctx->gi_res, sizeof(svo_res) * n_samples, "gi bounce results")) ||
         (rc = tgx_buffer_reserve(ctx, &ctx->gi_shadow, sizeof(svo_ray) * n_samples, "gi shadow rays")) ||
         (rc = tgx_buffer_reserve(ctx, &ctx->gi_shadow_res, sizeof(svo_res) * n_samples, "gi shadow results")) ||
         (rc = tgx_buffer_reserve(ctx, &ctx->gi_weight, 12ull * n_samples, "gi weights"))))
        return rc;
    svo_ray* rays = (svo_ray*)ctx->gi_rays.ptr;
    svo_res* res = (svo_res*)ctx->gi_res.ptr;
    svo_ray* shadow = (svo_ray*)ctx->gi_shadow.ptr;
    svo_res* sres = (svo_res*)ctx->gi_shadow_res.ptr;
    float* weight = (float*)ctx->gi_weight.ptr;
    float* contrib = (float*)ctx->gi_contrib.ptr;
    /* two sets: per-set counters and the shared queue block; otherwise all
     * in the first set's block */
    const bool two = n_sets == 2;
    uint32_t* cnt = two ? (uint32_t*)ctx->gi_counts.ptr + GI_CNT_QUEUE : sets[0].counters;
    hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(64), 0, ctx->stream, sets[0].counters, (uint64_t)GI_N_COUNTERS, 0u);
    if (two) {
        hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(64), 0, ctx->stream, sets[1].counters, (uint64_t)GI_N_COUNTERS, 0u);
        hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(64), 0, ctx->stream, cnt, (uint64_t)GI_N_COUNTERS, 0u);
    }
    for (uint32_t k = 0; k < n_sets; k++) {
        const gi_trace_set& t = sets[k];
        if (t.max_count == 0) continue;
        const gi_surf_rec* surf = t.surf ? t.surf : (const gi_surf_rec*)ctx->gi_surf.ptr;
        {
            tgx_launch_scope s(ctx, "gi_gen_k");
            hipLaunchKernelGGL(gi_gen_k, dim3(blocks_for((uint64_t)t.max_count * t.prm.spp)), dim3(GI_BLOCK), 0,
                               ctx->stream, surf, t.list,
                               t.count_dev, t.prm, t.frame_salt, rays, contrib, cnt, id_base[k], t.salt_add, t.max_count);
        }
        hipLaunchKernelGGL(gi_gen_tail_k, dim3(1), dim3(1), 0, ctx->stream, t.count_dev, t.max_count, t.prm.spp, cnt,
                           two ? t.counters : nullptr);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_gen_k");
    /* shade, trace and lit use what the lists share: sun, radiance, range */
    const tgx_gi_params& prm = sets[0].prm;
    if (ctx->tuning.trace_sort & 1u) {
        if ((rc = tgx_buffer_reserve(ctx, &ctx->gi_rays2, sizeof(svo_ray) * n_samples, "gi sorted rays"))) return rc;
        if ((rc = sort_rays(ctx, rays, cnt + 1, (uint32_t)n_samples, (svo_ray*)ctx->gi_rays2.ptr))) return rc;
        rays = (svo_ray*)ctx->gi_rays2.ptr;
    }
    if (!(ctx->tuning.trace_sort & 2u)) {
        /* bounce and shadow rays in one launch */
        trace_fuse F = {};
        F.contrib = contrib;
        F.sun0 = prm.sun_dir[0]; F.sun1 = prm.sun_dir[1]; F.sun2 = prm.sun_dir[2];
        /* the step lengths a shadow ray's set-up would divide out, in the
         * same IEEE single precision */
        for (int a = 0; a < 3; a++) {
            const float d = prm.sun_dir[a];
            const float td = d == 0.0f ? SVO_F32_MAX : 1.0f / (d > 0.0f ? d : -d);
            (a == 0 ? F.sun_td0 : a == 1 ? F.sun_td1 : F.sun_td2) = td;
        }
        F.sun_radiance = prm.sun_radiance;
        if (ctx->tuning.trace_regroup) {
            if ((rc = launch_trace_rg(ctx, rays, cnt + 1, (uint32_t)n_samples, prm.max_distance, cnt + 4, F,
                                      ctx->tuning.trace_regroup)))
                return rc;
        } else if ((rc = launch_trace(ctx, rays, cnt + 1, (uint32_t)n_samples, prm.max_distance, res, nullptr,
                                      &F))) {
            return rc;
        }
    } else {
    if ((rc = launch_trace(ctx, rays, cnt + 1, (uint32_t)n_samples, prm.max_distance, res, two ? nullptr : cnt + 2)))
        return rc;
    {
        tgx_launch_scope s(ctx, "gi_shade_k");
        hipLaunchKernelGGL(gi_shade_k, dim3((uint32_t)((n_samples + GI_BLOCK * SHADE_ITEMS - 1) / (GI_BLOCK * SHADE_ITEMS))),
                           dim3(GI_BLOCK), 0, ctx->stream, ctx->svo_n_inner, leaf_mats_of(ctx),
                           (const tgx_lut_entry*)ctx->luts.ptr, prm, rays, cnt + 1, res, shadow, weight, cnt,
                           two ? id_base[1] : 0xFFFFFFFFu, two ? sets[0].counters : nullptr,
                           two ? sets[1].counters : nullptr);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_shade_k");
    if (ctx->tuning.trace_sort & 2u) {
        /* bounce rays are consumed; their buffer takes the sorted shadow rays */
        svo_ray* dst = (svo_ray*)ctx->gi_rays.ptr;
        if ((rc = sort_rays(ctx, shadow, cnt + 3, (uint32_t)n_samples, dst))) return rc;
        shadow = dst;
    }
    if ((rc = launch_trace(ctx, shadow, cnt + 3, (uint32_t)n_samples, prm.max_distance, sres, cnt + 7))) return rc;
    {
        tgx_launch_scope s(ctx, "gi_lit_k");
        hipLaunchKernelGGL(gi_lit_k, dim3(blocks_for(n_samples)), dim3(256), 0, ctx->stream, shadow, cnt + 3, sres,
                           weight, contrib);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_lit_k");
    }
    /* the fused trace leaves its counts to the resolve pass */
    for (uint32_t k = 0; k < n_sets; k++) {
        const gi_trace_set& t = sets[k];
        if (t.max_count == 0) continue;
        const gi_surf_rec* surf = t.surf ? t.surf : (const gi_surf_rec*)ctx->gi_surf.ptr;
        tgx_launch_scope s(ctx, "gi_resolve_k");
        const uint32_t rb = (uint32_t)(((uint64_t)t.max_count + RESOLVE_BLOCK - 1) / RESOLVE_BLOCK);
        hipLaunchKernelGGL(gi_resolve_k, dim3(rb < RESOLVE_MAX_BLOCKS ? rb : RESOLVE_MAX_BLOCKS), dim3(RESOLVE_BLOCK), 0,
                           ctx->stream, surf, t.list, t.count_dev, t.max_count, contrib + 3ull * id_base[k], t.prm.spp, ctx->svo_n_inner,
                           leaf_mats_of(ctx), (const tgx_lut_entry*)ctx->luts.ptr, t.e_entry, t.e_var, t.radiance_dev,
                           t.irradiance_dev, fused ? (two ? t.counters : cnt) : nullptr);
    }
    TGX_LAUNCH_CHECK(ctx, "gi_resolve_k");
    return TGX_OK;
}

int tgx_gi_trace_list(tgx_context* ctx, const tgx_gi_params* prm, const uint32_t* list, const uint32_t* count_dev,
                      uint32_t max_count, uint32_t* cnt, float* e_entry, float* e_var, float* radiance_dev,
                      float* irradiance_dev, const uint32_t* frame_salt)
{
    const gi_trace_set t = {*prm, list, count_dev, max_count, cnt, e_entry, e_var, radiance_dev, irradiance_dev,
                            frame_salt, nullptr, 0u};
    return tgx_gi_trace_sets(ctx, &t, 1);
}

extern "C" {

int tgx_svo_trace(tgx_context* ctx, const float* origins, const float* directions, uint32_t n, float max_distance,
                  tgx_svo_hit* out_host)
{
    if (!ctx || (n && (!origins || !directions || !out_host))) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (int rc0 = tgx_svo_check_current(ctx)) return rc0;
    if (n == 0) return TGX_OK;
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_rays, sizeof(svo_ray) * (size_t)n, "svo rays")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_hits, (sizeof(svo_res) + sizeof(tgx_svo_hit)) * (size_t)n, "svo hits")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_counts, 4ull * GI_CNT_TOTAL, "gi counters")))
        return rc;
    std::vector<svo_ray> h(n);
    for (uint32_t i = 0; i < n; i++)
        h[i] = {origins[3 * i], origins[3 * i + 1], origins[3 * i + 2], directions[3 * i], directions[3 * i + 1],
                directions[3 * i + 2], i, 0u};
    svo_ray* d_rays = (svo_ray*)ctx->svo_rays.ptr;
    svo_res* d_res = (svo_res*)ctx->svo_hits.ptr;
    tgx_svo_hit* d_hits = (tgx_svo_hit*)(d_res + n);
    uint32_t* cnt = (uint32_t*)ctx->gi_counts.ptr + GI_CNT_SVO;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(d_rays, h.data(), sizeof(svo_ray) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(64), 0, ctx->stream, cnt, (uint64_t)GI_N_COUNTERS, 0u);
    hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(64), 0, ctx->stream, cnt + 6, 1ull, n);
    if ((rc = launch_trace(ctx, d_rays, cnt + 6, n, max_distance, d_res, nullptr))) return rc;
    hipLaunchKernelGGL(svo_hits_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, (const svo_gnode*)ctx->svo_nodes.ptr,
                       d_res, n, d_hits);
    TGX_LAUNCH_CHECK(ctx, "svo_hits_k");
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(out_host, d_hits, sizeof(tgx_svo_hit) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_svo_visibility_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, uint64_t* vis_dev)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!cam || !vis_dev) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "null argument");
    if (int rc0 = tgx_svo_check_current(ctx)) return rc0;
    if (w == 0 || h == 0 || w > 65535 || h > 65535 || (uint64_t)w * h >= (1ull << 31))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "invalid resolution %ux%u", w, h);
    if (!(cam->far_plane > 0.0f)) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "far plane must be > 0");
    const uint32_t n = w * h;
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_rays, sizeof(svo_ray) * (size_t)n, "svo rays")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_hits, sizeof(svo_res) * (size_t)n, "svo hits")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_counts, 4ull * GI_CNT_TOTAL, "gi counters")))
        return rc;
    gi_cam gc;
    for (int a = 0; a < 3; a++) {
        gc.pos[a] = cam->camera[a];
        gc.bl[a] = cam->ray_bl[a]; gc.br[a] = cam->ray_br[a];
        gc.tr[a] = cam->ray_tr[a]; gc.tl[a] = cam->ray_tl[a];
    }
    svo_ray* d_rays = (svo_ray*)ctx->svo_rays.ptr;
    svo_res* d_res = (svo_res*)ctx->svo_hits.ptr;
    uint32_t* cnt = (uint32_t*)ctx->gi_counts.ptr + GI_CNT_SVO;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(64), 0, ctx->stream, cnt, (uint64_t)GI_N_COUNTERS, 0u);
    hipLaunchKernelGGL(gi_fill_k, dim3(1), dim3(64), 0, ctx->stream, cnt + 6, 1ull, n);
    {
        tgx_launch_scope s(ctx, "svo_primary_rays_k");
        hipLaunchKernelGGL(svo_primary_rays_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, gc, w, h, d_rays);
    }
    TGX_LAUNCH_CHECK(ctx, "svo_primary_rays_k");
    if ((rc = launch_trace(ctx, d_rays, cnt + 6, n, SVO_F32_MAX, d_res, nullptr))) return rc;
    {
        tgx_launch_scope s(ctx, "svo_vis_pack_k");
        hipLaunchKernelGGL(svo_vis_pack_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, d_res, n, cam->far_plane,
                           vis_dev);
    }
    TGX_LAUNCH_CHECK(ctx, "svo_vis_pack_k");
    return TGX_OK;
}

int tgx_gi_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis_dev,
                  const tgx_gi_params* prm, float* radiance_dev, float* irradiance_dev)
{
    if (!ctx || !cam || !vis_dev || !prm || !radiance_dev) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    int rc;
    if ((rc = tgx_gi_validate(ctx, w, h, prm))) return rc;
    if ((rc = tgx_gi_surfaces(ctx, cam, w, h, vis_dev, radiance_dev, irradiance_dev))) return rc;
    uint32_t* cnt = (uint32_t*)ctx->gi_counts.ptr;
    ctx->gi_spp = prm->spp;
    return tgx_gi_trace_list(ctx, prm, (const uint32_t*)ctx->gi_list.ptr, cnt + GI_CNT_VISIBLE, w * h,
                             cnt + GI_CNT_PLAIN, nullptr, nullptr, radiance_dev, irradiance_dev, nullptr);
}

int tgx_gi_last_stats(tgx_context* ctx, tgx_gi_stats* out)
{
    if (!ctx || !out) return TGX_ERR_INVALID_ARG;
    if (!ctx->gi_counts.ptr) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no GI call yet");
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t c[GI_CNT_TOTAL];
    TGX_HIP_CHECK(ctx, hipMemcpy(c, ctx->gi_counts.ptr, sizeof(c), hipMemcpyDeviceToHost));
    out->visible_pixels = c[GI_CNT_VISIBLE];
    out->bounce_rays = c[GI_CNT_PLAIN + 1];
    out->bounce_hits = c[GI_CNT_PLAIN + 2];
    out->shadow_rays = c[GI_CNT_PLAIN + 3];
    return TGX_OK;
}

int tgx_frame_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const tgx_gi_params* params,
                     uint64_t* vis_dev, float* radiance_dev, float* irradiance_dev)
{
    int rc = tgx_visibility_device(ctx, cam, 1, w, h, vis_dev);
    if (rc) return rc;
    return tgx_gi_device(ctx, cam, w, h, vis_dev, params, radiance_dev, irradiance_dev);
}

} /* extern "C" */
