/*
 * tgx_svo.hip — sparse voxel octree build on gfx950.
 *
 * Build (replaces tg_svo_create, tg_sparse_voxel_octree.c:466-542, a
 * single-threaded depth-first recursion on the host):
 *   1. svo_frames_k      per cluster pointer: emptiness flag and the three
 *                        cluster transforms of the reference (world->cluster
 *                        :23-60, cluster->world :62-100 and its inverse).
 *   2. per level         (node, cluster pointer) pairs; svo_children_k runs
 *                        tg_intersect_aabb_obb_ignore_contact for the 8
 *                        children of each pair, a scan sizes the next level
 *                        and svo_emit_k writes it. Pairs stay ordered by the
 *                        parent pair, so every node's clusters stay in
 *                        ascending pointer order (the reference's list order)
 *                        without sorting between levels.
 *   3. per level         stable radix sort of the pairs by Morton code; each
 *                        distinct code is a node. Nodes are numbered by
 *                        (level, Morton code), children of a node are then
 *                        contiguous in child-index order; svo_nodes_k links
 *                        each node to its parent (binary search one level up).
 *   4. svo_leaf_k        one workgroup per 32^3 leaf: for each listed
 *                        cluster in order, the trimmed block box is point
 *                        sampled in the reference's exact fp32 order
 *                        (:242-315); solid bits collect in LDS, the first
 *                        cluster to set a voxel writes its material.
 * Traversal and GI live in tgx_gi.hip.
 */
#include "tgx_svo_dev.h"
#include "tgx_m4.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

/* Per-cluster-pointer transforms (rows 0..2 of each m4). */
struct svo_frame {
    float    ws2ms[12];
    float    cs2ws[12];
    float    ws2cs[12];
    uint32_t cidx;
    uint32_t lut;         /* (lut_idx & 255) << 8 */
    uint32_t pad[2];
};
static_assert(sizeof(svo_frame) == 160, "svo_frame layout");

/* rows 0..2 of m times (x, y, z, w), tg_math.c:2428-2438 order */
__device__ static inline f3 mulv(const float* m, float x, float y, float z, float w)
{
    f3 r;
    r.x = x * m[0] + y * m[1] + z * m[2] + w * m[3];
    r.y = x * m[4] + y * m[5] + z * m[6] + w * m[7];
    r.z = x * m[8] + y * m[9] + z * m[10] + w * m[11];
    return r;
}

__device__ static inline uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ static inline uint32_t compact3(uint32_t v)
{
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0x030000FFu;
    v = (v | (v >> 16)) & 0x000003FFu;
    return v;
}
__device__ static inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

/* ------------------------------------------------------------------------ */
/* tg_intersect_aabb_obb_ignore_contact against the cluster box [0,8]^3     */
/* (tg_physics.c:226-392), corners c[8] in cluster space                    */
/* ------------------------------------------------------------------------ */

__device__ static inline f3 f3_sub(f3 a, f3 b) { f3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
__device__ static inline float f3_dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ static inline float f3_mag(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ static inline f3 f3_normalized(f3 a)
{
    const float m = f3_mag(a);
    f3 r = {a.x / m, a.y / m, a.z / m};
    return r;
}

__device__ static bool obb_overlaps_cluster(const f3* c)
{
    bool sep;
    sep = true; for (int i = 0; i < 8; i++) sep = sep && c[i].x <= 0.0f; if (sep) return false;
    sep = true; for (int i = 0; i < 8; i++) sep = sep && c[i].x >= 8.0f; if (sep) return false;
    sep = true; for (int i = 0; i < 8; i++) sep = sep && c[i].y <= 0.0f; if (sep) return false;
    sep = true; for (int i = 0; i < 8; i++) sep = sep && c[i].y >= 8.0f; if (sep) return false;
    sep = true; for (int i = 0; i < 8; i++) sep = sep && c[i].z <= 0.0f; if (sep) return false;
    sep = true; for (int i = 0; i < 8; i++) sep = sep && c[i].z >= 8.0f; if (sep) return false;

    const f3 nxp = f3_normalized(f3_sub(c[1], c[0]));
    const f3 nyp = f3_normalized(f3_sub(c[2], c[0]));
    const f3 nzp = f3_normalized(f3_sub(c[4], c[0]));
    f3 nrm[6];
    nrm[0] = {-nxp.x, -nxp.y, -nxp.z}; nrm[1] = nxp;
    nrm[2] = {-nyp.x, -nyp.y, -nyp.z}; nrm[3] = nyp;
    nrm[4] = {-nzp.x, -nzp.y, -nzp.z}; nrm[5] = nzp;
    const float l0 = f3_mag(c[0]);
    const float l1 = f3_mag(c[7]);
    f3 n0 = {0.0f, 0.0f, 0.0f}, n1 = {0.0f, 0.0f, 0.0f};
    if (l0 != 0.0f) { n0.x = c[0].x / l0; n0.y = c[0].y / l0; n0.z = c[0].z / l0; }
    if (l1 != 0.0f) { n1.x = c[7].x / l1; n1.y = c[7].y / l1; n1.z = c[7].z / l1; }
    float dist[6];
    dist[0] = l0 * f3_dot(n0, nrm[0]);
    dist[1] = l1 * f3_dot(n1, nrm[1]);
    dist[2] = l0 * f3_dot(n0, nrm[2]);
    dist[3] = l1 * f3_dot(n1, nrm[3]);
    dist[4] = l0 * f3_dot(n0, nrm[4]);
    dist[5] = l1 * f3_dot(n1, nrm[5]);
    for (int p = 0; p < 6; p++) {
        sep = true;
        for (int i = 0; i < 8 && sep; i++) {
            const f3 pt = {(i & 1) ? 8.0f : 0.0f, (i & 2) ? 8.0f : 0.0f, (i & 4) ? 8.0f : 0.0f};
            sep = (f3_dot(nrm[p], pt) - dist[p]) >= 0.0f;
        }
        if (sep) return false;
    }
    return true;
}

/* ------------------------------------------------------------------------ */
/* Build kernels                                                            */
/* ------------------------------------------------------------------------ */

__global__ void svo_fill_k(uint32_t* p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void svo_frames_k(const tgx_object* __restrict__ objects,
                                                    const uint32_t* __restrict__ cptrs,
                                                    const uint32_t* __restrict__ c2o,
                                                    const uint32_t* __restrict__ bits, uint32_t n_cptr,
                                                    svo_frame* __restrict__ frames, uint32_t* __restrict__ nonempty)
{
    const uint32_t cp = blockIdx.x * blockDim.x + threadIdx.x;
    if (cp >= n_cptr) return;
    const uint32_t cidx = cptrs[cp];
    uint32_t any = 0;
    for (int i = 0; i < 16; i++) any |= bits[(size_t)cidx * 16u + i];
    nonempty[cp] = any ? 1u : 0u;
    const tgx_object ob = objects[c2o[cidx]];
    const uint32_t rel = cp - ob.first_cluster_pointer;
    const uint32_t nx = ob.n_cluster_pointers_per_dim[0], ny = ob.n_cluster_pointers_per_dim[1];
    const uint32_t rx = rel % nx, ry = (rel / nx) % ny, rz = rel / (nx * ny);
    const tgx_m4 w2c = tgx_world_to_cluster(ob.rotation, ob.translation, ob.n_cluster_pointers_per_dim, rx, ry, rz);
    const tgx_m4 c2w = tgx_cluster_to_world(ob.rotation, ob.translation, ob.n_cluster_pointers_per_dim, rx, ry, rz);
    const tgx_m4 w2c2 = tgx_m4_inverse(c2w);
    svo_frame f;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) {
            f.ws2ms[4 * r + c] = TGX_M(w2c, r, c);
            f.cs2ws[4 * r + c] = TGX_M(c2w, r, c);
            f.ws2cs[4 * r + c] = TGX_M(w2c2, r, c);
        }
    f.cidx = cidx;
    f.lut = (ob.lut_idx & 0xFFu) << 8;
    f.pad[0] = f.pad[1] = 0u;
    frames[cp] = f;
}

/* Level 0: the root's pairs are the non-empty cluster pointers in order. */
__global__ void svo_compact_k(const uint32_t* __restrict__ nonempty, const uint32_t* __restrict__ pos, uint32_t n,
                              uint32_t* __restrict__ out_coord, uint32_t* __restrict__ out_cptr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !nonempty[i]) return;
    out_coord[pos[i]] = 0u;
    out_cptr[pos[i]] = i;
}

/* Packed node coordinates: x | y << 10 | z << 20. */
__device__ static inline void unpack(uint32_t p, uint32_t& x, uint32_t& y, uint32_t& z)
{
    x = p & 0x3FFu; y = (p >> 10) & 0x3FFu; z = p >> 20;
}

__global__ __launch_bounds__(256) void svo_children_k(const uint32_t* __restrict__ coord, const uint32_t* __restrict__ cptr,
                                                      uint32_t n, const svo_frame* __restrict__ frames, float mnx,
                                                      float mny, float mnz, float side,
                                                      uint32_t* __restrict__ masks, uint32_t* __restrict__ counts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x, y, z;
    unpack(coord[i], x, y, z);
    /* Parent box in world space: exact integers (tg_sparse_voxel_octree.c:344,359-365). */
    const float pminx = mnx + (float)x * side, pminy = mny + (float)y * side, pminz = mnz + (float)z * side;
    const float pmaxx = pminx + side, pmaxy = pminy + side, pmaxz = pminz + side;
    const float cex = (pmaxx - pminx) * 0.5f, cey = (pmaxy - pminy) * 0.5f, cez = (pmaxz - pminz) * 0.5f;
    const float* m = frames[cptr[i]].ws2ms;
    uint32_t mask = 0;
    for (uint32_t c = 0; c < 8; c++) {
        const float cminx = pminx + (float)(c % 2) * cex;
        const float cminy = pminy + (float)((c / 2) % 2) * cey;
        const float cminz = pminz + (float)((c / 4) % 2) * cez;
        const float cmaxx = cminx + cex, cmaxy = cminy + cey, cmaxz = cminz + cez;
        f3 k[8];
        for (int j = 0; j < 8; j++)
            k[j] = mulv(m, (j & 1) ? cmaxx : cminx, (j & 2) ? cmaxy : cminy, (j & 4) ? cmaxz : cminz, 1.0f);
        if (obb_overlaps_cluster(k)) mask |= 1u << c;
    }
    masks[i] = mask;
    counts[i] = (uint32_t)__popc(mask);
}

__global__ void svo_emit_k(const uint32_t* __restrict__ coord, const uint32_t* __restrict__ cptr, uint32_t n,
                           const uint32_t* __restrict__ masks, const uint32_t* __restrict__ off,
                           uint32_t* __restrict__ out_coord, uint32_t* __restrict__ out_cptr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x, y, z;
    unpack(coord[i], x, y, z);
    const uint32_t mask = masks[i];
    uint32_t o = off[i];
    for (uint32_t c = 0; c < 8; c++) {
        if (!(mask & (1u << c))) continue;
        const uint32_t cx = 2u * x + (c & 1u), cy = 2u * y + ((c >> 1) & 1u), cz = 2u * z + ((c >> 2) & 1u);
        out_coord[o] = cx | (cy << 10) | (cz << 20);
        out_cptr[o] = cptr[i];
        o++;
    }
}

__global__ void svo_morton_k(const uint32_t* __restrict__ coord, uint32_t n, uint32_t* __restrict__ key)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x, y, z;
    unpack(coord[i], x, y, z);
    key[i] = morton3(x, y, z);
}

__global__ void svo_unique_flags_k(const uint32_t* __restrict__ key, uint32_t n, uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

/* One thread per pair starting a node: creates the node (canonical index
 * base + rank) and links it to its parent one level up. */
__global__ void svo_nodes_k(const uint32_t* __restrict__ key, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ rank, uint32_t n, uint32_t level, uint32_t base,
                            uint32_t is_leaf_level, const uint32_t* __restrict__ parent_keys, uint32_t n_parent,
                            uint32_t parent_base, svo_gnode* __restrict__ nodes, uint32_t* __restrict__ first_pair)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t r = rank[i];
    const uint32_t idx = base + r;
    const uint32_t k = key[i];
    first_pair[r] = i;
    svo_gnode nd;
    nd.first_child = SVO_NONE;
    nd.leaf_idx = is_leaf_level ? r : SVO_NONE;
    nd.masks = level << 16;
    nd.cx = compact3(k);
    nd.cy = compact3(k >> 1);
    nd.cz = compact3(k >> 2);
    nd.leaf_n = 0;
    /* parent: Morton code >> 3 among the sorted keys of the level above */
    const uint32_t pk = k >> 3;
    uint32_t lo = 0, hi = n_parent;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (parent_keys[mid] < pk) lo = mid + 1; else hi = mid;
    }
    const uint32_t parent = parent_base + lo;
    nd.parent = parent;
    nodes[idx] = nd;
    const uint32_t bit = 1u << (k & 7u);
    atomicOr(&nodes[parent].masks, is_leaf_level ? (bit | (bit << 8)) : bit);
    atomicMin(&nodes[parent].first_child, idx);
}

/* Keys of the distinct nodes of a level, in order. */
__global__ void svo_node_keys_k(const uint32_t* __restrict__ key, const uint32_t* __restrict__ flag,
                                const uint32_t* __restrict__ rank, uint32_t n, uint32_t* __restrict__ node_keys)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) node_keys[rank[i]] = key[i];
}

/* Trilinear interpolation of the block corners (tg_sparse_voxel_octree.c:242-291). */
__device__ static inline void sample_center(const f3* bc, uint32_t x, uint32_t y, uint32_t z, float& cx, float& cy,
                                            float& cz)
{
    const float tz = ((float)z + 0.5f) / 32.0f;
    const float omtz = 1.0f - tz;
    const float pz0x = omtz * bc[0].x + tz * bc[4].x, pz0y = omtz * bc[0].y + tz * bc[4].y, pz0z = omtz * bc[0].z + tz * bc[4].z;
    const float pz1x = omtz * bc[1].x + tz * bc[5].x, pz1y = omtz * bc[1].y + tz * bc[5].y, pz1z = omtz * bc[1].z + tz * bc[5].z;
    const float pz2x = omtz * bc[2].x + tz * bc[6].x, pz2y = omtz * bc[2].y + tz * bc[6].y, pz2z = omtz * bc[2].z + tz * bc[6].z;
    const float pz3x = omtz * bc[3].x + tz * bc[7].x, pz3y = omtz * bc[3].y + tz * bc[7].y, pz3z = omtz * bc[3].z + tz * bc[7].z;
    const float ty = ((float)y + 0.5f) / 32.0f;
    const float omty = 1.0f - ty;
    const float py0x = omty * pz0x + ty * pz2x, py0y = omty * pz0y + ty * pz2y, py0z = omty * pz0z + ty * pz2z;
    const float py1x = omty * pz1x + ty * pz3x, py1y = omty * pz1y + ty * pz3y, py1z = omty * pz1z + ty * pz3z;
    const float tx = ((float)x + 0.5f) / 32.0f;
    const float omtx = 1.0f - tx;
    cx = omtx * py0x + tx * py1x;
    cy = omtx * py0y + tx * py1y;
    cz = omtx * py0z + tx * py1z;
}

/* Leaf sampling, in two passes over the same leaves with the same
 * arithmetic. kMats false: the leaf's bits, its cluster list and its solid
 * voxel count. kMats true (after the counts are scanned into mat_off): the
 * materials of its solid voxels only, in bit order, at mat_off[leaf] in the
 * pool, and the per-word prefix counts that index them (wpre[w] = solid
 * voxels in the leaf's words before w). A voxel's material is that of the
 * first cluster in list order that samples it solid. */
template <bool kMats>
__global__ __launch_bounds__(256) void svo_leaf_k(const svo_gnode* __restrict__ nodes, uint32_t leaf_base,
                                                  const uint32_t* __restrict__ pair_cptr,
                                                  const uint32_t* __restrict__ leaf_first_pair, uint32_t n_pairs,
                                                  uint32_t n_leaves, const svo_frame* __restrict__ frames,
                                                  const uint32_t* __restrict__ bits,
                                                  const uint8_t* __restrict__ materials, float mnx, float mny,
                                                  float mnz, uint32_t quirk, uint32_t* __restrict__ leaf_bits,
                                                  uint32_t* __restrict__ leaf_solid, const uint32_t* __restrict__ mat_off,
                                                  uint16_t* __restrict__ wpre, uint16_t* __restrict__ pool,
                                                  uint32_t* __restrict__ leaf_n,
                                                  uint32_t* __restrict__ list_tmp, svo_gnode* __restrict__ nodes_rw,
                                                  const uint32_t* __restrict__ reuse, const uint32_t* __restrict__ old_bits,
                                                  const uint32_t* __restrict__ old_off, const uint16_t* __restrict__ old_wpre,
                                                  const uint16_t* __restrict__ old_pool, const uint32_t* __restrict__ old_n,
                                                  const uint32_t* __restrict__ old_list_off,
                                                  const uint32_t* __restrict__ old_list)
{
    __shared__ uint32_t lbits[TGX_SVO_BLOCK_WORDS];
    __shared__ uint32_t fbits[kMats ? TGX_SVO_BLOCK_WORDS : 1];
    __shared__ uint16_t lpre[kMats ? TGX_SVO_BLOCK_WORDS : 1];
    __shared__ uint32_t wsum[4];
    __shared__ int any_flag;
    __shared__ uint32_t n_contrib, last_cidx;
    const uint32_t leaf = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    if (reuse && reuse[leaf] != SVO_NONE) {
        /* an update: same candidate clusters in the same order, none of
         * them new, so the previous build's leaf is this leaf */
        const uint32_t j = reuse[leaf];
        if constexpr (!kMats) {
            const uint4* sb = (const uint4*)(old_bits + (size_t)j * TGX_SVO_BLOCK_WORDS);
            uint4* db = (uint4*)(leaf_bits + (size_t)leaf * TGX_SVO_BLOCK_WORDS);
            for (uint32_t w = tid; w < TGX_SVO_BLOCK_WORDS / 4u; w += blockDim.x) db[w] = sb[w];
            const uint32_t n = old_n[j], b = leaf_first_pair[leaf], o = old_list_off[j];
            for (uint32_t i = tid; i < n; i += blockDim.x) list_tmp[b + i] = old_list[o + i];
            if (tid == 0) {
                leaf_n[leaf] = n;
                nodes_rw[leaf_base + leaf].leaf_n = n;
                leaf_solid[leaf] = old_off[j + 1] - old_off[j];
            }
        } else {
            const uint2* sw = (const uint2*)(old_wpre + (size_t)j * TGX_SVO_BLOCK_WORDS);
            uint2* dw = (uint2*)(wpre + (size_t)leaf * TGX_SVO_BLOCK_WORDS);
            for (uint32_t w = tid; w < TGX_SVO_BLOCK_WORDS / 4u; w += blockDim.x) dw[w] = sw[w];
            const uint32_t o = old_off[j], n = old_off[j + 1] - o, d = mat_off[leaf];
            for (uint32_t i = tid; i < n; i += blockDim.x) pool[d + i] = old_pool[o + i];
        }
        return;
    }
    const svo_gnode nd = nodes[leaf_base + leaf];
    const float pminx = mnx + (float)nd.cx * 32.0f, pminy = mny + (float)nd.cy * 32.0f, pminz = mnz + (float)nd.cz * 32.0f;
    const float pmaxx = pminx + 32.0f, pmaxy = pminy + 32.0f, pmaxz = pminz + 32.0f;
    for (uint32_t w = tid; w < TGX_SVO_BLOCK_WORDS; w += blockDim.x) lbits[w] = 0u;
    if (tid == 0) { n_contrib = 0u; last_cidx = SVO_NONE; }
    uint16_t* mats = nullptr;
    if constexpr (kMats) {
        /* the leaf's final bits (first pass) and their per-word prefix */
        const uint32_t* fb = leaf_bits + (size_t)leaf * TGX_SVO_BLOCK_WORDS;
        uint32_t c[4], tot = 0;
        for (uint32_t i = 0; i < 4u; i++) {
            const uint32_t w = 4u * tid + i;
            fbits[w] = fb[w];
            c[i] = (uint32_t)__popc(fbits[w]);
            tot += c[i];
        }
        /* exclusive scan of the 256 threads' totals: wave scan, then waves */
        const uint32_t lane = tid & 63u, wv = tid >> 6;
        uint32_t incl = tot;
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint32_t u = __shfl_up(incl, d);
            if (lane >= d) incl += u;
        }
        if (lane == 63u) wsum[wv] = incl;
        __syncthreads();
        uint32_t run = incl - tot;
        for (uint32_t k = 0; k < wv; k++) run += wsum[k];
        for (uint32_t i = 0; i < 4u; i++) {
            lpre[4u * tid + i] = (uint16_t)run;
            wpre[(size_t)leaf * TGX_SVO_BLOCK_WORDS + 4u * tid + i] = (uint16_t)run;
            run += c[i];
        }
        mats = pool + mat_off[leaf];
    }
    const uint32_t b = leaf_first_pair[leaf];
    const uint32_t e = leaf + 1 < n_leaves ? leaf_first_pair[leaf + 1] : n_pairs;
    __syncthreads();
    for (uint32_t k = b; k < e; k++) {
        const svo_frame& f = frames[pair_cptr[k]];
        /* cluster corners in world space and their AABB (:134-197) */
        f3 cw[8];
        for (int i = 0; i < 8; i++)
            cw[i] = mulv(f.cs2ws, (i & 1) ? 8.0f : 0.0f, (i & 2) ? 8.0f : 0.0f, (i & 4) ? 8.0f : 0.0f, 1.0f);
        const int c3 = quirk ? 2 : 3;
        const float minx = fmin2(fmin2(fmin2(cw[0].x, cw[1].x), fmin2(cw[2].x, cw[c3].x)), fmin2(fmin2(cw[4].x, cw[6].x), fmin2(cw[5].x, cw[7].x)));
        const float miny = fmin2(fmin2(fmin2(cw[0].y, cw[1].y), fmin2(cw[2].y, cw[c3].y)), fmin2(fmin2(cw[4].y, cw[6].y), fmin2(cw[5].y, cw[7].y)));
        const float minz = fmin2(fmin2(fmin2(cw[0].z, cw[1].z), fmin2(cw[2].z, cw[c3].z)), fmin2(fmin2(cw[4].z, cw[6].z), fmin2(cw[5].z, cw[7].z)));
        const float maxx = fmax2(fmax2(fmax2(cw[0].x, cw[1].x), fmax2(cw[2].x, cw[c3].x)), fmax2(fmax2(cw[4].x, cw[6].x), fmax2(cw[5].x, cw[7].x)));
        const float maxy = fmax2(fmax2(fmax2(cw[0].y, cw[1].y), fmax2(cw[2].y, cw[c3].y)), fmax2(fmax2(cw[4].y, cw[6].y), fmax2(cw[5].y, cw[7].y)));
        const float maxz = fmax2(fmax2(fmax2(cw[0].z, cw[1].z), fmax2(cw[2].z, cw[c3].z)), fmax2(fmax2(cw[4].z, cw[6].z), fmax2(cw[5].z, cw[7].z)));
        /* block corners in cluster space (:199-211) */
        f3 bc[8];
        for (int i = 0; i < 8; i++)
            bc[i] = mulv(f.ws2cs, (i & 1) ? pmaxx : pminx, (i & 2) ? pmaxy : pminy, (i & 4) ? pmaxz : pminz, 1.0f);
        /* trimmed box (:218-233), clamped into [0, 32] */
        const float tlo[3] = {fmax2(pminx, floorf(minx)) - pminx, fmax2(pminy, floorf(miny)) - pminy, fmax2(pminz, floorf(minz)) - pminz};
        const float thi[3] = {fmin2(pmaxx, ceilf(maxx)) - pminx, fmin2(pmaxy, ceilf(maxy)) - pminy, fmin2(pmaxz, ceilf(maxz)) - pminz};
        uint32_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (uint32_t)fmin2(32.0f, fmax2(0.0f, tlo[a]));
            hi[a] = (uint32_t)fmin2(32.0f, fmax2(0.0f, thi[a]));
        }
        const uint32_t nx = hi[0] > lo[0] ? hi[0] - lo[0] : 0u;
        const uint32_t ny = hi[1] > lo[1] ? hi[1] - lo[1] : 0u;
        const uint32_t nz = hi[2] > lo[2] ? hi[2] - lo[2] : 0u;
        const uint32_t nv = nx * ny * nz;
        const uint64_t coff = (uint64_t)f.cidx * 512u;
        if (tid == 0) any_flag = 0;
        __syncthreads();
        bool any = false;
        for (uint32_t j = tid; j < nv; j += blockDim.x) {
            const uint32_t x = lo[0] + j % nx, y = lo[1] + (j / nx) % ny, z = lo[2] + j / (nx * ny);
            float cx, cy, cz;
            sample_center(bc, x, y, z, cx, cy, cz);
            if (cx < 0.0f || cx >= 8.0f || cy < 0.0f || cy >= 8.0f || cz < 0.0f || cz >= 8.0f) continue;
            const uint32_t v = 64u * (uint32_t)cz + 8u * (uint32_t)cy + (uint32_t)cx;
            const uint64_t cv = coff + v;
            if (!(bits[cv >> 5] & (1u << (uint32_t)(cv & 31u)))) continue;
            any = true;
            const uint32_t bv = 1024u * z + 32u * y + x;
            const uint32_t m = 1u << (bv & 31u);
            /* Within one cluster each block voxel is visited once; earlier
             * clusters' bits are complete (barrier), so the first cluster in
             * list order writes the material. */
            if (!(lbits[bv >> 5] & m)) {
                atomicOr(&lbits[bv >> 5], m);
                if constexpr (kMats)
                    mats[lpre[bv >> 5] + (uint32_t)__popc(fbits[bv >> 5] & (m - 1u))] =
                        (uint16_t)(f.lut | (materials ? (uint32_t)materials[cv] : 0u));
            }
        }
        if constexpr (!kMats) {
            if (any) any_flag = 1;
            __syncthreads();
            if (tid == 0 && any_flag && last_cidx != f.cidx) {
                list_tmp[b + n_contrib] = f.cidx;
                n_contrib++;
                last_cidx = f.cidx;
            }
        }
        __syncthreads();
    }
    if constexpr (!kMats) {
        uint32_t* out = leaf_bits + (size_t)leaf * TGX_SVO_BLOCK_WORDS;
        uint32_t cnt = 0;
        for (uint32_t w = tid; w < TGX_SVO_BLOCK_WORDS; w += blockDim.x) {
            out[w] = lbits[w];
            cnt += (uint32_t)__popc(lbits[w]);
        }
        for (uint32_t d = 32; d > 0; d >>= 1) cnt += __shfl_xor(cnt, d);
        if ((tid & 63u) == 0u) wsum[tid >> 6] = cnt;
        __syncthreads();
        if (tid == 0) {
            leaf_n[leaf] = n_contrib;
            nodes_rw[leaf_base + leaf].leaf_n = n_contrib;
            leaf_solid[leaf] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        }
    }
}

/* Update: for each new leaf, the previous build's leaf with the same key
 * whose candidate clusters (cluster indices in pointer order) are the same
 * and none of them placed since that build; SVO_NONE otherwise. */
__global__ __launch_bounds__(256) void svo_match_k(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ first,
                                                   uint32_t n_pairs, uint32_t n_leaves,
                                                   const uint32_t* __restrict__ pair_cptr,
                                                   const svo_frame* __restrict__ frames, const uint8_t* __restrict__ dirty,
                                                   const uint32_t* __restrict__ prev_keys, uint32_t n_prev,
                                                   const uint32_t* __restrict__ prev_off,
                                                   const uint32_t* __restrict__ prev_cand, uint32_t* __restrict__ reuse,
                                                   uint32_t* __restrict__ n_reused)
{
    /* one wave per leaf: the search is wave-uniform, the candidates are
     * compared 64 at a time */
    const uint32_t l = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u;
    const uint32_t lane = threadIdx.x & 63u;
    if (l >= n_leaves) return;
    const uint32_t key = keys[l];
    uint32_t lo = 0, hi = n_prev;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prev_keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    uint32_t r = SVO_NONE;
    if (lo < n_prev && prev_keys[lo] == key) {
        const uint32_t b = first[l], e = l + 1 < n_leaves ? first[l + 1] : n_pairs;
        const uint32_t pb = prev_off[lo], pe = prev_off[lo + 1];
        if (e - b == pe - pb) {
            uint32_t bad = 0u;
            for (uint32_t k = lane; k < e - b; k += 64u) {
                const uint32_t c = frames[pair_cptr[b + k]].cidx;
                bad |= (c != prev_cand[pb + k] || dirty[c] != 0u) ? 1u : 0u;
            }
            if (__ballot(bad != 0u) == 0ull) r = lo;
        }
    }
    if (lane == 0) {
        reuse[l] = r;
        if (r != SVO_NONE) atomicAdd(n_reused, 1u);
    }
}

/* candidate cluster pointers -> cluster indices (kept for the next update) */
__global__ void svo_cand_k(const uint32_t* __restrict__ pair_cptr, uint32_t n, const svo_frame* __restrict__ frames,
                           uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = frames[pair_cptr[i]].cidx;
}

__global__ void svo_list_compact_k(const uint32_t* __restrict__ leaf_first_pair, const uint32_t* __restrict__ leaf_n,
                                   const uint32_t* __restrict__ list_off, uint32_t n_leaves,
                                   const uint32_t* __restrict__ list_tmp, uint32_t* __restrict__ list)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_leaves) return;
    const uint32_t b = leaf_first_pair[l], o = list_off[l];
    for (uint32_t i = 0; i < leaf_n[l]; i++) list[o + i] = list_tmp[b + i];
}

/* Compact inner nodes for traversal (8 B): x = first child, y = parent
 * (24 bits, 0xFFFFFF for the root) | child-valid mask << 24. */
__global__ void svo_inner_k(const svo_gnode* __restrict__ nodes, uint32_t n_inner, uint2* __restrict__ inner)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inner) return;
    const svo_gnode n = nodes[i];
    inner[i] = make_uint2(n.first_child, (n.parent & 0xFFFFFFu) | ((n.masks & 0xFFu) << 24));
}

/* Per leaf, one bit per 8^3 brick (bx + 4 by + 16 bz) that holds a solid
 * voxel. One wave per leaf, one lane per brick. */
__global__ __launch_bounds__(64) void svo_bricks_k(const uint32_t* __restrict__ leaf_bits,
                                                   unsigned long long* __restrict__ bricks)
{
    const uint32_t t = threadIdx.x;
    const uint32_t bx = t & 3u, by = (t >> 2) & 3u, bz = t >> 4;
    const uint32_t* g = leaf_bits + (size_t)blockIdx.x * TGX_SVO_BLOCK_WORDS;
    uint32_t acc = 0;
    for (uint32_t z = 8u * bz; z < 8u * bz + 8u; z++)
        for (uint32_t y = 8u * by; y < 8u * by + 8u; y++) acc |= (g[32u * z + y] >> (8u * bx)) & 0xFFu;
    const unsigned long long m = __ballot(acc != 0u);
    if (t == 0) bricks[blockIdx.x] = m;
}

/* ------------------------------------------------------------------------ */
/* Host side                                                                */
/* ------------------------------------------------------------------------ */

void tgx_svo_release(tgx_context* ctx)
{
    tgx_buffer* bufs[] = {&ctx->svo_nodes, &ctx->svo_leaf_bits, &ctx->svo_leaf_mats, &ctx->svo_leaf_moff,
                          &ctx->svo_leaf_wpre, &ctx->svo_old_moff, &ctx->svo_old_wpre, &ctx->svo_leaf_n,
                          &ctx->svo_list_off, &ctx->svo_list, &ctx->svo_frames, &ctx->svo_rays,
                          &ctx->svo_hits, &ctx->prim_tmp, &ctx->gi_counts, &ctx->gi_surf,
                          &ctx->gi_rays, &ctx->gi_res, &ctx->gi_shadow, &ctx->gi_shadow_res,
                          &ctx->gi_weight, &ctx->gi_contrib, &ctx->gi_list, &ctx->gi_rays2, &ctx->gi_sort, &ctx->gi_bsurf, &ctx->gi_blist, &ctx->gi_bcount, &ctx->svo_inner,
                          &ctx->svo_bricks};
    for (tgx_buffer* b : bufs) tgx_buffer_release(b);
    tgx_buffer* prev[] = {&ctx->svo_prev_keys, &ctx->svo_prev_off, &ctx->svo_prev_cand, &ctx->svo_old_bits,
                          &ctx->svo_old_mats, &ctx->svo_old_n, &ctx->svo_old_list_off, &ctx->svo_old_list,
                          &ctx->svo_arena};
    for (tgx_buffer* b : prev) tgx_buffer_release(b);
    ctx->svo_built = false;
    ctx->svo_prev_valid = false;
}

namespace {
/* Temporary device arrays of one build, carved from the context's build
 * arena (kept between builds, so a rebuild of a similar scene allocates
 * nothing). What does not fit is allocated on its own and freed when the
 * build ends, and the arena grows to the whole need for the next build. */
struct dbuf {
    void* p = nullptr;
    template <typename T> T* as() { return (T*)p; }
};
int dalloc(tgx_context* ctx, dbuf& b, size_t bytes, const char* what)
{
    bytes = (bytes + 255u) & ~(size_t)255u;
    if (bytes == 0) bytes = 256;
    if (ctx->svo_arena.ptr && ctx->svo_arena_used + bytes <= ctx->svo_arena.bytes) {
        b.p = (char*)ctx->svo_arena.ptr + ctx->svo_arena_used;
        ctx->svo_arena_used += bytes;
        return TGX_OK;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return tgx_fail(ctx, TGX_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) for %s", bytes, what);
    }
    ctx->svo_arena_spill.push_back(b.p);
    ctx->svo_arena_spill_bytes += bytes;
    return TGX_OK;
}
/* Opens the arena for one build; on every exit frees the spills and
 * records the size the next build needs. */
struct arena_scope {
    tgx_context* ctx;
    explicit arena_scope(tgx_context* c) : ctx(c)
    {
        ctx->svo_arena_used = 0;
        ctx->svo_arena_spill_bytes = 0;
        if (ctx->svo_arena_want > ctx->svo_arena.bytes) {
            tgx_buffer_release(&ctx->svo_arena);
            if (tgx_buffer_reserve(ctx, &ctx->svo_arena, ctx->svo_arena_want, "svo build arena")) {
                tgx_buffer_release(&ctx->svo_arena);  /* the build spills instead */
                ctx->last_error.clear();
            }
        }
    }
    ~arena_scope()
    {
        if (!ctx->svo_arena_spill.empty()) {
            (void)hipStreamSynchronize(ctx->stream);
            for (void* p : ctx->svo_arena_spill) (void)hipFree(p);
            ctx->svo_arena_spill.clear();
            const size_t need = ctx->svo_arena_used + ctx->svo_arena_spill_bytes;
            ctx->svo_arena_want = need + need / 4u;
        }
    }
};
/* Pinned scratch of the SVO code: bytes [64, 128) of the context's pinned
 * page (bytes [0, 64) hold the visibility statistics readback). */
inline uint32_t* svo_pinned(tgx_context* ctx) { return (uint32_t*)((char*)ctx->host_pinned + 64); }
int read_u32(tgx_context* ctx, const uint32_t* dev, uint32_t* out)
{
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(svo_pinned(ctx), dev, 4, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *out = *svo_pinned(ctx);
    return TGX_OK;
}
/* total of an exclusive scan = last offset + last count */
int scan_total(tgx_context* ctx, const uint32_t* counts, const uint32_t* offs, uint32_t n, uint32_t* total)
{
    if (n == 0) { *total = 0; return TGX_OK; }
    uint32_t* h = svo_pinned(ctx);
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(h, offs + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(h + 1, counts + n - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *total = h[0] + h[1];
    return TGX_OK;
}
}  // namespace

extern "C" {

static int svo_build(tgx_context* ctx, const float mn[3], uint32_t depth, uint32_t flags, bool update);

int tgx_svo_build(tgx_context* ctx, const float mn[3], uint32_t depth, uint32_t flags)
{
    return svo_build(ctx, mn, depth, flags, false);
}

int tgx_svo_update(tgx_context* ctx, uint32_t* reused_leaves, uint32_t* voxelized_leaves)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!ctx->svo_built) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no SVO built: call tgx_svo_build first");
    const float mn[3] = {ctx->svo_info.min[0], ctx->svo_info.min[1], ctx->svo_info.min[2]};
    const int rc = svo_build(ctx, mn, ctx->svo_info.depth, ctx->svo_info.flags, ctx->dynamic && ctx->svo_prev_valid);
    if (!rc) {
        if (reused_leaves) *reused_leaves = ctx->svo_reused;
        if (voxelized_leaves) *voxelized_leaves = ctx->svo_info.n_leaves - ctx->svo_reused;
    }
    return rc;
}

static int svo_build(tgx_context* ctx, const float mn[3], uint32_t depth, uint32_t flags, bool update)
{
    if (ctx && ctx->n_graphs > 0)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "a frame graph uses the SVO: destroy it before rebuilding");
    if (!ctx || !mn) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "no scene uploaded");
    if (depth < TGX_SVO_MIN_DEPTH || depth > TGX_SVO_MAX_DEPTH)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "depth %u outside [%u, %u]", depth, TGX_SVO_MIN_DEPTH, TGX_SVO_MAX_DEPTH);
    for (int a = 0; a < 3; a++)
        if (!(mn[a] == floorf(mn[a])) || fabsf(mn[a]) + (float)(1u << depth) >= 16777216.0f)
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "svo min must be integral and |min| + 2^depth < 2^24");
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    ctx->svo_reused = 0;
    ctx->svo_built = false;
    /* The buffers stay allocated and are reused when large enough. An
     * update keeps the last build's leaves aside (double buffering). */
    if (update) {
        std::swap(ctx->svo_leaf_bits, ctx->svo_old_bits);
        std::swap(ctx->svo_leaf_mats, ctx->svo_old_mats);
        std::swap(ctx->svo_leaf_moff, ctx->svo_old_moff);
        std::swap(ctx->svo_leaf_wpre, ctx->svo_old_wpre);
        std::swap(ctx->svo_leaf_n, ctx->svo_old_n);
        std::swap(ctx->svo_list_off, ctx->svo_old_list_off);
        std::swap(ctx->svo_list, ctx->svo_old_list);
    } else {
        ctx->svo_prev_valid = false;
    }
    arena_scope arena(ctx);
    int rc;
    const uint32_t n_cptr = ctx->n_cluster_pointers;
    if (n_cptr == 0) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "the scene has no clusters");
    const uint32_t leaf_level = depth - 5u;

    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_frames, sizeof(svo_frame) * (size_t)n_cptr, "svo frames"))) return rc;
    svo_frame* frames = (svo_frame*)ctx->svo_frames.ptr;
    dbuf nonempty, pos;
    if ((rc = dalloc(ctx, nonempty, 4ull * n_cptr, "svo flags")) || (rc = dalloc(ctx, pos, 4ull * n_cptr, "svo positions"))) return rc;
    {
        tgx_launch_scope s(ctx, "svo_frames_k");
        hipLaunchKernelGGL(svo_frames_k, dim3(blocks_for(n_cptr)), dim3(256), 0, ctx->stream,
                           (const tgx_object*)ctx->objects.ptr, (const uint32_t*)ctx->cluster_pointers.ptr,
                           (const uint32_t*)ctx->cluster_idx_to_object.ptr, (const uint32_t*)ctx->bits.ptr, n_cptr,
                           frames, nonempty.as<uint32_t>());
    }
    TGX_LAUNCH_CHECK(ctx, "svo_frames_k");
    if ((rc = tgx_prim_exclusive_scan_u32(ctx, nonempty.as<uint32_t>(), pos.as<uint32_t>(), n_cptr))) return rc;
    uint32_t n_pairs = 0;
    if ((rc = scan_total(ctx, nonempty.as<uint32_t>(), pos.as<uint32_t>(), n_cptr, &n_pairs))) return rc;

    /* Per level: pairs (coords, cptr) in emission order, then sorted. */
    std::vector<dbuf> lvl_coord(leaf_level + 1), lvl_cptr(leaf_level + 1);
    std::vector<uint32_t> lvl_n(leaf_level + 1, 0);
    if ((rc = dalloc(ctx, lvl_coord[0], 4ull * n_pairs, "svo pairs")) || (rc = dalloc(ctx, lvl_cptr[0], 4ull * n_pairs, "svo pairs"))) return rc;
    lvl_n[0] = n_pairs;
    hipLaunchKernelGGL(svo_compact_k, dim3(blocks_for(n_cptr)), dim3(256), 0, ctx->stream, nonempty.as<uint32_t>(),
                       pos.as<uint32_t>(), n_cptr, lvl_coord[0].as<uint32_t>(), lvl_cptr[0].as<uint32_t>());
    TGX_LAUNCH_CHECK(ctx, "svo_compact_k");

    for (uint32_t l = 0; l < leaf_level; l++) {
        const uint32_t n = lvl_n[l];
        dbuf masks, counts, offs;
        if ((rc = dalloc(ctx, masks, 4ull * n, "svo masks")) || (rc = dalloc(ctx, counts, 4ull * n, "svo counts")) ||
            (rc = dalloc(ctx, offs, 4ull * n, "svo offsets")))
            return rc;
        const float side = (float)(1u << (depth - l));
        if (n) {
            tgx_launch_scope s(ctx, "svo_children_k");
            hipLaunchKernelGGL(svo_children_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, lvl_coord[l].as<uint32_t>(),
                               lvl_cptr[l].as<uint32_t>(), n, frames, mn[0], mn[1], mn[2], side, masks.as<uint32_t>(),
                               counts.as<uint32_t>());
        }
        TGX_LAUNCH_CHECK(ctx, "svo_children_k");
        if ((rc = tgx_prim_exclusive_scan_u32(ctx, counts.as<uint32_t>(), offs.as<uint32_t>(), n))) return rc;
        uint32_t next = 0;
        if ((rc = scan_total(ctx, counts.as<uint32_t>(), offs.as<uint32_t>(), n, &next))) return rc;
        if ((rc = dalloc(ctx, lvl_coord[l + 1], 4ull * next, "svo pairs")) || (rc = dalloc(ctx, lvl_cptr[l + 1], 4ull * next, "svo pairs"))) return rc;
        if (n)
            hipLaunchKernelGGL(svo_emit_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, lvl_coord[l].as<uint32_t>(),
                               lvl_cptr[l].as<uint32_t>(), n, masks.as<uint32_t>(), offs.as<uint32_t>(),
                               lvl_coord[l + 1].as<uint32_t>(), lvl_cptr[l + 1].as<uint32_t>());
        TGX_LAUNCH_CHECK(ctx, "svo_emit_k");
        lvl_n[l + 1] = next;
    }

    /* Sort each level by Morton code (stable: clusters stay in pointer
     * order within a node) and count the distinct nodes. */
    std::vector<dbuf> lvl_keys(leaf_level + 1), lvl_scptr(leaf_level + 1), lvl_flag(leaf_level + 1),
        lvl_rank(leaf_level + 1), lvl_node_keys(leaf_level + 1);
    std::vector<uint32_t> lvl_nodes(leaf_level + 1, 0), lvl_base(leaf_level + 1, 0);
    lvl_nodes[0] = 1;
    uint32_t total_nodes = 1;
    for (uint32_t l = 1; l <= leaf_level; l++) {
        const uint32_t n = lvl_n[l];
        dbuf raw;
        if ((rc = dalloc(ctx, raw, 4ull * n, "svo keys")) || (rc = dalloc(ctx, lvl_keys[l], 4ull * n, "svo keys")) ||
            (rc = dalloc(ctx, lvl_scptr[l], 4ull * n, "svo cptr")) || (rc = dalloc(ctx, lvl_flag[l], 4ull * n, "svo flags")) ||
            (rc = dalloc(ctx, lvl_rank[l], 4ull * n, "svo ranks")))
            return rc;
        if (n) hipLaunchKernelGGL(svo_morton_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, lvl_coord[l].as<uint32_t>(), n, raw.as<uint32_t>());
        TGX_LAUNCH_CHECK(ctx, "svo_morton_k");
        if ((rc = tgx_prim_sort_pairs_u32(ctx, raw.as<uint32_t>(), lvl_keys[l].as<uint32_t>(), lvl_cptr[l].as<uint32_t>(),
                                          lvl_scptr[l].as<uint32_t>(), n, 3u * l)))
            return rc;
        if (n) hipLaunchKernelGGL(svo_unique_flags_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, lvl_keys[l].as<uint32_t>(), n, lvl_flag[l].as<uint32_t>());
        TGX_LAUNCH_CHECK(ctx, "svo_unique_flags_k");
        if ((rc = tgx_prim_exclusive_scan_u32(ctx, lvl_flag[l].as<uint32_t>(), lvl_rank[l].as<uint32_t>(), n))) return rc;
        uint32_t cnt = 0;
        if ((rc = scan_total(ctx, lvl_flag[l].as<uint32_t>(), lvl_rank[l].as<uint32_t>(), n, &cnt))) return rc;
        lvl_nodes[l] = cnt;
        lvl_base[l] = total_nodes;
        total_nodes += cnt;
        if ((rc = dalloc(ctx, lvl_node_keys[l], 4ull * cnt, "svo node keys"))) return rc;
        if (n) hipLaunchKernelGGL(svo_node_keys_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, lvl_keys[l].as<uint32_t>(),
                                  lvl_flag[l].as<uint32_t>(), lvl_rank[l].as<uint32_t>(), n, lvl_node_keys[l].as<uint32_t>());
        TGX_LAUNCH_CHECK(ctx, "svo_node_keys_k");
    }
    const uint32_t n_leaves = lvl_nodes[leaf_level];

    if (total_nodes >= 0xFFFFFFu)
        return tgx_fail(ctx, TGX_ERR_CAPACITY, "the SVO has %u nodes; traversal holds node indices in 24 bits", total_nodes);
    /* Nodes: root, then each level linked to its parents. */
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_nodes, sizeof(svo_gnode) * (size_t)total_nodes, "svo nodes"))) return rc;
    svo_gnode* nodes = (svo_gnode*)ctx->svo_nodes.ptr;
    {
        const svo_gnode root = {SVO_NONE, SVO_NONE, SVO_NONE, 0u, 0u, 0u, 0u, 0u};
        TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        memcpy((char*)ctx->host_pinned + 256, &root, sizeof(root));
        TGX_HIP_CHECK(ctx, hipMemcpyAsync(nodes, (char*)ctx->host_pinned + 256, sizeof(root), hipMemcpyHostToDevice, ctx->stream));
    }
    dbuf root_key;
    if ((rc = dalloc(ctx, root_key, 4, "svo root key"))) return rc;
    hipLaunchKernelGGL(svo_fill_k, dim3(1), dim3(64), 0, ctx->stream, root_key.as<uint32_t>(), 1ull, 0u);
    std::vector<dbuf> lvl_first(leaf_level + 1);
    for (uint32_t l = 1; l <= leaf_level; l++) {
        const uint32_t n = lvl_n[l];
        if ((rc = dalloc(ctx, lvl_first[l], 4ull * (lvl_nodes[l] + 1), "svo first pair"))) return rc;
        const uint32_t* pkeys = l == 1 ? root_key.as<uint32_t>() : lvl_node_keys[l - 1].as<uint32_t>();
        if (n) {
            tgx_launch_scope s(ctx, "svo_nodes_k");
            hipLaunchKernelGGL(svo_nodes_k, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, lvl_keys[l].as<uint32_t>(),
                               lvl_flag[l].as<uint32_t>(), lvl_rank[l].as<uint32_t>(), n, l, lvl_base[l],
                               l == leaf_level ? 1u : 0u, pkeys, lvl_nodes[l - 1], lvl_base[l - 1], nodes,
                               lvl_first[l].as<uint32_t>());
        }
        TGX_LAUNCH_CHECK(ctx, "svo_nodes_k");
    }

    /* Leaves. */
    const size_t nl = n_leaves ? n_leaves : 1;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_bits, 4ull * TGX_SVO_BLOCK_WORDS * nl, "svo leaf bits")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_moff, 4ull * (nl + 1), "svo material offsets")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_wpre, 2ull * TGX_SVO_BLOCK_WORDS * nl, "svo material prefixes")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_n, 4ull * nl, "svo leaf n")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_list_off, 4ull * (nl + 1), "svo list offsets")))
        return rc;
    const uint32_t n_leaf_pairs = lvl_n[leaf_level];
    dbuf list_tmp, solid;
    if ((rc = dalloc(ctx, list_tmp, 4ull * (n_leaf_pairs ? n_leaf_pairs : 1), "svo list")) ||
        (rc = dalloc(ctx, solid, 4ull * nl, "svo solid counts")))
        return rc;
    dbuf reuse;
    if (update && n_leaves) {
        if ((rc = dalloc(ctx, reuse, 4ull * n_leaves + 4, "svo reuse"))) return rc;
        uint32_t* n_reused = reuse.as<uint32_t>() + n_leaves;
        hipLaunchKernelGGL(svo_fill_k, dim3(1), dim3(64), 0, ctx->stream, n_reused, 1ull, 0u);
        hipLaunchKernelGGL(svo_match_k, dim3((n_leaves + 3u) / 4u), dim3(256), 0, ctx->stream,
                           lvl_node_keys[leaf_level].as<uint32_t>(), lvl_first[leaf_level].as<uint32_t>(), n_leaf_pairs,
                           n_leaves, lvl_scptr[leaf_level].as<uint32_t>(), frames, (const uint8_t*)ctx->cl_dirty.ptr, (const uint32_t*)ctx->svo_prev_keys.ptr,
                           ctx->svo_prev_leaves, (const uint32_t*)ctx->svo_prev_off.ptr,
                           (const uint32_t*)ctx->svo_prev_cand.ptr, reuse.as<uint32_t>(), n_reused);
        TGX_LAUNCH_CHECK(ctx, "svo_match_k");
        if ((rc = read_u32(ctx, n_reused, &ctx->svo_reused))) return rc;
    }
    /* Leaf sampling in two passes: bits, lists and solid counts; then,
     * with the counts scanned into material offsets, the materials of the
     * solid voxels only. */
    const uint32_t quirk = (flags & TGX_SVO_FLAG_REFERENCE_CORNER_QUIRK) ? 1u : 0u;
    const uint32_t* rz = update ? reuse.as<uint32_t>() : nullptr;
    if (n_leaves) {
        tgx_launch_scope s(ctx, "svo_leaf_k");
        hipLaunchKernelGGL(svo_leaf_k<false>, dim3(n_leaves), dim3(256), 0, ctx->stream, nodes, lvl_base[leaf_level],
                           lvl_scptr[leaf_level].as<uint32_t>(), lvl_first[leaf_level].as<uint32_t>(), n_leaf_pairs,
                           n_leaves, frames, (const uint32_t*)ctx->bits.ptr, tgx_scene_materials(ctx),
                           mn[0], mn[1], mn[2], quirk, (uint32_t*)ctx->svo_leaf_bits.ptr, solid.as<uint32_t>(),
                           (const uint32_t*)nullptr, (uint16_t*)nullptr, (uint16_t*)nullptr,
                           (uint32_t*)ctx->svo_leaf_n.ptr, list_tmp.as<uint32_t>(), nodes, rz,
                           (const uint32_t*)ctx->svo_old_bits.ptr, (const uint32_t*)ctx->svo_old_moff.ptr,
                           (const uint16_t*)ctx->svo_old_wpre.ptr, (const uint16_t*)ctx->svo_old_mats.ptr,
                           (const uint32_t*)ctx->svo_old_n.ptr, (const uint32_t*)ctx->svo_old_list_off.ptr,
                           (const uint32_t*)ctx->svo_old_list.ptr);
    }
    TGX_LAUNCH_CHECK(ctx, "svo_leaf_k");
    if ((rc = tgx_prim_exclusive_scan_u32(ctx, solid.as<uint32_t>(), (uint32_t*)ctx->svo_leaf_moff.ptr, n_leaves))) return rc;
    uint32_t n_solid = 0;
    if ((rc = scan_total(ctx, solid.as<uint32_t>(), (const uint32_t*)ctx->svo_leaf_moff.ptr, n_leaves, &n_solid))) return rc;
    *(uint32_t*)((char*)ctx->host_pinned + 336) = n_solid;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync((uint32_t*)ctx->svo_leaf_moff.ptr + n_leaves, (char*)ctx->host_pinned + 336, 4,
                                      hipMemcpyHostToDevice, ctx->stream));
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_mats, 2ull * (n_solid ? n_solid : 1), "svo leaf materials"))) return rc;
    ctx->svo_n_solid = n_solid;
    if (n_leaves) {
        tgx_launch_scope s(ctx, "svo_leaf_mats_k");
        hipLaunchKernelGGL(svo_leaf_k<true>, dim3(n_leaves), dim3(256), 0, ctx->stream, nodes, lvl_base[leaf_level],
                           lvl_scptr[leaf_level].as<uint32_t>(), lvl_first[leaf_level].as<uint32_t>(), n_leaf_pairs,
                           n_leaves, frames, (const uint32_t*)ctx->bits.ptr, tgx_scene_materials(ctx),
                           mn[0], mn[1], mn[2], quirk, (uint32_t*)ctx->svo_leaf_bits.ptr, (uint32_t*)nullptr,
                           (const uint32_t*)ctx->svo_leaf_moff.ptr, (uint16_t*)ctx->svo_leaf_wpre.ptr,
                           (uint16_t*)ctx->svo_leaf_mats.ptr, (uint32_t*)nullptr, (uint32_t*)nullptr, nodes, rz,
                           (const uint32_t*)ctx->svo_old_bits.ptr, (const uint32_t*)ctx->svo_old_moff.ptr,
                           (const uint16_t*)ctx->svo_old_wpre.ptr, (const uint16_t*)ctx->svo_old_mats.ptr,
                           (const uint32_t*)ctx->svo_old_n.ptr, (const uint32_t*)ctx->svo_old_list_off.ptr,
                           (const uint32_t*)ctx->svo_old_list.ptr);
    }
    TGX_LAUNCH_CHECK(ctx, "svo_leaf_k");
    if ((rc = tgx_prim_exclusive_scan_u32(ctx, (const uint32_t*)ctx->svo_leaf_n.ptr, (uint32_t*)ctx->svo_list_off.ptr, n_leaves))) return rc;
    uint32_t n_list = 0;
    if ((rc = scan_total(ctx, (const uint32_t*)ctx->svo_leaf_n.ptr, (const uint32_t*)ctx->svo_list_off.ptr, n_leaves, &n_list))) return rc;
    *(uint32_t*)((char*)ctx->host_pinned + 320) = n_list;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync((uint32_t*)ctx->svo_list_off.ptr + n_leaves, (char*)ctx->host_pinned + 320, 4,
                                      hipMemcpyHostToDevice, ctx->stream));
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_list, 4ull * (n_list ? n_list : 1), "svo list"))) return rc;
    if (n_leaves)
        hipLaunchKernelGGL(svo_list_compact_k, dim3(blocks_for(n_leaves)), dim3(256), 0, ctx->stream,
                           lvl_first[leaf_level].as<uint32_t>(), (const uint32_t*)ctx->svo_leaf_n.ptr,
                           (const uint32_t*)ctx->svo_list_off.ptr, n_leaves, list_tmp.as<uint32_t>(),
                           (uint32_t*)ctx->svo_list.ptr);
    TGX_LAUNCH_CHECK(ctx, "svo_list_compact_k");

    /* Traversal layout: compact inner nodes and leaf brick masks. */
    const uint32_t n_inner = lvl_base[leaf_level];
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_inner, 8ull * n_inner, "svo inner nodes")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_bricks, 8ull * nl, "svo leaf bricks")))
        return rc;
    hipLaunchKernelGGL(svo_inner_k, dim3(blocks_for(n_inner)), dim3(256), 0, ctx->stream, (const svo_gnode*)nodes,
                       n_inner, (uint2*)ctx->svo_inner.ptr);
    TGX_LAUNCH_CHECK(ctx, "svo_inner_k");
    if (n_leaves)
        hipLaunchKernelGGL(svo_bricks_k, dim3(n_leaves), dim3(64), 0, ctx->stream, (const uint32_t*)ctx->svo_leaf_bits.ptr,
                           (unsigned long long*)ctx->svo_bricks.ptr);
    TGX_LAUNCH_CHECK(ctx, "svo_bricks_k");
    if (ctx->dynamic) {
        /* what the next tgx_svo_update compares against */
        if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_prev_keys, 4ull * nl, "svo previous keys")) ||
            (rc = tgx_buffer_reserve(ctx, &ctx->svo_prev_off, 4ull * (nl + 1), "svo previous offsets")) ||
            (rc = tgx_buffer_reserve(ctx, &ctx->svo_prev_cand, 4ull * (n_leaf_pairs ? n_leaf_pairs : 1), "svo previous candidates")))
            return rc;
        if (n_leaves) {
            TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->svo_prev_keys.ptr, lvl_node_keys[leaf_level].p, 4ull * n_leaves,
                                              hipMemcpyDeviceToDevice, ctx->stream));
            TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->svo_prev_off.ptr, lvl_first[leaf_level].p, 4ull * n_leaves,
                                              hipMemcpyDeviceToDevice, ctx->stream));
        }
        hipLaunchKernelGGL(svo_fill_k, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)ctx->svo_prev_off.ptr + n_leaves,
                           1ull, n_leaf_pairs);
        if (n_leaf_pairs)
            hipLaunchKernelGGL(svo_cand_k, dim3(blocks_for(n_leaf_pairs)), dim3(256), 0, ctx->stream,
                               lvl_scptr[leaf_level].as<uint32_t>(), n_leaf_pairs, frames,
                               (uint32_t*)ctx->svo_prev_cand.ptr);
        TGX_LAUNCH_CHECK(ctx, "svo_cand_k");
        const uint64_t dirty_words = (ctx->n_clusters + 3ull) / 4ull;
        hipLaunchKernelGGL(svo_fill_k, dim3((uint32_t)((dirty_words + 255) / 256)), dim3(256), 0, ctx->stream,
                           (uint32_t*)ctx->cl_dirty.ptr, dirty_words, 0u);
        TGX_LAUNCH_CHECK(ctx, "svo_fill_k");
        ctx->svo_prev_leaves = n_leaves;
    }
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->svo_prev_valid = ctx->dynamic;
    ctx->svo_n_inner = n_inner;

    ctx->svo_info.min[0] = mn[0]; ctx->svo_info.min[1] = mn[1]; ctx->svo_info.min[2] = mn[2];
    ctx->svo_info.depth = depth;
    ctx->svo_info.n_nodes = total_nodes;
    ctx->svo_info.n_leaves = n_leaves;
    ctx->svo_info.n_list_entries = n_list;
    ctx->svo_info.flags = flags;
    ctx->svo_built = true;
    ctx->svo_stale = false;
    return TGX_OK;
}

int tgx_svo_device_bytes(tgx_context* ctx, uint64_t* bytes, uint64_t* solid_voxels)
{
    if (!ctx || !bytes) return TGX_ERR_INVALID_ARG;
    if (!ctx->svo_built) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no SVO built");
    const tgx_svo_info& in = ctx->svo_info;
    const uint64_t nl = in.n_leaves;
    *bytes = (uint64_t)sizeof(svo_gnode) * in.n_nodes                  /* node records           */
             + (4ull * TGX_SVO_BLOCK_WORDS + 2ull * TGX_SVO_BLOCK_WORDS) * nl /* bits, word prefixes */
             + 4ull * (nl + 1) + 2ull * ctx->svo_n_solid               /* material offsets, pool */
             + 4ull * nl + 4ull * (nl + 1) + 4ull * in.n_list_entries  /* cluster lists          */
             + 8ull * ctx->svo_n_inner + 8ull * nl;                    /* traversal layout       */
    if (solid_voxels) *solid_voxels = ctx->svo_n_solid;
    return TGX_OK;
}

int tgx_svo_get_info(tgx_context* ctx, tgx_svo_info* out)
{
    if (!ctx || !out) return TGX_ERR_INVALID_ARG;
    if (!ctx->svo_built) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no SVO built");
    *out = ctx->svo_info;
    return TGX_OK;
}

}  /* extern "C" */

/* SVO files: "TGXSVO01", a header, the build's own node records, the leaf
 * counts, list offsets and lists, then per leaf its 1024 bit words and the
 * materials of its solid voxels only (in bit order), and an FNV-1a 64
 * checksum of everything before it. */
namespace {
struct svo_file_header {
    char         magic[8];
    uint32_t     version;
    uint32_t     n_inner;
    tgx_svo_info info;
    uint32_t     n_clusters;   /* of the scene the SVO was built from */
    uint32_t     n_luts;
    uint64_t     n_solid;      /* solid voxels over all leaves */
};
const char k_svo_magic[8] = {'T', 'G', 'X', 'S', 'V', 'O', '0', '1'};
constexpr uint32_t k_svo_chunk = 2048u;   /* leaves per staged chunk */

struct svo_stream {
    FILE*    f = nullptr;
    uint64_t h = 0xCBF29CE484222325ull;
    bool     bad = false;
    void hash(const void* p, size_t n)
    {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 0x100000001B3ull; }
    }
    void put(const void* p, size_t n)
    {
        if (bad || (n && fwrite(p, 1, n, f) != n)) { bad = true; return; }
        hash(p, n);
    }
    void get(void* p, size_t n)
    {
        if (bad || (n && fread(p, 1, n, f) != n)) { bad = true; return; }
        hash(p, n);
    }
    ~svo_stream() { if (f) fclose(f); }
};

inline uint32_t popc32(uint32_t v) { return (uint32_t)__builtin_popcount(v); }

/* The invariants tgx_svo_build's layout guarantees and the traversal relies
 * on: every index a kernel follows stays inside its array. */
const char* check_svo_nodes(const std::vector<svo_gnode>& g, uint32_t n_inner, uint32_t depth)
{
    const uint32_t n = (uint32_t)g.size(), leaf_level = depth - 5u;
    if (n == 0 || n_inner == 0 || n_inner > n) return "node counts";
    if (g[0].parent != SVO_NONE || ((g[0].masks >> 16) & 0xFFu) != 0u || g[0].cx || g[0].cy || g[0].cz) return "root";
    uint64_t claimed = 0;   /* children named by some parent; each names its parent */
    for (uint32_t i = 0; i < n; i++) {
        const svo_gnode& a = g[i];
        const uint32_t lvl = (a.masks >> 16) & 0xFFu, valid = a.masks & 0xFFu, leafm = (a.masks >> 8) & 0xFFu;
        if (a.masks >> 24) return "node masks";
        if (i >= n_inner) {
            if (lvl != leaf_level || valid || leafm || a.first_child != SVO_NONE || a.leaf_idx != i - n_inner)
                return "leaf node";
            continue;
        }
        if (lvl >= leaf_level || a.leaf_idx != SVO_NONE) return "inner node";
        const uint32_t kids = popc32(valid);
        if (kids == 0) {
            if (a.first_child != SVO_NONE || leafm) return "childless node";
            continue;
        }
        claimed += kids;
        const bool at_leaves = lvl + 1u == leaf_level;
        if (leafm != (at_leaves ? valid : 0u)) return "leaf mask";
        const uint64_t f = a.first_child, e = f + kids;
        if (f <= i || e > (at_leaves ? n : n_inner) || (at_leaves && f < n_inner)) return "child range";
        uint32_t slot = 0;
        for (uint32_t rc = 0; rc < 8u; rc++) {
            if (!(valid & (1u << rc))) continue;
            const svo_gnode& c = g[f + slot++];
            if (c.parent != i || ((c.masks >> 16) & 0xFFu) != lvl + 1u || c.cx != 2u * a.cx + (rc & 1u) ||
                c.cy != 2u * a.cy + ((rc >> 1) & 1u) || c.cz != 2u * a.cz + (rc >> 2))
                return "child record";
        }
    }
    if (claimed != n - 1u) return "unreachable nodes";
    return nullptr;
}
}  // namespace

extern "C" {

int tgx_svo_save(tgx_context* ctx, const char* path)
{
    if (!ctx || !path) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    /* the file records the scene's counts: a stale SVO would not match them */
    if (int rc0 = tgx_svo_check_current(ctx)) return rc0;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const tgx_svo_info in = ctx->svo_info;
    const uint32_t nl = in.n_leaves;
    std::vector<svo_gnode> g(in.n_nodes);
    std::vector<uint32_t> leaf_n(nl), off(nl + 1u), list(in.n_list_entries);
    TGX_HIP_CHECK(ctx, hipMemcpy(g.data(), ctx->svo_nodes.ptr, sizeof(svo_gnode) * g.size(), hipMemcpyDeviceToHost));
    if (nl) TGX_HIP_CHECK(ctx, hipMemcpy(leaf_n.data(), ctx->svo_leaf_n.ptr, 4ull * nl, hipMemcpyDeviceToHost));
    TGX_HIP_CHECK(ctx, hipMemcpy(off.data(), ctx->svo_list_off.ptr, 4ull * (nl + 1u), hipMemcpyDeviceToHost));
    if (!list.empty())
        TGX_HIP_CHECK(ctx, hipMemcpy(list.data(), ctx->svo_list.ptr, 4ull * list.size(), hipMemcpyDeviceToHost));
    /* the leaves' materials are stored as in the file: solid voxels only,
     * in bit order, from the leaf's material offset */
    std::vector<uint32_t> moff(nl + 1u);
    TGX_HIP_CHECK(ctx, hipMemcpy(moff.data(), ctx->svo_leaf_moff.ptr, 4ull * (nl + 1u), hipMemcpyDeviceToHost));
    std::vector<uint32_t> bits((size_t)k_svo_chunk * TGX_SVO_BLOCK_WORDS);
    std::vector<uint16_t> mats;
    const uint64_t n_solid = moff[nl];
    svo_stream s;
    s.f = fopen(path, "wb");
    if (!s.f) return tgx_fail(ctx, TGX_ERR_IO, "cannot write %s", path);
    svo_file_header hd = {};
    memcpy(hd.magic, k_svo_magic, 8);
    hd.version = 1u;
    hd.n_inner = ctx->svo_n_inner;
    hd.info = in;
    hd.n_clusters = ctx->n_clusters;
    hd.n_luts = ctx->n_luts;
    hd.n_solid = n_solid;
    s.put(&hd, sizeof(hd));
    s.put(g.data(), sizeof(svo_gnode) * g.size());
    s.put(leaf_n.data(), 4ull * nl);
    s.put(off.data(), 4ull * (nl + 1u));
    s.put(list.data(), 4ull * list.size());
    for (uint32_t l0 = 0; l0 < nl && !s.bad; l0 += k_svo_chunk) {
        const uint32_t m = std::min(k_svo_chunk, nl - l0);
        TGX_HIP_CHECK(ctx, hipMemcpy(bits.data(), (const uint32_t*)ctx->svo_leaf_bits.ptr + (size_t)l0 * TGX_SVO_BLOCK_WORDS,
                                     4ull * TGX_SVO_BLOCK_WORDS * m, hipMemcpyDeviceToHost));
        mats.resize((size_t)moff[l0 + m] - moff[l0]);
        if (!mats.empty())
            TGX_HIP_CHECK(ctx, hipMemcpy(mats.data(), (const uint16_t*)ctx->svo_leaf_mats.ptr + moff[l0],
                                         2ull * mats.size(), hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < m; k++) {
            s.put(&bits[(size_t)k * TGX_SVO_BLOCK_WORDS], 4ull * TGX_SVO_BLOCK_WORDS);
            s.put(mats.data() + (moff[l0 + k] - moff[l0]), 2ull * (moff[l0 + k + 1] - moff[l0 + k]));
        }
    }
    const uint64_t h = s.h;
    s.put(&h, sizeof(h));
    if (s.bad) return tgx_fail(ctx, TGX_ERR_IO, "write error on %s", path);
    return TGX_OK;
}

int tgx_svo_load(tgx_context* ctx, const char* path)
{
    if (!ctx || !path) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (ctx->n_graphs > 0)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "a frame graph uses the SVO: destroy it before loading");
    if (!ctx->has_scene) return tgx_fail(ctx, TGX_ERR_NO_SCENE, "load the SVO's scene first");
    svo_stream s;
    s.f = fopen(path, "rb");
    if (!s.f) return tgx_fail(ctx, TGX_ERR_IO, "cannot open %s", path);
    if (fseek(s.f, 0, SEEK_END) != 0) return tgx_fail(ctx, TGX_ERR_IO, "cannot seek in %s", path);
    const long fsize = ftell(s.f);
    rewind(s.f);
    svo_file_header hd;
    s.get(&hd, sizeof(hd));
    if (s.bad || memcmp(hd.magic, k_svo_magic, 8) != 0 || hd.version != 1u)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s is not an SVO file", path);
    const tgx_svo_info in = hd.info;
    /* the fixed-size parts must fit in the file before anything is allocated */
    const uint64_t fixed = sizeof(hd) + sizeof(svo_gnode) * (uint64_t)in.n_nodes + 4ull * (2ull * in.n_leaves + 1u) +
                           4ull * in.n_list_entries + 4ull * TGX_SVO_BLOCK_WORDS * in.n_leaves + 2ull * hd.n_solid + 8u;
    bool bad_min = false;
    for (int a = 0; a < 3; a++)
        bad_min |= !(in.min[a] == floorf(in.min[a])) ||
                   !(fabsf(in.min[a]) + (float)(1u << (in.depth & 15u)) < 16777216.0f);
    if (in.depth < TGX_SVO_MIN_DEPTH || in.depth > TGX_SVO_MAX_DEPTH || bad_min || in.n_nodes == 0 ||
        in.n_nodes >= 0xFFFFFFu || hd.n_inner > in.n_nodes || in.n_leaves != in.n_nodes - hd.n_inner ||
        hd.n_solid > (uint64_t)in.n_leaves * TGX_SVO_BLOCK_VOXELS || fsize < 0 || fixed != (uint64_t)fsize)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: inconsistent SVO header or file size", path);
    if (hd.n_clusters != ctx->n_clusters || hd.n_luts != ctx->n_luts)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s was built from a scene of %u clusters and %u LUTs, not %u and %u",
                        path, hd.n_clusters, hd.n_luts, ctx->n_clusters, ctx->n_luts);
    const uint32_t nl = in.n_leaves;
    std::vector<svo_gnode> g(in.n_nodes);
    std::vector<uint32_t> leaf_n(nl), off(nl + 1u), list(in.n_list_entries);
    s.get(g.data(), sizeof(svo_gnode) * g.size());
    s.get(leaf_n.data(), 4ull * nl);
    s.get(off.data(), 4ull * (nl + 1u));
    s.get(list.data(), 4ull * list.size());
    if (s.bad) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s is truncated", path);
    if (const char* why = check_svo_nodes(g, hd.n_inner, in.depth))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: bad SVO nodes (%s)", path, why);
    if (off[0] != 0u || off[nl] != in.n_list_entries) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: bad list offsets", path);
    for (uint32_t l = 0; l < nl; l++)
        if (off[l + 1] < off[l] || off[l + 1] - off[l] != leaf_n[l] ||
            g[hd.n_inner + l].leaf_n != leaf_n[l])
            return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: bad leaf lists", path);
    for (uint32_t c : list)
        if (c >= ctx->n_clusters) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: cluster index out of range", path);
    for (const svo_gnode& a : g)
        if (a.leaf_n && a.leaf_idx == SVO_NONE) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: bad node", path);

    /* Leaf data in two passes over the file: the first checks every leaf
     * and the checksum with nothing on the device touched, the second
     * uploads (and hashes again, against a file changed in between). */
    std::vector<uint32_t> bits((size_t)k_svo_chunk * TGX_SVO_BLOCK_WORDS), moff(nl + 1u);
    std::vector<uint16_t> mats, wpre((size_t)k_svo_chunk * TGX_SVO_BLOCK_WORDS);
    const long leaf_data_at = ftell(s.f);
    const uint64_t h_leaf = s.h;
    auto leaf_pass = [&](bool upload) -> int {
        uint64_t n_solid = 0;
        for (uint32_t l0 = 0; l0 < nl; l0 += k_svo_chunk) {
            const uint32_t m = std::min(k_svo_chunk, nl - l0);
            const uint64_t chunk_first = n_solid;
            mats.clear();
            for (uint32_t k = 0; k < m && !s.bad; k++) {
                uint32_t* b = &bits[(size_t)k * TGX_SVO_BLOCK_WORDS];
                s.get(b, 4ull * TGX_SVO_BLOCK_WORDS);
                uint32_t cnt = 0;
                for (uint32_t w = 0; w < TGX_SVO_BLOCK_WORDS; w++) {
                    wpre[(size_t)k * TGX_SVO_BLOCK_WORDS + w] = (uint16_t)cnt;
                    cnt += popc32(b[w]);
                }
                moff[l0 + k] = (uint32_t)n_solid;
                n_solid += cnt;
                if (s.bad || n_solid > hd.n_solid) { s.bad = true; break; }
                const size_t at = mats.size();
                mats.resize(at + cnt);
                s.get(mats.data() + at, 2ull * cnt);
                for (uint32_t j = 0; j < cnt && !s.bad; j++)
                    if (ctx->n_luts && (uint32_t)(mats[at + j] >> 8) >= ctx->n_luts) s.bad = true;
            }
            if (s.bad) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: truncated or bad leaf data", path);
            if (!upload) continue;
            TGX_HIP_CHECK(ctx, hipMemcpy((uint32_t*)ctx->svo_leaf_bits.ptr + (size_t)l0 * TGX_SVO_BLOCK_WORDS,
                                         bits.data(), 4ull * TGX_SVO_BLOCK_WORDS * m, hipMemcpyHostToDevice));
            TGX_HIP_CHECK(ctx, hipMemcpy((uint16_t*)ctx->svo_leaf_wpre.ptr + (size_t)l0 * TGX_SVO_BLOCK_WORDS,
                                         wpre.data(), 2ull * TGX_SVO_BLOCK_WORDS * m, hipMemcpyHostToDevice));
            if (!mats.empty())
                TGX_HIP_CHECK(ctx, hipMemcpy((uint16_t*)ctx->svo_leaf_mats.ptr + chunk_first, mats.data(),
                                             2ull * mats.size(), hipMemcpyHostToDevice));
        }
        if (n_solid != hd.n_solid) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: solid voxel count mismatch", path);
        moff[nl] = (uint32_t)n_solid;
        return TGX_OK;
    };
    int rc;
    if (leaf_data_at < 0) return tgx_fail(ctx, TGX_ERR_IO, "cannot tell the position in %s", path);
    if ((rc = leaf_pass(false))) return rc;
    const uint64_t h = s.h;
    uint64_t stored = 0;
    if (fread(&stored, sizeof(stored), 1, s.f) != 1 || stored != h)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: checksum mismatch", path);

    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->svo_built = false;
    ctx->svo_prev_valid = false;
    ctx->svo_reused = 0;
    const size_t nlr = nl ? nl : 1;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->svo_nodes, sizeof(svo_gnode) * g.size(), "svo nodes")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_bits, 4ull * TGX_SVO_BLOCK_WORDS * nlr, "svo leaf bits")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_mats, 2ull * (hd.n_solid ? hd.n_solid : 1), "svo leaf materials")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_moff, 4ull * (nlr + 1), "svo material offsets")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_wpre, 2ull * TGX_SVO_BLOCK_WORDS * nlr, "svo material prefixes")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_leaf_n, 4ull * nlr, "svo leaf n")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_list_off, 4ull * (nlr + 1), "svo list offsets")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_list, 4ull * (list.empty() ? 1 : list.size()), "svo list")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_inner, 8ull * hd.n_inner, "svo inner nodes")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->svo_bricks, 8ull * nlr, "svo leaf bricks")))
        return rc;
    if (fseek(s.f, leaf_data_at, SEEK_SET) != 0) return tgx_fail(ctx, TGX_ERR_IO, "cannot seek in %s", path);
    s.h = h_leaf;
    if ((rc = leaf_pass(true))) return rc;
    if (s.h != h) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s changed while loading", path);
    TGX_HIP_CHECK(ctx, hipMemcpy(ctx->svo_leaf_moff.ptr, moff.data(), 4ull * (nl + 1u), hipMemcpyHostToDevice));
    ctx->svo_n_solid = hd.n_solid;
    TGX_HIP_CHECK(ctx, hipMemcpy(ctx->svo_nodes.ptr, g.data(), sizeof(svo_gnode) * g.size(), hipMemcpyHostToDevice));
    if (nl) TGX_HIP_CHECK(ctx, hipMemcpy(ctx->svo_leaf_n.ptr, leaf_n.data(), 4ull * nl, hipMemcpyHostToDevice));
    TGX_HIP_CHECK(ctx, hipMemcpy(ctx->svo_list_off.ptr, off.data(), 4ull * (nl + 1u), hipMemcpyHostToDevice));
    if (!list.empty())
        TGX_HIP_CHECK(ctx, hipMemcpy(ctx->svo_list.ptr, list.data(), 4ull * list.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(svo_inner_k, dim3(blocks_for(hd.n_inner)), dim3(256), 0, ctx->stream,
                       (const svo_gnode*)ctx->svo_nodes.ptr, hd.n_inner, (uint2*)ctx->svo_inner.ptr);
    TGX_LAUNCH_CHECK(ctx, "svo_inner_k");
    if (nl)
        hipLaunchKernelGGL(svo_bricks_k, dim3(nl), dim3(64), 0, ctx->stream, (const uint32_t*)ctx->svo_leaf_bits.ptr,
                           (unsigned long long*)ctx->svo_bricks.ptr);
    TGX_LAUNCH_CHECK(ctx, "svo_bricks_k");
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->svo_info = in;
    ctx->svo_n_inner = hd.n_inner;
    ctx->svo_built = true;
    ctx->svo_stale = false;
    return TGX_OK;
}

int tgx_svo_export(tgx_context* ctx, tgx_svo_node* nodes_out, uint32_t* leaf_bits, uint16_t* leaf_materials,
                   uint32_t* leaf_n, uint32_t* list_offsets, uint32_t* list)
{
    if (!ctx) return TGX_ERR_INVALID_ARG;
    if (!ctx->svo_built) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no SVO built");
    const tgx_svo_info& in = ctx->svo_info;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (nodes_out) {
        std::vector<svo_gnode> g(in.n_nodes);
        TGX_HIP_CHECK(ctx, hipMemcpy(g.data(), ctx->svo_nodes.ptr, sizeof(svo_gnode) * in.n_nodes, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < in.n_nodes; i++) {
            tgx_svo_node& o = nodes_out[i];
            o.level = (g[i].masks >> 16) & 0xFFu;
            o.coord[0] = g[i].cx; o.coord[1] = g[i].cy; o.coord[2] = g[i].cz;
            o.valid_mask = g[i].masks & 0xFFu;
            o.leaf_mask = (g[i].masks >> 8) & 0xFFu;
            o.first_child = g[i].first_child;
            o.leaf_idx = g[i].leaf_idx;
        }
    }
    if (leaf_bits && in.n_leaves)
        TGX_HIP_CHECK(ctx, hipMemcpy(leaf_bits, ctx->svo_leaf_bits.ptr, 4ull * TGX_SVO_BLOCK_WORDS * in.n_leaves, hipMemcpyDeviceToHost));
    if (leaf_materials && in.n_leaves) {
        /* the dense canonical grid from the solid voxels' materials */
        const uint32_t nl = in.n_leaves;
        std::vector<uint32_t> moff(nl + 1u), bw((size_t)k_svo_chunk * TGX_SVO_BLOCK_WORDS);
        std::vector<uint16_t> pm;
        TGX_HIP_CHECK(ctx, hipMemcpy(moff.data(), ctx->svo_leaf_moff.ptr, 4ull * (nl + 1u), hipMemcpyDeviceToHost));
        for (uint32_t l0 = 0; l0 < nl; l0 += k_svo_chunk) {
            const uint32_t m = std::min(k_svo_chunk, nl - l0);
            TGX_HIP_CHECK(ctx, hipMemcpy(bw.data(), (const uint32_t*)ctx->svo_leaf_bits.ptr + (size_t)l0 * TGX_SVO_BLOCK_WORDS,
                                         4ull * TGX_SVO_BLOCK_WORDS * m, hipMemcpyDeviceToHost));
            pm.resize((size_t)moff[l0 + m] - moff[l0]);
            if (!pm.empty())
                TGX_HIP_CHECK(ctx, hipMemcpy(pm.data(), (const uint16_t*)ctx->svo_leaf_mats.ptr + moff[l0],
                                             2ull * pm.size(), hipMemcpyDeviceToHost));
            uint16_t* out = leaf_materials + (size_t)l0 * TGX_SVO_BLOCK_VOXELS;
            memset(out, 0, 2ull * TGX_SVO_BLOCK_VOXELS * m);
            size_t j = 0;
            for (size_t w = 0; w < (size_t)m * TGX_SVO_BLOCK_WORDS; w++)
                for (uint32_t word = bw[w]; word; word &= word - 1u)
                    out[32u * w + (uint32_t)__builtin_ctz(word)] = pm[j++];
        }
    }
    if (leaf_n && in.n_leaves)
        TGX_HIP_CHECK(ctx, hipMemcpy(leaf_n, ctx->svo_leaf_n.ptr, 4ull * in.n_leaves, hipMemcpyDeviceToHost));
    if (list_offsets)
        TGX_HIP_CHECK(ctx, hipMemcpy(list_offsets, ctx->svo_list_off.ptr, 4ull * (in.n_leaves + 1), hipMemcpyDeviceToHost));
    if (list && in.n_list_entries)
        TGX_HIP_CHECK(ctx, hipMemcpy(list, ctx->svo_list.ptr, 4ull * in.n_list_entries, hipMemcpyDeviceToHost));
    return TGX_OK;
}

} /* extern "C" */
