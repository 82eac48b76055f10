/*
 * tgx.h — C ABI of the MI355X (gfx950) voxel ray-tracing hot path.
 *
 * This library replaces the GPU half of tg's voxel ray tracer: the
 * Vulkan/GLSL passes behind tg_raytracer_render() and the CPU SVO build
 * behind tg_svo_create(). The host keeps tg's object/cluster data model
 * byte for byte; everything crosses this boundary as plain pointers and
 * sizes. Every entry point returns a status code (tgx_status); the
 * reference asserts instead (tg_common.h:32-47), so callers that relied on
 * a crash now get TGX_ERR_* and a message from tgx_last_error().
 *
 * Threading: one context per host thread. All work of a context is
 * ordered on the context's own HIP stream.
 *
 * Reference citations are file:line under torben-gross/tg.
 */
#ifndef TGX_H
#define TGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Constants of the data model (tg_graphics_core.h:18-23, cluster.inc:1-6)   */
/* ------------------------------------------------------------------------ */

#define TGX_CLUSTER_SIDE            8u    /* voxels per cluster edge          */
#define TGX_VOXELS_PER_CLUSTER      512u  /* 8^3, 9-bit voxel index 64z+8y+x  */
#define TGX_BITS_WORDS_PER_CLUSTER  16u   /* 512 solid bits as 16 x u32       */
#define TGX_CLUSTER_POINTER_BITS    31u   /* 31-bit cluster pointer field     */
#define TGX_LUT_ENTRIES             256u  /* entries per material LUT         */
#define TGX_VISIBILITY_EMPTY        0xFFFFFFFFFFFFFFFFull /* clear.comp:19   */

/* Visibility word: depth24 << 40 | cluster_pointer31 << 9 | voxel9
 * (visibility.frag:194-201). */
#define TGX_VIS_DEPTH_SHIFT         40u
#define TGX_VIS_CPTR_SHIFT          9u
#define TGX_VIS_CPTR_MASK           0x7FFFFFFFu
#define TGX_VIS_VOXEL_MASK          0x1FFu

/* Iteration cap of every DDA loop (oracle and kernels alike). An 8^3 walk
 * takes at most 22 steps; the cap only ends walks whose step sizes have
 * overflowed to infinity (direction components near FLT_MIN). */
#define TGX_DDA_MAX_ITERATIONS      64u

/* ------------------------------------------------------------------------ */
/* Byte layouts (std430 of the reference's SSBOs)                           */
/* ------------------------------------------------------------------------ */

/* Object record, 96 B. buffers.inc:6-13, tgvk_raytracer.c:35-42.
 * rotation is a column-major m4 (tg_math.h:240-275): rotation[4*c + r].
 * lut_idx occupies the reference's `pad` word (buffers.inc:11) and selects
 * the object's 256-entry material LUT. */
typedef struct tgx_object {
    uint32_t n_cluster_pointers_per_dim[3];
    uint32_t first_cluster_pointer;
    float    translation[3];
    uint32_t lut_idx;
    float    rotation[16];
} tgx_object;

/* Camera record, 88 B. buffers.inc:30-39, tgvk_raytracer.c:65-74.
 * camera.xyz = position; ray_* = normalised corner rays
 * (tgvk_core.c:435-444); w components are unused. */
typedef struct tgx_camera {
    float camera[4];
    float ray_bl[4];
    float ray_br[4];
    float ray_tr[4];
    float ray_tl[4];
    float near_plane;
    float far_plane;
} tgx_camera;

/* Material LUT entry, 16 B: fp32 albedo and roughness. The reference packs
 * RGBA8 (tgvk_raytracer.c:1131-1135); this path carries fp32 albedo. */
typedef struct tgx_lut_entry {
    float albedo[3];
    float roughness;
} tgx_lut_entry;

/* Host-side scene description; arrays are caller-owned and copied.
 *   cluster_pointers[cptr]          = cluster idx (cluster_pointer_ssbo)
 *   cluster_idx_to_object[cidx]     = object idx
 *   bits[16*cidx + w]               = solid bits, voxel v of cluster c is
 *                                     bit (512c+v)%32 of word (512c+v)/32
 *   materials[512*cidx + v]         = LUT index of voxel v (may be NULL)
 * (tgvk_raytracer.c:851-978, visibility.frag:139-148, shading.frag:128-131) */
typedef struct tgx_scene_desc {
    const tgx_object*    objects;
    uint32_t             n_objects;
    const uint32_t*      cluster_pointers;
    uint32_t             n_cluster_pointers;
    const uint32_t*      cluster_idx_to_object;
    uint32_t             n_clusters;
    const uint32_t*      bits;
    const uint8_t*       materials;
    const tgx_lut_entry* luts;
    uint32_t             n_luts;
} tgx_scene_desc;

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */

typedef enum tgx_status {
    TGX_OK                 = 0,
    TGX_ERR_INVALID_ARG    = 1,
    TGX_ERR_CAPACITY       = 2,
    TGX_ERR_HIP            = 3,
    TGX_ERR_NO_SCENE       = 4,
    TGX_ERR_NO_DEVICE      = 5,
    TGX_ERR_OUT_OF_MEMORY  = 6,
    TGX_ERR_NOT_BUILT      = 7,
    TGX_ERR_IO             = 8
} tgx_status;

typedef struct tgx_context tgx_context;

/* ------------------------------------------------------------------------ */
/* Context                                                                  */
/* ------------------------------------------------------------------------ */

/* Replaces tg_raytracer_create's device half (tgvk_raytracer.c:662-789):
 * selects HIP device `device`, creates the context's stream. */
int         tgx_create(int device, tgx_context** out_ctx);
/* Replaces tg_raytracer_destroy: frees every device buffer of the context. */
int         tgx_destroy(tgx_context* ctx);
/* Message of the last failing call on this context ("" if none). The
 * reference has no error channel (tgvk_common.h:25-35 asserts). */
const char* tgx_last_error(const tgx_context* ctx);
/* Blocks until all work on the context's stream has finished
 * (the fence wait of tgvk_raytracer.c:1560-1561). */
int         tgx_synchronize(tgx_context* ctx);

/* ------------------------------------------------------------------------ */
/* Scene                                                                    */
/* ------------------------------------------------------------------------ */

/* Uploads the whole scene (objects, cluster maps, solid bits, materials,
 * LUTs) into library-owned device buffers, replacing any previous scene.
 * materials may be NULL (every voxel then has material 0); luts may be
 * NULL (GI, cache and shading calls then refuse). Replacing the scene
 * marks the SVO stale: tracing calls return TGX_ERR_NOT_BUILT until
 * tgx_svo_update or tgx_svo_build. Refused (TGX_ERR_CAPACITY) while a
 * frame graph exists, as are tgx_scene_reserve and tgx_gen_scene_device.
 * Replaces the staging pushes of tg_raytracer_create_object
 * (tgvk_raytracer.c:805-992) and tg_raytracer_color_lut_set (:1122-1142).
 * Errors: TGX_ERR_CAPACITY if n_cluster_pointers >= 2^31 (the 31-bit
 * field, tg_graphics_core.h:21-22); TGX_ERR_INVALID_ARG if a cluster
 * pointer or object index is out of range. */
int tgx_scene_upload(tgx_context* ctx, const tgx_scene_desc* scene);

/* Copies the context's scene arrays back to host memory (each pointer may
 * be NULL; sizes are those of the uploaded or generated scene). */
int tgx_scene_export(tgx_context* ctx, tgx_object* objects, uint32_t* cluster_pointers,
                     uint32_t* cluster_idx_to_object, uint32_t* bits, uint8_t* materials,
                     tgx_lut_entry* luts);

/* Array sizes of the context's scene as tgx_scene_export writes them. For
 * a dynamic scene: n_objects and n_clusters are the capacities (records of
 * free object slots are zero, tg_object_is_initialized at
 * tgvk_raytracer.c:1069-1077) and n_cluster_pointers the live pointers. */
/* Sets one LUT entry of the scene (albedo and roughness, fp32): the
 * counterpart of tg_raytracer_color_lut_set (tgvk_raytracer.c:1122-1142),
 * which packs an RGBA8 colour into its single LUT. Entries are indexed
 * (lut_idx, entry < 256); SVO leaves store indices, so no rebuild is
 * needed, and frames (captured ones too) read the new value from their next
 * launch on. TGX_ERR_INVALID_ARG outside the scene's LUTs. */
int tgx_lut_set(tgx_context* ctx, uint32_t lut_idx, uint32_t entry, const tgx_lut_entry* value);

typedef struct tgx_scene_counts {
    uint32_t n_objects;
    uint32_t n_cluster_pointers;
    uint32_t n_clusters;
    uint32_t n_luts;
    uint32_t has_materials;
} tgx_scene_counts;
int tgx_scene_get_counts(tgx_context* ctx, tgx_scene_counts* out);

/* ------------------------------------------------------------------------ */
/* Dynamic scenes: objects created and destroyed one at a time              */
/* ------------------------------------------------------------------------ */

/* Empty scene with fixed capacities and the LUT array (the scene half of
 * tg_raytracer_create, tgvk_raytracer.c:687-712): object and cluster
 * indices are handed out from LIFO free stacks that start as
 * [capacity - 1, ..., 0]. Replaces any uploaded scene. */
int tgx_scene_reserve(tgx_context* ctx, uint32_t object_capacity, uint32_t cluster_capacity,
                      const tgx_lut_entry* luts, uint32_t n_luts);

/* tg_raytracer_create_object (tgvk_raytracer.c:805-868): pops an object
 * index and n = nx ny nz cluster indices, appends n cluster pointers after
 * the live ones, writes the object record (first pointer, extent in
 * clusters, translation, column-major rotation, LUT index) and both cluster
 * maps. bits (16 u32 per cluster) and materials (512 B per cluster), in
 * relative-pointer order x fastest, fill the new clusters (NULL: empty /
 * zero). tg fixes the rotation (7 degrees times the index about y) and
 * fills voxels on the host; here the caller passes both. */
int tgx_object_create(tgx_context* ctx, const uint32_t n_clusters_per_dim[3], const float translation[3],
                      const float rotation[16], uint32_t lut_idx, const uint32_t* bits,
                      const uint8_t* materials, uint32_t* out_object_idx);

/* tg_raytracer_destroy_object (tgvk_raytracer.c:994-1067): pushes the
 * object's cluster indices back in pointer order, shifts every later
 * cluster pointer down over the gap (on the device), lowers the first
 * pointer of every later object, clears the record and pushes the object
 * index back. Neither call touches the SVO: run tgx_svo_update (or
 * tgx_svo_build) before tracing the edited scene. Both calls are refused
 * while a frame graph exists. */
int tgx_object_destroy(tgx_context* ctx, uint32_t object_idx);

/* Scene files (host only; no context needed). tg keeps its scene in memory
 * (tgvk_raytracer.c:805-992); this format stores the same arrays byte for
 * byte with an FNV-1a 64 checksum. save: TGX_ERR_IO if the file cannot be
 * written. counts / load: TGX_ERR_IO if it cannot be opened,
 * TGX_ERR_INVALID_ARG if it is not a scene file, is truncated or fails the
 * checksum. load fills caller arrays sized by `counts` (as
 * tgx_scene_file_counts returned them; materials may be NULL, a file
 * without materials loads zeros) and refuses, with TGX_ERR_INVALID_ARG and
 * before writing any array, a file whose header counts differ from them. */
int tgx_scene_save(const char* path, const tgx_scene_desc* scene);
int tgx_scene_file_counts(const char* path, tgx_scene_counts* out);
int tgx_scene_load(const char* path, const tgx_scene_counts* counts, tgx_object* objects,
                   uint32_t* cluster_pointers, uint32_t* cluster_idx_to_object, uint32_t* bits,
                   uint8_t* materials, tgx_lut_entry* luts);

/* ------------------------------------------------------------------------ */
/* Primary visibility                                                       */
/* ------------------------------------------------------------------------ */

/* Clears n_pixels visibility words at device pointer vis_dev to all ones.
 * Replaces clear.comp (clear.comp:15-20) / tg_raytracer_clear
 * (tgvk_raytracer.c:1556-1603). */
int tgx_visibility_clear_device(tgx_context* ctx, uint64_t* vis_dev, uint64_t n_pixels);

/* Renders the visibility buffers of n_cams cameras at w x h into host
 * memory out_host[n_cams * w * h] (row-major, y = 0 is the top row).
 * Each word is the minimum over all cluster pointers whose 8^3 DDA hits a
 * solid voxel with depth <= 1 of the packed word (visibility.frag:22-208),
 * or TGX_VISIBILITY_EMPTY. Replaces the visibility draw of
 * tg_raytracer_render (tgvk_raytracer.c:1275-1317). */
int tgx_visibility(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams,
                   uint32_t w, uint32_t h, uint64_t* out_host);

/* As tgx_visibility, but writes into the device buffer out_dev
 * (n_cams * w * h words) and returns without waiting for the GPU. */
int tgx_visibility_device(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams,
                          uint32_t w, uint32_t h, uint64_t* out_dev);

/* Scratch of tgx_visibility_device_noalloc: one array of this many bytes
 * per GPU, held from the library's first context on that GPU until the
 * process ends. */
#define TGX_VIS_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* As tgx_visibility_device, but allocates nothing: no hipMalloc, no
 * hipFree, no allocating library routine, no HIP objects. No counterpart in
 * the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops). Every work buffer
 * (camera records, per-(camera, cluster pointer) records, tile counts and
 * lists, cull flags) comes from a fixed device array of
 * TGX_VIS_NOALLOC_SCRATCH_BYTES per GPU. Cameras run in batches of at most
 * 16 that fit it with at least 64 tile-list entries per tile; a call where
 * even one camera does not fit (116 B + 92 B per cluster pointer + 520 B
 * per 16x16 tile + 1 B per object + 1 B per cluster, 256-B aligned parts)
 * fails with TGX_ERR_CAPACITY. The results equal tgx_visibility_device's.
 * Every call on a GPU shares that GPU's array: calls that overlap on one
 * GPU (two streams, two contexts, two threads) race. */
int tgx_visibility_device_noalloc(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams,
                                  uint32_t w, uint32_t h, uint64_t* out_dev);

/* Host-side decode of one visibility word, replacing
 * tg_raytracer_get_hovered_voxel's unpack (tgvk_raytracer.c:1639-1653).
 * The 31-bit field is returned as what the GPU wrote: a cluster POINTER
 * (visibility.frag:199). Returns 1 if the word is a hit (depth < 1). */
int tgx_visibility_decode(uint64_t word, float* out_depth, uint32_t* out_cluster_pointer,
                          uint32_t* out_voxel_idx);

/* ------------------------------------------------------------------------ */
/* Sparse voxel octree (tg_sparse_voxel_octree.{c,h}, svo.inc)              */
/* ------------------------------------------------------------------------ */

#define TGX_SVO_BLOCK_SIDE          32u     /* leaf block edge (svo.inc:2)       */
#define TGX_SVO_BLOCK_WORDS         1024u   /* 32^3 bits as u32                  */
#define TGX_SVO_BLOCK_VOXELS        32768u
#define TGX_SVO_MIN_DEPTH           6u      /* root side 2^depth voxels          */
#define TGX_SVO_MAX_DEPTH           13u
/* Iteration cap of the traversal loop: the advance by exit + FLT_EPSILON
 * (tg_sparse_voxel_octree.c:709-712) can stop moving once |coords| >= 1. */
#define TGX_SVO_MAX_STEPS           4096u
/* Reproduce tg's leaf-bounds quirk: the cluster AABB of a leaf is taken
 * over corners 0,1,2,2,4,5,6,7 (tg_sparse_voxel_octree.c:183,191). Off by
 * default: the build uses all eight corners. */
#define TGX_SVO_FLAG_REFERENCE_CORNER_QUIRK 1u

typedef struct tgx_svo_info {
    float    min[3];          /* world position of the root's min corner     */
    uint32_t depth;           /* root side = 2^depth voxels, leaves 32^3     */
    uint32_t n_nodes;         /* inner and leaf nodes, root included         */
    uint32_t n_leaves;
    uint32_t n_list_entries;  /* total length of the leaf cluster lists      */
    uint32_t flags;
} tgx_svo_info;

/* Canonical node record. Nodes are ordered by (level, Morton code of
 * coord); the children of a node are contiguous in child-index order
 * (bit 1 = +x, 2 = +y, 4 = +z; tg_sparse_voxel_octree.c:356-361). The
 * reference allocates the same nodes in depth-first order with u16
 * relative child pointers (tg_sparse_voxel_octree.h:6-11); here pointers are
 * absolute u32 so depth 11-13 trees fit. */
typedef struct tgx_svo_node {
    uint32_t level;           /* 0 = root                                     */
    uint32_t coord[3];        /* node min = svo min + coord * 2^(depth-level) */
    uint32_t valid_mask;      /* children present                            */
    uint32_t leaf_mask;       /* children that are 32^3 leaf blocks           */
    uint32_t first_child;     /* node index of the first child, or ~0         */
    uint32_t leaf_idx;        /* leaves: index into the leaf arrays, else ~0  */
} tgx_svo_node;

/* Result of one SVO ray (tg_svo_traverse, tg_sparse_voxel_octree.c:558-740).
 * The hit key (block, voxel) names the voxel independently of node layout. */
typedef struct tgx_svo_hit {
    uint32_t hit;             /* 1 if a solid voxel was reached within range  */
    float    distance;        /* origin to the voxel entry                    */
    uint32_t block[3];        /* leaf block coord, in units of 32 voxels      */
    uint32_t voxel;           /* 1024 z + 32 y + x within the block           */
    int32_t  normal[3];       /* entry face of the voxel, axis-aligned        */
    uint32_t steps;           /* traversal loop iterations                    */
} tgx_svo_hit;

/* Builds the SVO of the uploaded scene: root [min, min + 2^depth)^3, 32^3
 * leaf blocks, every non-empty cluster (tg_svo_create,
 * tg_sparse_voxel_octree.c:466-542). Inner levels keep a cluster in a
 * child's list iff tg_intersect_aabb_obb_ignore_contact passes
 * (tg_physics.c:226-392); leaves point-sample block-voxel centres into
 * their clusters (tg_sparse_voxel_octree.c:102-317). Each leaf also
 * stores, per solid voxel, (lut_idx << 8 | material) of the first cluster
 * in its list that sampled it. min must be integral; depth in
 * [TGX_SVO_MIN_DEPTH, TGX_SVO_MAX_DEPTH]. */
int tgx_svo_build(tgx_context* ctx, const float min[3], uint32_t depth, uint32_t flags);

/* Rebuilds the SVO after object creates and destroys, with the last
 * build's min, depth and flags: the incremental update the reference
 * promises (README.md:43) but does not have (tg_svo_create runs once,
 * tgvk_raytracer.c:1187-1196). The nodes, lists and inner levels are
 * rebuilt; a leaf whose key and candidate clusters (indices, in pointer
 * order) equal those of a leaf of the last build, none of them placed since,
 * copies that leaf instead of re-sampling 32^3 voxels. The result equals
 * tgx_svo_build's on the same scene, byte for byte. Without a previous
 * build on the current dynamic scene (tgx_scene_reserve) every leaf is
 * sampled. reused_leaves / voxelized_leaves (may be NULL) receive the
 * split. TGX_ERR_NOT_BUILT without a previous tgx_svo_build. */
int tgx_svo_update(tgx_context* ctx, uint32_t* reused_leaves, uint32_t* voxelized_leaves);
int tgx_svo_get_info(tgx_context* ctx, tgx_svo_info* out);

/* Copies the SVO out in canonical form: nodes[n_nodes]; per leaf bits
 * (1024 u32), materials (32768 u16), n (clusters that sampled a solid
 * voxel), list offsets (n_leaves + 1) and cluster-index list. Any output
 * pointer may be NULL. */
int tgx_svo_export(tgx_context* ctx, tgx_svo_node* nodes, uint32_t* leaf_bits,
                   uint16_t* leaf_materials, uint32_t* leaf_n, uint32_t* list_offsets,
                   uint32_t* list);

/* SVO files. tg builds its SVO in memory on the first frame and keeps no
 * copy (tg_svo_create, tgvk_raytracer.c:1187-1196); a saved SVO loads in
 * place of that build. save writes the nodes, leaf lists, per-leaf bits and
 * the materials of solid voxels only, with an FNV-1a 64 checksum;
 * TGX_ERR_NOT_BUILT without an SVO, TGX_ERR_IO if the file cannot be
 * written. load needs the scene the SVO was built from (its cluster and LUT
 * counts are checked) and no frame graph; it checks the checksum, the file
 * size and every index the traversal follows (child ranges, parent links,
 * node levels and coordinates, leaf lists, cluster and LUT indices) and
 * fails with TGX_ERR_INVALID_ARG on any mismatch, TGX_ERR_IO if the file
 * cannot be opened. Every check, the checksum included, runs before the
 * device is touched, so a rejected file leaves the context's SVO as it was;
 * only a device error or a file changed during the load leaves no SVO built.
 * The loaded SVO exports and traces as the saved one did; the next
 * tgx_svo_update samples every leaf. */
/* Device memory the built SVO's data takes (node records, leaf bits,
 * material offsets and word prefixes, the solid voxels' materials, cluster
 * lists and the traversal layout) and its solid voxel count. A leaf keeps
 * materials for its solid voxels only (2 B each, in bit order) plus 2 KiB of
 * per-word prefix counts, where a dense grid would take 64 KiB. */
int tgx_svo_device_bytes(tgx_context* ctx, uint64_t* bytes, uint64_t* solid_voxels);

int tgx_svo_save(tgx_context* ctx, const char* path);
int tgx_svo_load(tgx_context* ctx, const char* path);

/* Traces n rays (world-space origins and directions, 3 floats each)
 * through the SVO; a hit farther than max_distance is reported as a miss.
 * Follows tg_svo_traverse: the origin is taken relative to the root
 * centre, the descent picks the child containing the current position,
 * leaves run tg_amanatides_woo (tg_amanatides_woo.c:3-114) and misses
 * advance to the child border plus FLT_EPSILON and pop finished nodes. */
int tgx_svo_trace(tgx_context* ctx, const float* origins, const float* directions, uint32_t n,
                  float max_distance, tgx_svo_hit* out_host);

/* SVO primary visibility, the second producer of the visibility word
 * (debug_visibility_svo.frag:27-70): one ray per pixel from the camera
 * along the normalised corner-ray lerp, traced through the SVO (the
 * traversal of tgx_svo_trace, no range limit). A hit at distance t with
 * d = t / far <= 1 writes u64(d * 16777215) << 40 | leaf node index << 9 |
 * (block voxel index % 512), the node index in the canonical order of
 * tgx_svo_export; other pixels get TGX_VISIBILITY_EMPTY. The reference's
 * GLSL traversal reports enter/far in a mixed frame; this follows the C
 * traversal's distance (tg_sparse_voxel_octree.c:558-740). */
int tgx_svo_visibility_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h,
                              uint64_t* vis_dev);

/* ------------------------------------------------------------------------ */
/* One-bounce global illumination through the SVO                           */
/* ------------------------------------------------------------------------ */

/* tg has no GI estimator: shading.frag binds the SVO but never traces it
 * (shading.frag:40-43) and the secondary-ray code is a disabled CPU debug
 * block (tgvk_raytracer.c:1390-1434). This estimator is defined here and
 * computed identically by the kernels and the CPU oracle:
 *   - primary surface: decode the visibility word, re-intersect the hit
 *     voxel in cluster space, major-axis face normal rotated to world space
 *     and the world hit point (shading.frag:116-229);
 *   - spp bounce rays per visible pixel, cosine-weighted about the normal:
 *     xorshift32 (util.inc:11-25) cube rejection to a unit vector s, then
 *     d = normalize(n + s), started sqrt(3) voxels along the normal n
 *     (tg's debug bounce starts sqrt(3) along d, tgvk_raytracer.c:1419,
 *     which lets grazing rays start behind an adjacent wall), traced
 *     through the SVO up to max_distance; a hit at negative distance (the
 *     start lies in solid) counts as occluded;
 *   - at a bounce hit facing the sun: LUT albedo of the voxel's material
 *     and one shadow ray towards the sun, started 1/64 voxel off the hit
 *     face, same range;
 *   - E = (1/spp) sum_k albedo_k * max(0, n_k . sun) * radiance / pi
 *     over lit bounce hits; radiance L = albedo_p * E (per RGB channel).
 *     Each lit hit returns the Lambertian radiance albedo_k (n_k . sun)
 *     radiance / pi, and under cosine-weighted sampling the mean of the
 *     returned radiance is the irradiance at the pixel divided by pi: E is
 *     that quantity (called "irradiance" in the outputs below), and L is
 *     the pixel's outgoing Lambertian radiance.
 * RNG state of (pixel i, sample s): murmur3_fmix32 (util.inc:47-56) of
 * i * spp + s + 0x9E3779B9 * (frame + 1), with 0 replaced by 0x6D2B79F5. */
typedef struct tgx_gi_params {
    float    sun_dir[3];      /* unit vector towards the sun                */
    float    sun_radiance;
    uint32_t spp;             /* bounce rays per visible pixel (1..64)      */
    float    max_distance;    /* range of bounce and shadow rays (voxels)   */
    uint32_t frame;           /* decorrelates the random streams            */
    uint32_t pad;
} tgx_gi_params;

/* Ray counts of the last GI call. */
typedef struct tgx_gi_stats {
    uint64_t visible_pixels;
    uint64_t bounce_rays;
    uint64_t bounce_hits;
    uint64_t shadow_rays;
} tgx_gi_stats;

/* Computes per-pixel RGB radiance (and optionally E, the irradiance / pi) of one camera
 * from its visibility buffer vis_dev (w * h words, as written by
 * tgx_visibility_device) with the SVO of tgx_svo_build. Outputs are device
 * buffers of 3 * w * h floats, row-major RGB; pixels without a hit get 0.
 * Returns without waiting for the GPU. */
int tgx_gi_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h,
                  const uint64_t* vis_dev, const tgx_gi_params* params,
                  float* radiance_dev, float* irradiance_dev);
/* Synchronises and returns the ray counts of the last tgx_gi_device call. */
int tgx_gi_last_stats(tgx_context* ctx, tgx_gi_stats* out);

/* One frame of the hot path: primary visibility of one camera into vis_dev
 * followed by tgx_gi_device into radiance_dev. */
int tgx_frame_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h,
                     const tgx_gi_params* params, uint64_t* vis_dev, float* radiance_dev,
                     float* irradiance_dev);

/* Pick readback: replaces tg_raytracer_get_hovered_voxel
 * (tgvk_raytracer.c:1605-1655), which reads one visibility word back and
 * returns its 31-bit field as a "cluster index" although the GPU wrote a
 * cluster pointer (the caller then indexes cluster_idx_to_object with it,
 * tg_application.c:228). Here each result carries the pointer, the cluster
 * index it maps to and the object, all resolved on the device. */
typedef struct tgx_pick_result {
    uint32_t hit;              /* depth < 1 */
    float    depth;            /* (word >> 40) / 16777215 */
    uint32_t cluster_pointer;  /* the word's 31-bit field (UINT32_MAX on a miss) */
    uint32_t cluster_idx;      /* cluster_pointers[cluster_pointer] */
    uint32_t object_idx;       /* cluster_idx_to_object[cluster_idx] */
    uint32_t voxel;            /* 9-bit voxel index 64 z + 8 y + x (UINT32_MAX on a miss) */
    float    position[3];      /* camera + depth far dir (shading.frag:231), if cam given */
} tgx_pick_result;

/* Picks n pixels (xy[2 i], xy[2 i + 1]) of a w x h visibility buffer on the
 * device; cam may be NULL (position then 0). Synchronous. */
int tgx_pick(tgx_context* ctx, const tgx_camera* cam, const uint64_t* vis_dev, uint32_t w, uint32_t h,
             const uint32_t* xy, uint32_t n, tgx_pick_result* out_host);

/* ------------------------------------------------------------------------ */
/* Shading pass over the visibility buffer (shading.frag:114-337)           */
/* ------------------------------------------------------------------------ */

/* Views, numbered as tg's TG_DEBUG_SHOW_* (shading.frag:16-25). */
#define TGX_SHADE_LIT               0u  /* ambient + GGX-lit LUT albedo (shading.frag:308-322) */
#define TGX_SHADE_OBJECT_INDEX      1u  /* hash colour of the object index */
#define TGX_SHADE_DEPTH             2u  /* min(1, 8 depth) as grey */
#define TGX_SHADE_CLUSTER_INDEX     3u  /* hash colour of the cluster index */
#define TGX_SHADE_VOXEL_INDEX       4u  /* hash colour of the 9-bit voxel index */
#define TGX_SHADE_BLOCKS            5u  /* as CLUSTER_INDEX (shading.frag:244) */
#define TGX_SHADE_COLOR_LUT_INDEX   6u  /* hash colour of the material (LUT entry) */
#define TGX_SHADE_COLOR             7u  /* LUT albedo */
#define TGX_SHADE_NORMAL            8u  /* face normal * 0.5 + 0.5 */
#define TGX_SHADE_SHADING           9u  /* lit with white albedo (shading.frag:291-306) */

typedef struct tgx_shade_params {
    uint32_t view;            /* TGX_SHADE_* */
    float    light_dir[3];    /* normalised here; tg: (0, 0.8, 0.3) */
    float    light_radiance;  /* tg: 3.0 */
    float    ambient;         /* ambient = ambient * albedo; tg: 0.1 */
    float    metallic;        /* tg: 0.1 */
    float    roughness;       /* tg: 0.8; negative: the LUT entry's roughness */
    float    miss[4];         /* RGBA of pixels without a hit; tg: (1, 0, 1, 1) */
} tgx_shade_params;

/* Shades every pixel of one camera's visibility buffer vis_dev (w * h words
 * from tgx_visibility_device) into rgba_dev (4 floats per pixel, row-major),
 * as shading.frag's full-screen pass does (out_color). tg's LUT is RGBA8;
 * this path reads the fp32 albedo (and optionally roughness) of the LUT
 * entry. Returns without waiting for the GPU. */
int tgx_shade_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h,
                     const uint64_t* vis_dev, const tgx_shade_params* params, float* rgba_dev);

/* ------------------------------------------------------------------------ */
/* Tuning: launch shapes and queue policies. Results never depend on them;  */
/* tgx_autotune picks them by timing.                                        */
/* ------------------------------------------------------------------------ */

typedef struct tgx_tuning {
    uint32_t trace_sort;          /* bit 0: sort bounce rays, bit 1: sort shadow rays,
                                     by (direction octant, origin cell) before tracing (0) */
    uint32_t trace_refill;        /* idle lanes of a wave that trigger a queue refill, 1..64 (5) */
    uint32_t trace_blocks_per_cu; /* persistent 1024-lane trace workgroups per CU, 1..2 (2) */
    uint32_t trace_vote;          /* 0: every iteration runs node steps and DDA bursts for the
                                     lanes needing them; k in 1..64: a wave runs only DDA
                                     bursts once >= k/64 of its active lanes are in a leaf,
                                     else only node steps (0) */
    uint32_t trace_regroup;       /* fused GI launch: 0 runs each lane's phase in its own wave
                                     slot; k in 1..8 regroups the rays of a 1024-lane workgroup
                                     by phase after every step (one workgroup per CU), node-walk
                                     waves taking up to k node steps per step (0) */
    uint32_t trace_skip;          /* regrouped launch: 1 crosses 8^3 bricks without solid voxels
                                     in one exact step of the leaf DDA instead of voxel by
                                     voxel (0) */
    uint32_t vis_setup_blocks;    /* visibility: workgroups per camera of the grid-stride cull
                                     and tile-list scatter passes, 64..16384 (2048) */
    uint32_t trace_dda;           /* lane-machine launch: 1 runs a wave's leaf DDA without the
                                     iteration count (two steps per exit test) when no DDA lane's
                                     direction has a component below 2^-100 (1) */
    uint32_t vis_tile;            /* visibility: tile edge in pixels, 8, 16 or 32 (16); one
                                     workgroup of vis_tile^2 lanes walks each tile's list */
    uint32_t trace_cross;         /* uncounted DDA: 0 walks every voxel; k in 1..64 makes a
                                     wave cross 8^3 bricks without solid voxels in one exact
                                     step once >= k/64 of its walking lanes are in one (40) */
    uint32_t trace_node;          /* lane-machine launch: node-loop iterations a descending lane
                                     may take per step before its DDA burst, 1..4 (3) */
    uint32_t trace_burst;         /* uncounted DDA: iterations per burst, even, 2..32 (8) */
    uint32_t trace_chunk;         /* lane-machine launch: rays a wave takes from the queue per
                                     cursor atomic, kept for its next refills, 0..256 (32); 0
                                     takes only the idle lanes' rays. A wave takes only what it
                                     hands out once fewer than two chunks per wave are left */
    uint32_t trace_small;         /* lane-machine launch with more than one workgroup per CU: a
                                     queue of fewer than trace_small rays per lane of one
                                     workgroup per CU is traced by one workgroup per CU (the
                                     others exit at once), 0..64 (8); 0 keeps every workgroup */
} tgx_tuning;

int tgx_tuning_get(tgx_context* ctx, tgx_tuning* out);
int tgx_tuning_set(tgx_context* ctx, const tgx_tuning* tuning);

/* ------------------------------------------------------------------------ */
/* Batched frames as captured HIP graphs. The reference renders one camera  */
/* per tg_raytracer_render (tgvk_raytracer.c:1144-1553), recording and      */
/* submitting its visibility, SVO and shading command buffers every frame;  */
/* here a frame of n cameras is captured once and replayed.                 */
/* ------------------------------------------------------------------------ */

#define TGX_FRAME_NEURAL 1u   /* per camera the cache frame (tgx_nrc_frame_device) instead of tgx_gi_device */

typedef struct tgx_frame_graph tgx_frame_graph;

typedef struct tgx_frame_stats {
    uint64_t primary_rays;          /* n_cams * w * h */
    uint64_t visible_pixels;
    uint64_t bounce_rays;           /* plain GI, or the cache's fallback pixels */
    uint64_t shadow_rays;
    uint64_t bounce_hits;
    uint64_t train_rays;            /* bounce + shadow rays of training targets */
    uint64_t fallback_pixels;
} tgx_frame_stats;

/* Runs one frame of n_cams cameras at w x h without capture (the work a
 * graph records, see below); with stats_out it waits and fills the totals. */
int tgx_frame_batch(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                    const tgx_gi_params* gi, uint32_t flags, uint64_t* vis_dev, float* radiance_dev,
                    float* irradiance_dev, tgx_frame_stats* stats_out);

/* Captures one frame of n_cams cameras at w x h: visibility of all cameras
 * into vis_dev (n_cams * w * h words, camera-major), then per camera c GI
 * (or the cache frame) into radiance_dev + 3 c w h (and irradiance_dev,
 * which may be null). Runs the frame once uncaptured first (its results
 * are those of the first replay's inputs, advanced by one cache frame).
 * While any graph exists the context refuses calls that would allocate,
 * free or move buffers (scene upload, SVO build, cache creation, larger
 * resolutions). Captured frames trace rays unsorted (trace_sort ignored). */
int tgx_frame_graph_create(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                           const tgx_gi_params* gi, uint32_t flags, uint64_t* vis_dev, float* radiance_dev,
                           float* irradiance_dev, tgx_frame_graph** out);
/* Replays the frame on the context stream; returns without waiting. */
int tgx_frame_graph_launch(tgx_context* ctx, tgx_frame_graph* g);
/* Waits and returns the totals of the last replay. */
int tgx_frame_graph_stats(tgx_context* ctx, tgx_frame_graph* g, tgx_frame_stats* out);
int tgx_frame_graph_destroy(tgx_context* ctx, tgx_frame_graph* g);

/* Times the frame (uncaptured, `frames` runs per candidate after one
 * untimed run) by coordinate descent over the tuning knobs: trace launch
 * form {0, 1, 2, 4}, empty-brick skip, trace workgroups per CU {1, 2},
 * uncounted DDA, node iterations per step {1, 2, 3, 4}, empty-brick
 * crossing {0, 24, 32, 40, 48}, refill thresholds {4, 8, 12, 16, 24, 32}, visibility
 * workgroups per camera {256, 1024, 2048, 4096} and visibility tile edge
 * {8, 16, 32}. Keeps the fastest in the context's tuning and returns it in
 * best_out (may be null). Every candidate gives the same results; cache
 * frames advance training once per run, so a cache that is still learning
 * makes later candidates look faster (tune the plain form and keep its
 * tuning for the cache form: the kernels are the same). While a frame
 * graph exists, candidates that would need larger buffers are skipped. */
int tgx_autotune(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                 const tgx_gi_params* gi, uint32_t flags, uint64_t* vis_dev, float* radiance_dev,
                 float* irradiance_dev, uint32_t frames, tgx_tuning* best_out);

/* ------------------------------------------------------------------------ */
/* Neural irradiance cache. No counterpart in the reference: a small MLP    */
/* that predicts the irradiance E of tgx_gi_device from the primary hit,    */
/* trained online on samples the SVO tracer generates, with the tracer as   */
/* its fallback.                                                            */
/*                                                                          */
/* Network: 96 input features -> 64 -> 64 -> 64 -> 64 -> 3, ReLU after the  */
/* four hidden layers, no biases, linear output. Features: for axis a and   */
/* frequency k in 0..11, sin and cos of 2^k pi p_hat[a] at 24a + 2k (+1);   */
/* normal (72..74), albedo (75..77), constant 1 (78), zeros (79..95), with  */
/* p_hat = (p - origin) * scale in fp32. Weights and activations are BF16   */
/* on MFMA with FP32 accumulation; master weights and Adam are FP32; the    */
/* loss is the mean squared error over samples and RGB channels.            */
/* ------------------------------------------------------------------------ */

#define TGX_NRC_INPUTS      96u
#define TGX_NRC_HIDDEN      64u
#define TGX_NRC_N_PARAMS    18624u  /* W0[96][64], W1..W3[64][64], W4[64][3], row-major [in][out] */
#define TGX_NRC_BATCH_ALIGN 512u    /* train_batch granularity */

typedef struct tgx_nrc_config {
    float    origin[3];             /* position encoding: p_hat = (p - origin) * scale */
    float    scale;
    float    learning_rate;         /* Adam (1e-3) */
    float    beta1;                 /* 0.9 */
    float    beta2;                 /* 0.99 */
    float    epsilon;               /* 1e-8 */
    uint32_t seed;                  /* weight initialisation */
    uint32_t train_samples;         /* per frame of tgx_nrc_frame_device (65536) */
    uint32_t train_batch;           /* samples per Adam step, multiple of 512 (16384) */
    uint32_t train_spp;             /* SVO bounce samples per training target (4) */
    float    fallback_threshold;    /* cell relative squared error above which pixels are traced
                                       (relative to t^2, or to the tracer's variance) */
    uint32_t fallback_min_samples;  /* cells with fewer (decayed) training samples are traced */
    uint32_t cell_size;             /* edge of an error-estimate cell in voxels, power of two */
    uint32_t flags;                 /* TGX_NRC_FLAG_* */
    float    ema_decay;             /* 0: inference uses the trained weights. In [0, 1): inference
                                       uses their exponential moving average, e = d e + (1 - d) w
                                       after every Adam step (the trained weights still train) */
} tgx_nrc_config;

#define TGX_NRC_FLAG_NO_FALLBACK    1u  /* never trace: every visible pixel uses the cache */
#define TGX_NRC_FLAG_ALL_FALLBACK   2u  /* trace every visible pixel (plain GI through this path) */
#define TGX_NRC_FLAG_NO_TRAINING    4u  /* frames run inference only */
/* Fallback where the cell's squared bias exceeds fallback_threshold times
 * the variance the plain estimator would have at the frame's spp (measured
 * on the training targets), instead of fallback_threshold times t^2: with
 * threshold 1 a pixel is traced when tracing is expected to be closer. */
#define TGX_NRC_FLAG_FALLBACK_VS_TRACER 8u
/* Frames of several cameras (frame graphs, tgx_frame_batch): every camera
 * renders with the cache the previous frame left, then one training pass
 * takes train_samples / train_batch Adam steps, each on train_batch
 * samples of every camera, and the error estimate decays once per frame.
 * Without the flag the cameras run as consecutive single-camera frames
 * (train_samples of training per camera). One camera: the same frame
 * either way. */
#define TGX_NRC_FLAG_JOINT_TRAINING 16u

typedef struct tgx_nrc_stats {
    uint64_t visible_pixels;
    uint64_t fallback_pixels;       /* traced by the SVO tracer this frame */
    uint64_t train_samples;
    uint64_t train_rays;            /* bounce + shadow rays traced for training targets */
    uint64_t fallback_rays;         /* bounce + shadow rays traced for fallback pixels */
    uint32_t adam_steps;            /* total since tgx_nrc_create */
    float    train_loss;            /* mean loss of the frame's training batches */
} tgx_nrc_stats;

/* (Re)creates the cache: He-uniform weights from cfg->seed, zero Adam state
 * and an empty error estimate. */
int tgx_nrc_create(tgx_context* ctx, const tgx_nrc_config* cfg);

/* Master weights in the TGX_NRC_N_PARAMS layout (host arrays). set also
 * refreshes the BF16 copies and sets the moving average (ema_decay) to the
 * same weights; Adam state is kept. */
int tgx_nrc_get_params(tgx_context* ctx, float* params);
/* The fp32 weights inference packs to BF16: the moving average when
 * ema_decay > 0, else the master weights. New: the reference has no
 * learned part. */
int tgx_nrc_get_inference_params(tgx_context* ctx, float* params);
int tgx_nrc_set_params(tgx_context* ctx, const float* params);

/* Inference on n host samples of 9 floats (position, normal, albedo) into
 * out[3 n] (RGB irradiance). Synchronous. */
int tgx_nrc_infer(tgx_context* ctx, const float* samples, uint32_t n, float* out);

/* One Adam step on n host samples (9 floats each) with targets[3 n]. The
 * mean loss goes to *loss and, if grad is not null, the gradient (before
 * the step, TGX_NRC_N_PARAMS floats) to grad. Synchronous. */
int tgx_nrc_train_batch(tgx_context* ctx, const float* samples, const float* targets, uint32_t n,
                        float* loss, float* grad);

/* One frame with the cache: primary visibility into vis_dev, the surface of
 * every visible pixel, cfg->train_samples training targets traced with the
 * SVO estimator (gi->spp replaced by train_spp) at visible pixels chosen by
 * a hash of (sample, frame), Adam steps of train_batch samples, inference
 * for every visible pixel, and the plain SVO estimator (gi->spp) for pixels
 * whose error-estimate cell is unreliable. radiance = albedo * E. */
int tgx_nrc_frame_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h,
                         const tgx_gi_params* gi, uint64_t* vis_dev, float* radiance_dev,
                         float* irradiance_dev);
int tgx_nrc_last_stats(tgx_context* ctx, tgx_nrc_stats* out);

/* The cache's training state in a file: its configuration, the frame and
 * Adam-step counters, the fp32 master weights, both Adam moments, the
 * moving average and the error-estimate table, sealed by an FNV-1a 64
 * checksum. tgx_nrc_load re-creates the cache from it (TGX_ERR_IO if the
 * file cannot be opened; TGX_ERR_INVALID_ARG for a wrong size, magic,
 * layout, checksum, configuration or non-finite value, leaving the current
 * cache as it was; refused while a frame graph exists). Frames after a load
 * equal the frames the saved cache would have run next, bit for bit. New:
 * the reference has no learned part; this is the resume point of online
 * training. */
int tgx_nrc_save(tgx_context* ctx, const char* path);
int tgx_nrc_load(tgx_context* ctx, const char* path);

/* ------------------------------------------------------------------------ */
/* Device memory and profiling helpers (for callers without their own       */
/* allocator; the reference's equivalents are tgvk_memory.c:182-240)        */
/* ------------------------------------------------------------------------ */

int tgx_device_alloc(tgx_context* ctx, size_t bytes, void** out_dev);
int tgx_device_free(tgx_context* ctx, void* dev);
int tgx_memcpy_d2h(tgx_context* ctx, void* dst_host, const void* src_dev, size_t bytes);
int tgx_memcpy_h2d(tgx_context* ctx, void* dst_dev, const void* src_host, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (tgx_copy_k: 16-byte
 * loads and stores, bytes a multiple of 16); returns without waiting. Also
 * the known-size reference for HBM byte counters. */
int tgx_memcpy_d2d(tgx_context* ctx, void* dst_dev, const void* src_dev, size_t bytes);

/* Per-kernel timing with HIP events recorded on the context's stream.
 * When enabled, every kernel launch of the context is bracketed by events;
 * tgx_profile_get() returns, per kernel name, the launch count and the
 * summed event time in milliseconds. */
typedef struct tgx_kernel_stat {
    char     name[48];
    uint64_t launches;
    double   total_ms;
} tgx_kernel_stat;

int tgx_profile_enable(tgx_context* ctx, int enable);
int tgx_profile_reset(tgx_context* ctx);
int tgx_profile_get(tgx_context* ctx, tgx_kernel_stat* out, uint32_t capacity, uint32_t* out_n);

/* Statistics of the last visibility call (work items after culling). */
typedef struct tgx_visibility_stats {
    uint64_t clusters_total;     /* cluster pointers x cameras              */
    uint64_t clusters_binned;    /* (cluster, camera) pairs surviving cull  */
    uint64_t tile_entries;       /* (tile, cluster) pairs after binning     */
    uint32_t tiles;              /* tiles per camera                        */
    uint32_t tile_size;          /* tile edge in pixels                     */
} tgx_visibility_stats;

/* Synchronises the context stream and returns the statistics of the last
 * tgx_visibility / tgx_visibility_device call. */
int tgx_visibility_last_stats(tgx_context* ctx, tgx_visibility_stats* out);

/* Tile-list capacity in entries; 0 (the default) sizes it automatically
 * from the largest list total seen. A frame whose lists exceed the
 * capacity is still exact: its overflowing tiles test every record. No
 * reference counterpart (tg rasterises instead of binning). */
int tgx_visibility_set_list_capacity(tgx_context* ctx, uint32_t entries);

#ifdef __cplusplus
}
#endif

#endif /* TGX_H */
