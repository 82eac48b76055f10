/*
 * tgx_oracle.h — CPU oracle of the voxel ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar C restatement of tg's
 * algorithms, built with strict IEEE fp32 (-ffp-contract=off
 * -fno-fast-math). Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it, and only as the checker; the product library never
 * links it.
 *
 * Byte layouts are those of include/tgx.h (the data model).
 */
#ifndef TGX_ORACLE_H
#define TGX_ORACLE_H

#include "../include/tgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scene view used by every oracle function (arrays are not copied). */
typedef struct tgxo_scene {
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
} tgxo_scene;

/* Candidate word of one (pixel, cluster pointer) pair, or all ones
 * (visibility.frag:22-208). */
uint64_t tgxo_vis_pair(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                       uint32_t px, uint32_t py, uint32_t cptr);

/* Visibility buffer by exhaustive min over all (pixel, cluster pointer). */
void tgxo_visibility_brute(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                           uint64_t* out);

/* Same result through a conservative per-cluster screen rectangle. Rows
 * with y % row_step != 0 are left untouched (row_step = 1: all rows).
 * n_threads <= 0 uses the OpenMP default. Returns the number of
 * (pixel, cluster) pairs evaluated. */
uint64_t tgxo_visibility_culled(const tgxo_scene* s, const tgx_camera* cam, uint32_t w,
                                uint32_t h, uint32_t row_step, int n_threads, uint64_t* out);

/* Conservative pixel rectangle of one cluster pointer: returns 0 if no
 * pixel ray can hit it, else 1 with the inclusive rectangle. */
int tgxo_cluster_rect(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                      uint32_t cptr, uint32_t rect[4]);

/* Algorithmic bytes of a visibility frame: 64 x unique 64-B lines of the
 * input arrays that the culled oracle reads for pairs whose slab test
 * passes, plus 8 bytes per output pixel. */
uint64_t tgxo_visibility_alg_bytes(const tgxo_scene* s, const tgx_camera* cam, uint32_t w,
                                   uint32_t h);

/* Slab test (collide.inc:3-24, tg_physics.c:394-406). */
int tgxo_intersect_ray_aabb(const float o[3], const float d[3], const float mn[3],
                            const float mx[3], float* enter, float* exit_);

/* The reference's m4 helpers, exported for the closed-form tests. */
void tgxo_m4_mul(const float a[16], const float b[16], float out[16]);
void tgxo_m4_inverse(const float a[16], float out[16]);
void tgxo_world_to_cluster(const tgx_object* o, uint32_t rel_cptr, float out[16]);

/* ------------------------------------------------------------------------ */
/* Sparse voxel octree (tgx_oracle_svo.c)                                   */
/* ------------------------------------------------------------------------ */

typedef struct tgxo_svo tgxo_svo;

/* Depth-first build as tg_svo_create; NULL on bad depth or allocation. */
tgxo_svo* tgxo_svo_build(const tgxo_scene* s, const float min[3], uint32_t depth, uint32_t flags);
void      tgxo_svo_free(tgxo_svo* svo);
void      tgxo_svo_get_info(const tgxo_svo* svo, tgx_svo_info* out);
/* Canonical export (same layout as tgx_svo_export). */
int       tgxo_svo_export(const tgxo_svo* svo, tgx_svo_node* nodes, uint32_t* leaf_bits,
                          uint16_t* leaf_materials, uint32_t* leaf_n, uint32_t* list_offsets,
                          uint32_t* list);
void      tgxo_svo_traverse(const tgxo_svo* svo, const float origin[3], const float dir[3],
                            float max_distance, tgx_svo_hit* out);
void      tgxo_svo_trace(const tgxo_svo* svo, const float* origins, const float* dirs, uint32_t n,
                         float max_distance, tgx_svo_hit* out);
/* One-bounce GI of one camera from its visibility buffer (estimator of
 * tgx_gi_params); rows y % row_step != 0 are skipped. counts = visible
 * pixels, bounce rays, bounce hits, shadow rays. */
void      tgxo_gi(const tgxo_svo* svo, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis,
                  const tgx_gi_params* params, uint32_t row_step, float* radiance, float* irradiance,
                  uint64_t counts[4]);
/* Algorithmic bytes of one GI frame (the frame of tgxo_gi at row_step 1):
 * 64 x the unique 64-B lines the oracle's execution reads, per array in
 * the reference's layouts, plus the bytes it writes. out[0] = surfaces:
 * the visibility words and the scene lines of visible pixels (pointer map,
 * cluster map, object record, material); out[1] = traversals: SVO node
 * (u32 each, build order), leaf data (260 B) and voxel-pool (4 KiB per
 * leaf) lines of every bounce and shadow ray, the leaf-material (2 B per
 * voxel) and LUT lines of lit bounce hits; out[2] =
 * radiance written (12 B per pixel, 24 with irradiance); out[3] = total.
 * out[4..14] = the line counts: nodes, leaf data, voxel pool, leaf
 * materials, LUT (hits), pointer map, cluster map, objects, scene
 * materials, LUT (surfaces), visibility words. A line two kernels read
 * counts in both parts. Returns 0, or -1 if a bitmap could not be
 * allocated. */
int       tgxo_gi_alg_bytes(const tgxo_svo* svo, const tgx_camera* cam, uint32_t w, uint32_t h,
                            const uint64_t* vis, const tgx_gi_params* params, uint32_t with_irradiance,
                            uint64_t out[16]);
int       tgxo_gi_surface(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h, uint32_t px,
                          uint32_t py, uint64_t word, float out_p[3], float out_n[3], float out_albedo[3]);
/* Number of pixels of vis (w * h words of camera cam) that have a primary
 * surface by tgxo_gi_surface, i.e. the visible pixels a GI frame lists
 * (OpenMP over rows; n_threads <= 0 keeps the default). */
uint64_t  tgxo_gi_surface_count(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h,
                                const uint64_t* vis, int n_threads);
uint32_t  tgxo_gi_seed(uint32_t pixel, uint32_t sample, uint32_t spp, uint32_t frame);
/* Cosine-weighted bounce direction about unit n from one RNG seed. */
void      tgxo_gi_direction(const float n[3], uint32_t seed, float out[3]);
int       tgxo_aabb_obb_ignore_contact(const float mn[3], const float mx[3], const float corners[24]);
uint64_t  tgxo_svo_node_key(uint32_t level, const uint32_t coord[3]);

/* ------------------------------------------------------------------------ */
/* Shading pass (tgx_oracle_shade.c, shading.frag:114-337)                   */
/* ------------------------------------------------------------------------ */

void tgxo_shade_pixel(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h, uint32_t px,
                      uint32_t py, uint64_t word, const tgx_shade_params* params, float out[4]);
void tgxo_shade(const tgxo_scene* s, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis,
                const tgx_shade_params* params, float* rgba);

#ifdef __cplusplus
}
#endif

#endif
