/*
 * tgx_internal.h — context state shared by the library's translation units.
 */
#ifndef TGX_INTERNAL_H
#define TGX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/tgx.h"

/* Device buffer that only grows (never while a frame graph exists). */
struct tgx_buffer {
    void*  ptr = nullptr;
    size_t bytes = 0;
};

struct tgx_pending_event {
    int         stat_idx;
    hipEvent_t  start;
    hipEvent_t  stop;
};

struct tgx_context {
    int          device = 0;
    hipStream_t  stream = nullptr;
    std::string  last_error;

    /* Scene (device copies of tgx_scene_desc arrays). */
    bool         has_scene = false;
    uint32_t     n_objects = 0;
    uint32_t     n_cluster_pointers = 0;
    uint32_t     n_clusters = 0;
    uint32_t     n_luts = 0;
    bool         has_materials = false; /* the current scene carries materials (else all 0) */
    tgx_buffer   objects, cluster_pointers, cluster_idx_to_object, bits, materials, luts;

    /* Visibility work buffers. */
    tgx_buffer   vis_clusters;   /* compacted per-(camera, cluster) records */
    tgx_buffer   vis_solid;      /* u8 per cluster index: has a solid voxel */
    tgx_buffer   vis_objcnt;     /* 3 u32 per (camera, object): cull lists  */
    tgx_buffer   vis_occ;        /* occupied cluster pointers: flags, scan, list (3 n_cptr u32) */
    tgx_buffer   vis_obj_occ;    /* uint2 per object: first occupied entry, count */
    tgx_buffer   vis_scan_sums;  /* per-workgroup sums of the device scans + grand totals */
    uint32_t     vis_n_occ = 0;  /* occupied cluster pointers of the scene  */
    uint64_t     scene_version = 0;     /* bumped by every write of cluster bits */
    uint64_t     vis_solid_version = ~0ull; /* scene_version vis_solid was made for */
    tgx_buffer   vis_rec_flags;  /* 1 if a record is in the tile lists      */
    tgx_buffer   vis_global_keys;/* records with screen-sized rectangles    */
    tgx_buffer   vis_counters;   /* per-camera record counts + scalars      */
    tgx_buffer   vis_tile_count; /* per-tile entry counts                   */
    tgx_buffer   vis_tile_off;   /* per-tile list offsets (exclusive scan)  */
    tgx_buffer   vis_tile_list;  /* record indices per tile                 */
    tgx_buffer   vis_cams;       /* camera constants                        */
    tgx_buffer   vis_out;        /* staging output for tgx_visibility       */
    tgx_buffer   vis_stats_dev;  /* frame statistics (records, entries, max)*/
    uint32_t     vis_list_capacity = 0;   /* tile-list entries allocated    */
    uint32_t     vis_list_capacity_fixed = 0; /* caller override, 0 = auto */
    hipEvent_t   vis_stats_event = nullptr;
    bool         vis_stats_pending = false;
    void*        host_pinned = nullptr; /* small pinned readback area       */
    tgx_visibility_stats vis_stats = {};

    /* Sparse voxel octree (built by tgx_svo_build). */
    bool         svo_built = false;
    tgx_svo_info svo_info = {};
    tgx_buffer   svo_nodes;      /* svo_gnode[n_nodes], canonical order     */
    tgx_buffer   svo_leaf_bits;  /* 1024 u32 per leaf                       */
    tgx_buffer   svo_leaf_mats;  /* u16 per solid voxel, leaf by leaf in bit order */
    tgx_buffer   svo_leaf_moff;  /* n_leaves + 1: a leaf's first material   */
    tgx_buffer   svo_leaf_wpre;  /* 1024 u16 per leaf: solid voxels in earlier words */
    uint64_t     svo_n_solid = 0;/* solid voxels over all leaves            */
    tgx_buffer   svo_leaf_n;     /* clusters per leaf that sampled a voxel  */
    tgx_buffer   svo_list_off;   /* n_leaves + 1                            */
    tgx_buffer   svo_list;       /* cluster idx lists                       */
    tgx_buffer   svo_frames;     /* per-cptr cluster transforms (build)     */
    tgx_buffer   svo_inner;      /* uint2 per inner node (traversal layout) */
    tgx_buffer   svo_bricks;     /* u64 per leaf: 8^3 bricks holding solids */
    uint32_t     svo_n_inner = 0;/* inner nodes = index of the first leaf   */
    tgx_buffer   svo_rays;       /* staging for tgx_svo_trace               */
    tgx_buffer   svo_hits;
    tgx_buffer   prim_tmp;       /* rocPRIM scratch                         */
    tgx_buffer   gi_counts;      /* GI queue counters and trace cursors     */
    tgx_buffer   gi_surf;        /* per-pixel surface records               */
    tgx_buffer   gi_list;        /* visible pixel list, flags, scan         */
    tgx_buffer   gi_bsurf;       /* multi-camera frames: surfaces per camera */
    tgx_buffer   gi_blist;       /* ... pixel lists, flags, scans per camera */
    tgx_buffer   gi_bcount;      /* ... visible count per camera             */
    tgx_buffer   gi_rays;        /* bounce ray queue                        */
    tgx_buffer   gi_res;         /* bounce results                          */
    tgx_buffer   gi_shadow;      /* shadow ray queue                        */
    tgx_buffer   gi_shadow_res;  /* shadow results                          */
    tgx_buffer   gi_weight;      /* RGB weight per shadow ray               */
    tgx_buffer   gi_contrib;     /* RGB contribution per (pixel, sample)    */
    tgx_buffer   gi_rays2;       /* sorted bounce rays                      */
    tgx_buffer   gi_sort;        /* sort keys and indices                   */
    uint32_t     gi_spp = 0;     /* spp of the last GI call                 */
    uint32_t     rg_lds_set = 0; /* regrouped trace variants whose LDS limit is raised */
    tgx_tuning   tuning = {0u, 5u, 2u, 0u, 0u, 0u, 2048u, 1u, 16u, 40u, 3u, 8u, 32u, 8u};

    /* Neural irradiance cache (tgx_nrc.hip). */
    bool           nrc_ready = false;
    tgx_nrc_config nrc_cfg = {};
    bool           nrc_fb_traced = false; /* last frame traced a fallback list */
    uint32_t       nrc_qc_off = 0;        /* u32 offset of the last camera's counters in nrc_queue */
    tgx_buffer     nrc_master, nrc_m, nrc_v;   /* fp32 [N_PARAMS]           */
    tgx_buffer     nrc_fwd, nrc_bwd; /* bf16 MFMA A-operand fragments       */
    tgx_buffer     nrc_ema;          /* fp32 moving average of the weights  */
    tgx_buffer     nrc_efwd;         /* its bf16 forward fragments          */
    tgx_buffer     nrc_partials;     /* fp32 [chunks][N_PARAMS]             */
    tgx_buffer     nrc_grad;         /* fp32 [N_PARAMS], last step          */
    tgx_buffer     nrc_train;        /* bf16 [feature][batch] activations   */
    tgx_buffer     nrc_loss;         /* fp32 per training sample            */
    tgx_buffer     nrc_samples, nrc_targets, nrc_out; /* staging            */
    tgx_buffer     nrc_cells;        /* error-estimate table                */
    tgx_buffer     nrc_queue;        /* frame lists and counters            */
    tgx_buffer     nrc_dev;          /* device frame and Adam-step counters */
    tgx_nrc_stats  nrc_stats = {};

    /* Captured frames alive: while > 0, buffers are neither grown nor freed
     * (the graphs hold their addresses). */
    uint32_t     n_graphs = 0;
    tgx_buffer   frame_acc;      /* totals of the last tgx_frame_batch     */
    tgx_buffer   pick_buf;       /* tgx_pick coordinates and results       */

    /* Dynamic scene (tgx_dynamic.hip): free stacks, host copy of the
     * object records, staging and compaction scratch. */
    bool                   dynamic = false;
    bool                   svo_stale = false; /* scene edited since the last SVO build */
    std::vector<tgx_object> dyn_objects;
    std::vector<uint32_t>  dyn_free_objects, dyn_free_clusters;
    tgx_buffer             dyn_stage, dyn_tmp;
    tgx_buffer             cl_dirty;          /* u8 per cluster: placed since the last SVO build */

    /* What tgx_svo_update compares against: the last build's leaf keys,
     * per-leaf candidate offsets and candidate cluster indices, and (during
     * an update) the last build's leaf data. */
    bool         svo_prev_valid = false;
    uint32_t     svo_prev_leaves = 0;
    uint32_t     svo_reused = 0;
    tgx_buffer   svo_prev_keys, svo_prev_off, svo_prev_cand;
    tgx_buffer   svo_old_bits, svo_old_mats, svo_old_moff, svo_old_wpre, svo_old_n, svo_old_list_off, svo_old_list;
    /* Temporary arrays of a build: one arena, plus spills freed at its end. */
    tgx_buffer          svo_arena;
    size_t              svo_arena_used = 0, svo_arena_want = 0, svo_arena_spill_bytes = 0;
    std::vector<void*>  svo_arena_spill;

    /* Profiling. */
    bool                           profiling = false;
    std::vector<tgx_kernel_stat>   stats;
    std::vector<tgx_pending_event> pending;
};

int  tgx_fail(tgx_context* ctx, int code, const char* fmt, ...);
int  tgx_prim_exclusive_scan_u32(tgx_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n);
int  tgx_prim_sort_pairs_u32(tgx_context* ctx, const uint32_t* keys_in, uint32_t* keys_out,
                             const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, unsigned end_bit);
void tgx_svo_release(tgx_context* ctx);
void tgx_nrc_release(tgx_context* ctx);
/* Visibility camera records for captured frames: tgx_vis_cams_upload
 * allocates and fills them (free with hipFree); tgx_visibility_record is
 * tgx_visibility_device without host work (no camera upload, no list
 * capacity update, no statistics readback), for stream capture after one
 * ordinary call at the same size. */
int  tgx_vis_cams_upload(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                         void** dev_cams);
int  tgx_visibility_record(tgx_context* ctx, const void* dev_cams, const tgx_camera* cams, uint32_t n_cams,
                           uint32_t w, uint32_t h, uint64_t* out_dev);
/* The cache frame after visibility (vis_dev already holds the camera's words). */
extern "C" int tgx_nrc_frame_from_vis(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const tgx_gi_params* gi,
                            const uint64_t* vis_dev, float* radiance_dev, float* irradiance_dev);
/* The cache frames of n_cams cameras after their visibility (camera k's
 * words at vis_dev + k w h, outputs likewise): the same results as
 * tgx_nrc_frame_from_vis camera after camera, with every camera's training
 * targets traced in one launch. acc (may be null): frame totals. */
extern "C" int tgx_nrc_frames_from_vis(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w,
                                       uint32_t h, const tgx_gi_params* gi, const uint64_t* vis_dev,
                                       float* radiance_dev, float* irradiance_dev, uint32_t* acc);
int  tgx_buffer_reserve(tgx_context* ctx, tgx_buffer* b, size_t bytes, const char* what);
/* Material bytes of the current scene for the kernels: null when the scene
 * was uploaded without materials (every voxel then has material 0), never
 * a buffer left over from an earlier scene. */
static inline const uint8_t* tgx_scene_materials(const tgx_context* ctx)
{
    return ctx->has_materials ? (const uint8_t*)ctx->materials.ptr : nullptr;
}
/* Called by every call that replaces the scene (upload, reserve, device
 * generation) once its arguments are checked: refused while a frame graph
 * holds the scene's buffers; otherwise the SVO of the previous scene no
 * longer describes the scene and is marked stale. */
int  tgx_scene_replace_begin(tgx_context* ctx, const char* what);
/* TGX_OK if an SVO is built and describes the current scene: tracing calls
 * refuse an SVO made stale by scene replacement or object edits
 * (TGX_ERR_NOT_BUILT until tgx_svo_update or tgx_svo_build). */
int  tgx_svo_check_current(tgx_context* ctx);
/* Brings the per-cluster solid flags of the visibility cull up to date with
 * the scene's bits (a kernel over every cluster index when they changed). */
int  tgx_vis_solid_refresh(tgx_context* ctx);
void tgx_buffer_release(tgx_buffer* b);
/* Looks up, once per device, the device address of the scratch array of
 * tgx_visibility_device_noalloc (the context's device must be current). */
int  tgx_vis_scratch_init(tgx_context* ctx);

/* Kernel launch bracket: records HIP events on the context stream around a
 * launch when profiling is enabled and `on` is true (paths that must not
 * create HIP objects pass false). */
struct tgx_launch_scope {
    tgx_context* ctx;
    int          stat_idx = -1;
    hipEvent_t   start = nullptr, stop = nullptr;
    tgx_launch_scope(tgx_context* c, const char* name, bool on = true);
    ~tgx_launch_scope();
};

#define TGX_HIP_CHECK(ctx, call)                                                          \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess)                                                             \
            return tgx_fail((ctx), TGX_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                            hipGetErrorString(_e), __FILE__, __LINE__);                   \
    } while (0)

#define TGX_LAUNCH_CHECK(ctx, name)                                                       \
    do {                                                                                  \
        hipError_t _e = hipGetLastError();                                                \
        if (_e != hipSuccess)                                                             \
            return tgx_fail((ctx), TGX_ERR_HIP, "launch of %s failed: %s", (name),        \
                            hipGetErrorString(_e));                                       \
    } while (0)

#endif
