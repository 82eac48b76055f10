/*
 * tgx_gi_dev.h — GI pieces shared by the plain estimator (tgx_gi.hip) and
 * the neural irradiance cache (tgx_nrc.hip): the per-pixel surface record,
 * the counter layout, workgroup-aggregated queue appends and the host
 * passes that trace GI for a list of pixels.
 */
#ifndef TGX_GI_DEV_H
#define TGX_GI_DEV_H

#include "tgx_svo_dev.h"

/* Surface of a primary hit (40 B): world position, unit normal, albedo. */
struct gi_surf_rec {
    float    p[3];
    float    n[3];
    float    albedo[3];
    uint32_t valid;
};
static_assert(sizeof(gi_surf_rec) == 40, "gi_surf_rec layout");

/* A block of GI_N_COUNTERS u32 counters belongs to one traced list:
 * 1 bounce rays, 2 bounce hits, 3 shadow rays, 4 and 5 trace cursors
 * (bounce, shadow), 7 shadow hits. */
#define GI_N_COUNTERS  8
/* ctx->gi_counts layout */
#define GI_CNT_PLAIN   0u   /* tgx_gi_device and the cache's fallback pixels */
#define GI_CNT_TRAIN   8u   /* the cache's training targets */
#define GI_CNT_SVO     16u  /* tgx_svo_trace */
#define GI_CNT_VISIBLE 24u  /* visible pixels of the last surface pass */
#define GI_CNT_QUEUE   32u  /* shared ray queues of tgx_gi_trace_sets with two sets */
#define GI_CNT_SEG     40u  /* the trace launch's queue-segment cursors, 16 u32 apart */
#define GI_CNT_TOTAL   168u

/* Reserves `count` consecutive queue slots for the calling lane (counts may
 * differ per lane) with one atomic per workgroup; returns the lane's first
 * slot. Every thread of the workgroup must call it. */
template <uint32_t kBlock>
__device__ static inline uint32_t block_append(uint32_t* counter, uint32_t count)
{
    __shared__ uint32_t s_wave[kBlock / 64u + 1u];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = count;
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    if (lane == 63u) s_wave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < kBlock / 64u; w++) {
            const uint32_t c = s_wave[w];
            s_wave[w] = total;
            total += c;
        }
        s_wave[kBlock / 64u] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    return s_wave[kBlock / 64u] + s_wave[wave] + incl - count;
}

/* MurmurHash3 finaliser (util.inc:47-56) */
__device__ static inline uint32_t gi_fmix32(uint32_t v)
{
    v ^= v >> 16; v *= 0x85ebca6bu; v ^= v >> 13; v *= 0xc2b2ae35u; v ^= v >> 16;
    return v;
}

/* Checks resolution and estimator parameters for a GI call. */
int tgx_gi_validate(tgx_context* ctx, uint32_t w, uint32_t h, const tgx_gi_params* prm);

/* Surface pass: ctx->gi_surf[w h] (gi_surf_rec), ctx->gi_list = visible
 * pixels in ascending order, their count at gi_counts[GI_CNT_VISIBLE].
 * Outputs of pixels without a hit are set to 0 (either may be null). */
int tgx_gi_surfaces(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis_dev,
                    float* radiance_dev, float* irradiance_dev);

/* The surface pass into caller buffers: surf[w h], list_scratch[3 w h]
 * (list, flags, scan), *count = visible pixels. */
int tgx_gi_surfaces_at(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const uint64_t* vis_dev,
                       float* radiance_dev, float* irradiance_dev, gi_surf_rec* surf, uint32_t* list_scratch,
                       uint32_t* count);

/* Traces the estimator for the pixels list[0 .. *count_dev) (max_count
 * bounds the count) using the surfaces of the last surface pass. Per list
 * entry i, the irradiance E goes to e_entry[3 i] and the variance of that
 * mean (summed over RGB, 0 for spp 1) to e_var[i]; per pixel, albedo * E
 * goes to radiance_dev and E to irradiance_dev (each may be null).
 * counters is a block of GI_N_COUNTERS in gi_counts, cleared here. With
 * frame_salt (device u32, may be null) the bounce samples of frame
 * prm->frame + 0x632BE5AB (*frame_salt + 1) are drawn, so replayed work can
 * draw fresh samples. */
int tgx_gi_trace_list(tgx_context* ctx, const tgx_gi_params* prm, const uint32_t* list, const uint32_t* count_dev,
                      uint32_t max_count, uint32_t* counters, float* e_entry, float* e_var, float* radiance_dev,
                      float* irradiance_dev, const uint32_t* frame_salt);

/* One list of tgx_gi_trace_sets: the arguments of tgx_gi_trace_list. */
struct gi_trace_set {
    tgx_gi_params   prm;
    const uint32_t* list;
    const uint32_t* count_dev;
    uint32_t        max_count;
    uint32_t*       counters;
    float*          e_entry;
    float*          e_var;
    float*          radiance_dev;
    float*          irradiance_dev;
    const uint32_t* frame_salt;
    const gi_surf_rec* surf;        /* the list's surface records (null: the last surface pass's) */
    uint32_t        salt_add;       /* added to *frame_salt (work drawn ahead of the counter's ticks) */
};

/* Traces several lists through one shared ray queue and one trace launch,
 * so a small list does not pay a launch tail of its own. Each list's
 * results equal those of tgx_gi_trace_list on it alone (every sample is
 * drawn from its own list entry, sample and frame, and contributions are
 * summed per entry in sample order). The lists share prm except spp and
 * frame. Counters hold bounce rays (1), bounce hits (2) and shadow rays (3):
 * with two sets per list, the queues then using the block at GI_CNT_QUEUE;
 * with one set, or more than two, all in sets[0].counters (totals). */
int tgx_gi_trace_sets(tgx_context* ctx, const gi_trace_set* sets, uint32_t n_sets);

#endif
