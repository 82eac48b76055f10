/*
 * tgx_svo_dev.h — SVO node layout and small device helpers shared by the
 * SVO build (tgx_svo.hip) and the traversal / GI kernels (tgx_gi.hip).
 */
#ifndef TGX_SVO_DEV_H
#define TGX_SVO_DEV_H

#include "tgx_internal.h"

#define SVO_F32_MAX 3.402823466e+38f
#define SVO_EPSILON 1.192092896e-07f
#define SVO_NONE    0xFFFFFFFFu

/* Node layout on the device (32 B): canonical order, parent links. */
struct svo_gnode {
    uint32_t first_child;
    uint32_t parent;
    uint32_t leaf_idx;
    uint32_t masks;       /* valid | leaf << 8 | level << 16 */
    uint32_t cx, cy, cz;
    uint32_t leaf_n;
};
static_assert(sizeof(svo_gnode) == 32, "svo_gnode layout");

struct f3 { float x, y, z; };

__device__ static inline float fmin2(float a, float b) { return a < b ? a : b; }
__device__ static inline float fmax2(float a, float b) { return a > b ? a : b; }

static inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

#endif
