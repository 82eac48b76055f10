/*
 * tgx_gen.h — deterministic synthetic scenes and cameras (host C).
 *
 * These generators produce the benchmark and test workloads: tg itself
 * fills clusters with simplex-noise terrain on the host
 * (tgvk_raytracer.c:868-943) and builds its camera from Euler angles
 * (tgvk_core.c:382-444). The scenes here follow the same data model
 * (tgx.h) with seeded content so that the GPU path and the CPU oracle
 * consume identical bytes. No GPU is needed to call anything in this file.
 *
 * Random streams: splitmix64, one independent stream per purpose
 * (extents, transforms, solidity, materials, LUTs, camera), each seeded
 * with tgx_gen_stream_seed(seed, purpose).
 */
#ifndef TGX_GEN_H
#define TGX_GEN_H

#include "tgx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tgx_gen_kind {
    /* One object of 4x4x4 clusters, identity rotation, translation 0.
     * Solid ~ Bernoulli(0.3), materials ~ U{0..255}, one LUT. */
    TGX_GEN_SINGLE_BERNOULLI = 0,
    /* n_objects objects with per-axis extents U{1..max_extent} clusters,
     * uniform unit-quaternion rotations, translations U[0, world)^3,
     * solidity = value noise (period 32 voxels) below a threshold that
     * leaves about 40% of voxels solid, materials = quantised value noise
     * (period 16), n_luts LUTs, lut_idx ~ U{0..n_luts-1}. */
    TGX_GEN_NOISE_OBJECTS = 1
} tgx_gen_kind;

typedef struct tgx_gen_params {
    uint32_t kind;         /* tgx_gen_kind                                  */
    uint64_t seed;
    uint32_t n_objects;    /* NOISE_OBJECTS only                            */
    uint32_t max_extent;   /* clusters per axis, NOISE_OBJECTS only         */
    float    world;        /* translation range, NOISE_OBJECTS only         */
    uint32_t n_luts;       /* >= 1                                          */
} tgx_gen_params;

/* Sizes of the arrays a scene needs. */
typedef struct tgx_gen_sizes {
    uint32_t n_objects;
    uint32_t n_cluster_pointers;
    uint32_t n_clusters;
    uint32_t n_luts;
} tgx_gen_sizes;

/* Pass 1: computes the array sizes of the scene. */
int tgx_gen_scene_sizes(const tgx_gen_params* params, tgx_gen_sizes* out);

/* Pass 2: fills caller-allocated arrays sized by tgx_gen_scene_sizes:
 * objects[n_objects], cluster_pointers[n_cluster_pointers],
 * cluster_idx_to_object[n_clusters], bits[16*n_clusters],
 * materials[512*n_clusters], luts[256*n_luts]. Cluster indices are
 * allocated in pointer order, as a freshly created tg scene has them
 * (tgvk_raytracer.c:707-710, 851-866). */
int tgx_gen_scene_fill(const tgx_gen_params* params, tgx_object* objects,
                       uint32_t* cluster_pointers, uint32_t* cluster_idx_to_object,
                       uint32_t* bits, uint8_t* materials, tgx_lut_entry* luts);

/* Builds a camera record the way tg does (tgvk_core.c:382-444):
 * rotation = inverse(euler(pitch, yaw, roll)), projection =
 * perspective(fov_y, aspect, near, far), corner rays =
 * normalize((inverse(P * R) * (+-1, +-1, 1, 1)).xyz). */
int tgx_camera_make(const float position[3], float pitch, float yaw, float roll,
                    float fov_y, float aspect, float near_plane, float far_plane,
                    tgx_camera* out);

/* World-space forward direction of a camera with the given angles. */
void tgx_camera_forward(float pitch, float yaw, float out_forward[3]);

/* Seeded camera of a workload:
 *   SINGLE_BERNOULLI: 60 deg FOV, square aspect, 64 voxels from the object
 *                     centre looking at it, yaw U[0,2pi), pitch U[-0.6,0.6];
 *   NOISE_OBJECTS:    60 deg FOV, aspect w/h, at (world/2)^3 with yaw
 *                     U[0,2pi), pitch U[-0.5,0.5].
 * camera_index selects one of several seeded cameras (0 = the default).
 * near = 0.1, far = 1000 (tg_application.c:58-59). */
int tgx_gen_camera(const tgx_gen_params* params, uint32_t camera_index,
                   uint32_t w, uint32_t h, tgx_camera* out);

/* Builds the scene of `params` directly in the context's device buffers,
 * replacing any previous scene, as if tgx_gen_scene_fill's arrays had been
 * passed to tgx_scene_upload (identical bytes). NOISE_OBJECTS scenes are
 * voxelised on the GPU (the per-voxel noise of the 9.4 GB material array
 * of a 4096-object scene would take minutes on the host); the
 * single-object scene is generated on the host and uploaded. */
int tgx_gen_scene_device(tgx_context* ctx, const tgx_gen_params* params);

/* splitmix64 stream seed for (seed, purpose). */
uint64_t tgx_gen_stream_seed(uint64_t seed, uint64_t purpose);

#ifdef __cplusplus
}
#endif

#endif /* TGX_GEN_H */
