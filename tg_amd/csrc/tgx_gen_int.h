/*
 * tgx_gen_int.h — generator pieces shared by the host generator
 * (tgx_gen.c) and the device generator (tgx_gen_dev.hip).
 */
#ifndef TGX_GEN_INT_H
#define TGX_GEN_INT_H

#include "../../include/tgx_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Object records of a NOISE_OBJECTS scene (clusters numbered object by
 * object, pointer == index) and each object's noise-field offset
 * (3 doubles per object), exactly as tgx_gen_scene_fill draws them. */
int tgx_gen_noise_objects(const tgx_gen_params* p, tgx_object* objects, double* noise_offsets);

/* The LUT array of a scene (256 entries per LUT). */
void tgx_gen_luts(const tgx_gen_params* p, tgx_lut_entry* luts);

/* Key of the solidity (material = 0) or material (material = 1) noise. */
uint64_t tgx_gen_noise_key(uint64_t seed, int material);

#ifdef __cplusplus
}
#endif

#endif
