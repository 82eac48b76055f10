/*
 * tgx_noise.h — value noise of the synthetic scenes, shared by the host
 * generator (tgx_gen.c) and the device generator (tgx_gen_dev.hip).
 * Integer lattice hashing (splitmix64) and IEEE double arithmetic only, with
 * a fixed operation order, so both builds produce identical voxels.
 */
#ifndef TGX_NOISE_H
#define TGX_NOISE_H

#include <math.h>
#include <stdint.h>

#include "tgx_m4.h"

/* Threshold below which the solidity noise leaves ~40% of voxels solid
 * (measured over 2^22 samples; the fraction of a generated scene is checked
 * by the tests). */
#define TGX_NOISE_SOLID_THRESHOLD 0.4493

TGX_HD uint64_t tgx_splitmix64_next(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform double in [0, 1) at an integer lattice point. */
TGX_HD double tgx_noise_lattice(uint64_t key, int64_t x, int64_t y, int64_t z)
{
    uint64_t h = key;
    h ^= (uint64_t)x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)y * 0xC2B2AE3D27D4EB4Full;
    h ^= (uint64_t)z * 0x165667B19E3779F9ull;
    (void)tgx_splitmix64_next(&h);
    return (double)(tgx_splitmix64_next(&h) >> 11) * (1.0 / 9007199254740992.0);
}

TGX_HD double tgx_noise_smooth(double t) { return t * t * (3.0 - 2.0 * t); }

/* Trilinear interpolation of lattice values with smoothstep weights. */
TGX_HD double tgx_value_noise(uint64_t key, double x, double y, double z, double period)
{
    const double px = x / period, py = y / period, pz = z / period;
    const double fx0 = floor(px), fy0 = floor(py), fz0 = floor(pz);
    const int64_t ix = (int64_t)fx0, iy = (int64_t)fy0, iz = (int64_t)fz0;
    const double tx = tgx_noise_smooth(px - fx0), ty = tgx_noise_smooth(py - fy0), tz = tgx_noise_smooth(pz - fz0);
    const double c000 = tgx_noise_lattice(key, ix, iy, iz), c001 = tgx_noise_lattice(key, ix + 1, iy, iz);
    const double c010 = tgx_noise_lattice(key, ix, iy + 1, iz), c011 = tgx_noise_lattice(key, ix + 1, iy + 1, iz);
    const double c100 = tgx_noise_lattice(key, ix, iy, iz + 1), c101 = tgx_noise_lattice(key, ix + 1, iy, iz + 1);
    const double c110 = tgx_noise_lattice(key, ix, iy + 1, iz + 1), c111 = tgx_noise_lattice(key, ix + 1, iy + 1, iz + 1);
    const double x00 = c000 + (c001 - c000) * tx;
    const double x10 = c010 + (c011 - c010) * tx;
    const double x01 = c100 + (c101 - c100) * tx;
    const double x11 = c110 + (c111 - c110) * tx;
    const double y0 = x00 + (x10 - x00) * ty;
    const double y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

/* Voxel v (64 z + 8 y + x) of cluster (cx, cy, cz) of an object whose noise
 * field is offset by (ox, oy, oz): solid flag and material index. */
TGX_HD void tgx_noise_voxel(uint64_t key_solid, uint64_t key_mat, double ox, double oy, double oz, uint32_t cx,
                            uint32_t cy, uint32_t cz, uint32_t v, int* solid, uint8_t* material)
{
    const double x = ox + (double)(8u * cx + (v & 7u)) + 0.5;
    const double y = oy + (double)(8u * cy + ((v >> 3) & 7u)) + 0.5;
    const double z = oz + (double)(8u * cz + (v >> 6)) + 0.5;
    *solid = tgx_value_noise(key_solid, x, y, z, 32.0) < TGX_NOISE_SOLID_THRESHOLD;
    const double m = tgx_value_noise(key_mat, x, y, z, 16.0) * 256.0;
    *material = (uint8_t)(m >= 255.0 ? 255.0 : m);
}

#endif
