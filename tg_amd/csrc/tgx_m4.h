/*
 * tgx_m4.h — 4x4 matrix and 3-vector helpers shared by host C and HIP
 * device code of the product path.
 *
 * Storage is column-major, m[4*col + row], as tg's m4 (tg_math.h:240-275).
 * Every function fixes its floating-point operation order, so that host
 * and device builds (both compiled with -ffp-contract=off, IEEE fp32
 * division and square root) produce the same bits:
 *   mul:     r(i,j) = ((a(i,0)b(0,j) + a(i,1)b(1,j)) + a(i,2)b(2,j)) + a(i,3)b(3,j)
 *            (tg_math.c:2374-2399)
 *   mulv4:   r.i = ((x m(i,0) + y m(i,1)) + z m(i,2)) + w m(i,3)
 *            (tg_math.c:2428-2438)
 *   inverse: cofactor expansion of tg_math.c:2185-2234, one fp32 division.
 */
#ifndef TGX_M4_H
#define TGX_M4_H

#if defined(__HIPCC__)
#define TGX_HD __host__ __device__ static inline
#else
#define TGX_HD static inline
#endif

typedef struct tgx_m4 { float m[16]; } tgx_m4;
typedef struct tgx_v3 { float x, y, z; } tgx_v3;
typedef struct tgx_v4 { float x, y, z, w; } tgx_v4;

#define TGX_M(mat, r, c) ((mat).m[4 * (c) + (r)])

TGX_HD tgx_v3 tgx_v3_make(float x, float y, float z) { tgx_v3 v; v.x = x; v.y = y; v.z = z; return v; }

TGX_HD tgx_m4 tgx_m4_identity(void)
{
    tgx_m4 r;
    for (int i = 0; i < 16; i++) r.m[i] = 0.0f;
    r.m[0] = 1.0f; r.m[5] = 1.0f; r.m[10] = 1.0f; r.m[15] = 1.0f;
    return r;
}

TGX_HD tgx_m4 tgx_m4_translate(float x, float y, float z)
{
    tgx_m4 r = tgx_m4_identity();
    TGX_M(r, 0, 3) = x; TGX_M(r, 1, 3) = y; TGX_M(r, 2, 3) = z;
    return r;
}

TGX_HD tgx_m4 tgx_m4_mul(tgx_m4 a, tgx_m4 b)
{
    tgx_m4 r;
    for (int c = 0; c < 4; c++) {
        for (int i = 0; i < 4; i++) {
            float s = TGX_M(a, i, 0) * TGX_M(b, 0, c);
            s = s + TGX_M(a, i, 1) * TGX_M(b, 1, c);
            s = s + TGX_M(a, i, 2) * TGX_M(b, 2, c);
            s = s + TGX_M(a, i, 3) * TGX_M(b, 3, c);
            TGX_M(r, i, c) = s;
        }
    }
    return r;
}

TGX_HD tgx_v4 tgx_m4_mulv4(tgx_m4 m, float x, float y, float z, float w)
{
    tgx_v4 r;
    r.x = x * TGX_M(m, 0, 0) + y * TGX_M(m, 0, 1) + z * TGX_M(m, 0, 2) + w * TGX_M(m, 0, 3);
    r.y = x * TGX_M(m, 1, 0) + y * TGX_M(m, 1, 1) + z * TGX_M(m, 1, 2) + w * TGX_M(m, 1, 3);
    r.z = x * TGX_M(m, 2, 0) + y * TGX_M(m, 2, 1) + z * TGX_M(m, 2, 2) + w * TGX_M(m, 2, 3);
    r.w = x * TGX_M(m, 3, 0) + y * TGX_M(m, 3, 1) + z * TGX_M(m, 3, 2) + w * TGX_M(m, 3, 3);
    return r;
}

/* Cofactor inverse with the reference's sub-determinant names and order. */
TGX_HD tgx_m4 tgx_m4_inverse(tgx_m4 a)
{
#define A(r, c) TGX_M(a, r, c)
    const float m2323 = A(2,2) * A(3,3) - A(2,3) * A(3,2);
    const float m1323 = A(2,1) * A(3,3) - A(2,3) * A(3,1);
    const float m1223 = A(2,1) * A(3,2) - A(2,2) * A(3,1);
    const float m0323 = A(2,0) * A(3,3) - A(2,3) * A(3,0);
    const float m0223 = A(2,0) * A(3,2) - A(2,2) * A(3,0);
    const float m0123 = A(2,0) * A(3,1) - A(2,1) * A(3,0);
    const float m2313 = A(1,2) * A(3,3) - A(1,3) * A(3,2);
    const float m1313 = A(1,1) * A(3,3) - A(1,3) * A(3,1);
    const float m1213 = A(1,1) * A(3,2) - A(1,2) * A(3,1);
    const float m2312 = A(1,2) * A(2,3) - A(1,3) * A(2,2);
    const float m1312 = A(1,1) * A(2,3) - A(1,3) * A(2,1);
    const float m1212 = A(1,1) * A(2,2) - A(1,2) * A(2,1);
    const float m0313 = A(1,0) * A(3,3) - A(1,3) * A(3,0);
    const float m0213 = A(1,0) * A(3,2) - A(1,2) * A(3,0);
    const float m0312 = A(1,0) * A(2,3) - A(1,3) * A(2,0);
    const float m0212 = A(1,0) * A(2,2) - A(1,2) * A(2,0);
    const float m0113 = A(1,0) * A(3,1) - A(1,1) * A(3,0);
    const float m0112 = A(1,0) * A(2,1) - A(1,1) * A(2,0);

    const float det = 1.0f / (
        A(0,0) * (A(1,1) * m2323 - A(1,2) * m1323 + A(1,3) * m1223) -
        A(0,1) * (A(1,0) * m2323 - A(1,2) * m0323 + A(1,3) * m0223) +
        A(0,2) * (A(1,0) * m1323 - A(1,1) * m0323 + A(1,3) * m0123) -
        A(0,3) * (A(1,0) * m1223 - A(1,1) * m0223 + A(1,2) * m0123));

    tgx_m4 r;
    TGX_M(r,0,0) = det *  (A(1,1) * m2323 - A(1,2) * m1323 + A(1,3) * m1223);
    TGX_M(r,0,1) = det * -(A(0,1) * m2323 - A(0,2) * m1323 + A(0,3) * m1223);
    TGX_M(r,0,2) = det *  (A(0,1) * m2313 - A(0,2) * m1313 + A(0,3) * m1213);
    TGX_M(r,0,3) = det * -(A(0,1) * m2312 - A(0,2) * m1312 + A(0,3) * m1212);
    TGX_M(r,1,0) = det * -(A(1,0) * m2323 - A(1,2) * m0323 + A(1,3) * m0223);
    TGX_M(r,1,1) = det *  (A(0,0) * m2323 - A(0,2) * m0323 + A(0,3) * m0223);
    TGX_M(r,1,2) = det * -(A(0,0) * m2313 - A(0,2) * m0313 + A(0,3) * m0213);
    TGX_M(r,1,3) = det *  (A(0,0) * m2312 - A(0,2) * m0312 + A(0,3) * m0212);
    TGX_M(r,2,0) = det *  (A(1,0) * m1323 - A(1,1) * m0323 + A(1,3) * m0123);
    TGX_M(r,2,1) = det * -(A(0,0) * m1323 - A(0,1) * m0323 + A(0,3) * m0123);
    TGX_M(r,2,2) = det *  (A(0,0) * m1313 - A(0,1) * m0313 + A(0,3) * m0113);
    TGX_M(r,2,3) = det * -(A(0,0) * m1312 - A(0,1) * m0312 + A(0,3) * m0112);
    TGX_M(r,3,0) = det * -(A(1,0) * m1223 - A(1,1) * m0223 + A(1,2) * m0123);
    TGX_M(r,3,1) = det *  (A(0,0) * m1223 - A(0,1) * m0223 + A(0,2) * m0123);
    TGX_M(r,3,2) = det * -(A(0,0) * m1213 - A(0,1) * m0213 + A(0,2) * m0113);
    TGX_M(r,3,3) = det *  (A(0,0) * m1212 - A(0,1) * m0212 + A(0,2) * m0112);
#undef A
    return r;
}

/* World -> cluster space of cluster (rx, ry, rz) of an object with
 * n clusters per axis: T(-offset) * T(+half) * inverse(R) * T(-t),
 * associated left to right (visibility.frag:57-62,
 * tg_sparse_voxel_octree.c:53-57). The cluster then spans [0,8]^3. */
TGX_HD tgx_m4 tgx_world_to_cluster(const float rotation[16], const float translation[3],
                                   const unsigned n[3], unsigned rx, unsigned ry, unsigned rz)
{
    tgx_m4 rot;
    for (int i = 0; i < 16; i++) rot.m[i] = rotation[i];
    const float offx = (float)(rx * 8u), offy = (float)(ry * 8u), offz = (float)(rz * 8u);
    const float hx = 4.0f * (float)n[0], hy = 4.0f * (float)n[1], hz = 4.0f * (float)n[2];
    const tgx_m4 m0 = tgx_m4_translate(-translation[0], -translation[1], -translation[2]);
    const tgx_m4 m1 = tgx_m4_inverse(rot);
    const tgx_m4 m2 = tgx_m4_translate(hx, hy, hz);
    const tgx_m4 m3 = tgx_m4_translate(-offx, -offy, -offz);
    return tgx_m4_mul(tgx_m4_mul(tgx_m4_mul(m3, m2), m1), m0);
}

/* Cluster -> world space: T(t) * R * T(-half) * T(+offset), associated
 * left to right (tg_sparse_voxel_octree.c:92-97). */
TGX_HD tgx_m4 tgx_cluster_to_world(const float rotation[16], const float translation[3],
                                   const unsigned n[3], unsigned rx, unsigned ry, unsigned rz)
{
    tgx_m4 rot;
    for (int i = 0; i < 16; i++) rot.m[i] = rotation[i];
    const float offx = (float)(rx * 8u), offy = (float)(ry * 8u), offz = (float)(rz * 8u);
    const float hx = (float)n[0] * 4.0f, hy = (float)n[1] * 4.0f, hz = (float)n[2] * 4.0f;
    const tgx_m4 m1 = tgx_m4_translate(offx, offy, offz);
    const tgx_m4 m2 = tgx_m4_translate(-hx, -hy, -hz);
    const tgx_m4 m4_ = tgx_m4_translate(translation[0], translation[1], translation[2]);
    return tgx_m4_mul(tgx_m4_mul(tgx_m4_mul(m4_, rot), m2), m1);
}

#endif /* TGX_M4_H */
