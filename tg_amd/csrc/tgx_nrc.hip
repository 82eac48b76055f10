/*
 * tgx_nrc.hip — neural irradiance cache on gfx950: BF16 MFMA training and
 * inference with FP32 accumulation, FP32 Adam. No counterpart in the
 * reference; the network and features are described at tgx_nrc_config in
 * tgx.h.
 *
 * MFMA layout (v_mfma_f32_16x16x32_bf16, one wave): A is 16 x 32 with lane
 * l holding row l % 16, k = 8 (l / 16) + j; B is 32 x 16 with lane l
 * holding column l % 16, the same k; the accumulator holds rows
 * 4 (l / 16) + i of column l % 16. Layers are evaluated transposed,
 * Z^T [out x batch] = W^T [out x in] A^T [in x batch], 16 samples per tile,
 * so a lane holds hidden features 16 t + 4 g + i (tile t, g = l / 16) of
 * sample l % 16. The next layer's B operand for K step s takes tiles 2s and
 * 2s + 1 unchanged: its k slot 8 g + j is feature 16 (2s + j / 4) + 4 g +
 * j % 4, and the weights are packed into A fragments with that K order, so
 * no data moves between layers. The backward pass, dA_L = W_L dZ_(L+1),
 * uses the same layouts with W_L packed as the A operand. Weight gradients
 * dW_L = A_L^T dZ_(L+1) take K = batch: activations and gradients are
 * stored [feature][batch] so that each A / B fragment is one 16-byte load.
 *
 * Training step (samples padded to a multiple of 512):
 *   nrc_fwdbwd_k  forward, loss gradient and backward of 16 samples per
 *                 wave; stores A_0..A_4 and dZ_1..dZ_5 as bf16
 *   nrc_wgrad_k   one 16 x 16 weight-gradient tile over 512 samples per
 *                 wave (76 tiles x batch / 512 chunks)
 *   nrc_adam_k    per parameter: sum of the chunk partials in chunk order,
 *                 Adam, fp32 master weights and the bf16 fragments
 * Every reduction has a fixed order, so training is deterministic.
 */
#include "tgx_gi_dev.h"

#include <cmath>
#include <cstdio>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define NRC_N          TGX_NRC_N_PARAMS
#define NRC_FWD_ELEMS  19456u   /* bf16 A fragments of the forward pass */
#define NRC_BWD_ELEMS  14336u   /* bf16 A fragments of the backward pass */
#define NRC_CHUNK      512u     /* samples per weight-gradient partial */
#define NRC_WTILES     76u      /* 16 x 16 weight-gradient tiles */
#define NRC_ROWS       624u     /* [feature] rows of the training scratch */
#define NRC_CELLS      (1u << 18)
#define NRC_ADAM_BLOCKS ((NRC_N + 255u) / 256u)   /* workgroups of nrc_adam_k */
#define NRC_FIX        1048576.0f /* fixed point of the error estimate, 2^20 */
#define NRC_BLOCK      256u

/* parameter offsets (floats), row-major W[in][out] */
#define P_W0 0u
#define P_W1 6144u
#define P_W2 10240u
#define P_W3 14336u
#define P_W4 18432u
/* forward fragment bases (bf16 elements): layer 0 4 tiles x 3 K steps,
 * layers 1..3 4 x 2, layer 4 1 x 2 */
#define F_L0 0u
#define F_L1 6144u
#define F_L4 18432u
/* backward fragment bases: layers 1..3 4 x 2, layer 4 4 x 1 */
#define B_L1 0u
#define B_L4 12288u
/* rows of the training scratch */
#define R_ACT0 0u      /* 96 rows: A_0 (inputs) */
#define R_ACT1 96u     /* A_1..A_4: 64 rows each */
#define R_DZ1  352u    /* dZ_1..dZ_4: 64 rows each */
#define R_DZ5  608u    /* dZ_5 = dL/dy: 3 rows (16 reserved) */

struct nrc_cell {
    int32_t  err;  /* sum of (y - t)^2 - var(t), fixed point */
    uint32_t ref;  /* sum of t^2, fixed point */
    uint32_t n;    /* samples * 256 */
    uint32_t var;  /* sum of var(t), fixed point */
};

/* A training batch drawn from several cameras: `len` consecutive samples
 * per camera, camera k's surfaces, list, count and targets at k times the
 * strides (in records, entries, words and samples). One camera: len = the
 * batch, strides 0. */
struct nrc_segs {
    uint32_t len;
    uint32_t surf;
    uint32_t list;
    uint32_t count;
    uint32_t tgt;
};

struct nrc_enc {
    float ox, oy, oz, scale;
};

__device__ static inline f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* hidden feature held in k slot 8 g + j of K step s */
__device__ static inline uint32_t kfeat_hidden(uint32_t s, uint32_t g, uint32_t j)
{
    return 16u * (2u * s + j / 4u) + 4u * g + (j & 3u);
}

/* ------------------------------------------------------------------------ */
/* Features and the per-tile passes                                         */
/* ------------------------------------------------------------------------ */

__device__ static inline float feat_tail(const gi_surf_rec& r, uint32_t f)
{
    return f == 72u ? r.n[0] : f == 73u ? r.n[1] : f == 74u ? r.n[2] : f == 75u ? r.albedo[0]
         : f == 76u ? r.albedo[1] : f == 77u ? r.albedo[2] : f == 78u ? 1.0f : 0.0f;
}

/* B fragment of the input features for K step s (features 32 s + 8 g + j).
 * sin(2^k pi x) = sin(2 pi frac(2^(k-1) x)); v_sin / v_cos take turns. */
__device__ static inline bf16x8 input_frag(const gi_surf_rec& r, float px, float py, float pz, uint32_t s, uint32_t g)
{
    bf16x8 v;
#pragma unroll
    for (uint32_t jj = 0; jj < 8u; jj += 2u) {
        const uint32_t f = 32u * s + 8u * g + jj;
        float a, b;
        if (f < 72u) {
            const uint32_t ax = f / 24u, k = (f % 24u) / 2u;
            const float ph = ax == 0u ? px : (ax == 1u ? py : pz);
            const float x = ldexpf(ph, (int)k - 1);
            const float t = x - floorf(x);
            a = __builtin_amdgcn_sinf(t);
            b = __builtin_amdgcn_cosf(t);
        } else {
            a = feat_tail(r, f);
            b = feat_tail(r, f + 1u);
        }
        v[jj] = (__bf16)a;
        v[jj + 1] = (__bf16)b;
    }
    return v;
}

__device__ static inline void encode(const gi_surf_rec& r, const nrc_enc& e, uint32_t g, bf16x8 x[3])
{
    const float px = (r.p[0] - e.ox) * e.scale, py = (r.p[1] - e.oy) * e.scale, pz = (r.p[2] - e.oz) * e.scale;
#pragma unroll
    for (uint32_t s = 0; s < 3u; s++) x[s] = input_frag(r, px, py, pz, s, g);
}

__device__ static inline void relu_pack(const f32x4 acc[4], bf16x8 out[2])
{
#pragma unroll
    for (uint32_t s = 0; s < 2u; s++)
#pragma unroll
        for (uint32_t j = 0; j < 8u; j++) {
            const float v = acc[2u * s + j / 4u][j & 3u];
            out[s][j] = (__bf16)(v > 0.0f ? v : 0.0f);
        }
}

/* Forward pass of one 16-sample tile with the forward fragments in
 * registers (fragment f of the forward array is wf[64 f + lane]). h[L] =
 * activations after hidden layer L; y in lanes 0..15. */
__device__ static inline void forward_tile(const bf16x8 (&w)[NRC_FWD_ELEMS / 512u], const bf16x8 x[3],
                                           bf16x8 h[4][2], f32x4& y)
{
    f32x4 acc[4];
#pragma unroll
    for (uint32_t t = 0; t < 4u; t++) {
        acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t s = 0; s < 3u; s++) acc[t] = mfma16(w[t * 3u + s], x[s], acc[t]);
    }
    relu_pack(acc, h[0]);
#pragma unroll
    for (uint32_t L = 1; L < 4u; L++) {
#pragma unroll
        for (uint32_t t = 0; t < 4u; t++) {
            acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (uint32_t s = 0; s < 2u; s++) acc[t] = mfma16(w[12u + (L - 1u) * 8u + t * 2u + s], h[L - 1u][s], acc[t]);
        }
        relu_pack(acc, h[L]);
    }
    y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t s = 0; s < 2u; s++) y = mfma16(w[36u + s], h[3][s], y);
}

/* The 38 forward A fragments of one lane: fragment f of the forward
 * fragment array is wf[64 f + lane] (layer 0: 4 tiles x 3 K steps, layers
 * 1..3: 4 x 2 each, layer 4: 2). */
#define NRC_FWD_FRAGS (NRC_FWD_ELEMS / 512u)

/* Forward pass for inference with the fragments in registers: one
 * activation buffer, y in lanes 0..15. */
__device__ static inline f32x4 infer_tile(const bf16x8 (&w)[NRC_FWD_FRAGS], const bf16x8 x[3])
{
    f32x4 acc[4];
    bf16x8 h[2];
#pragma unroll
    for (uint32_t t = 0; t < 4u; t++) {
        acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (uint32_t s = 0; s < 3u; s++) acc[t] = mfma16(w[t * 3u + s], x[s], acc[t]);
    }
    relu_pack(acc, h);
#pragma unroll
    for (uint32_t L = 1; L < 4u; L++) {
#pragma unroll
        for (uint32_t t = 0; t < 4u; t++) {
            acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (uint32_t s = 0; s < 2u; s++) acc[t] = mfma16(w[12u + (L - 1u) * 8u + t * 2u + s], h[s], acc[t]);
        }
        relu_pack(acc, h);
    }
    f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t s = 0; s < 2u; s++) y = mfma16(w[36u + s], h[s], y);
    return y;
}

/* dZ = dA masked by the layer input's activation (same lane layout). */
__device__ static inline void mask_pack(const f32x4 acc[4], const bf16x8 act[2], bf16x8 out[2])
{
#pragma unroll
    for (uint32_t s = 0; s < 2u; s++)
#pragma unroll
        for (uint32_t j = 0; j < 8u; j++)
            out[s][j] = (float)act[s][j] > 0.0f ? (__bf16)acc[2u * s + j / 4u][j & 3u] : (__bf16)0.0f;
}

/* Backward pass with the backward fragments in registers (fragment f of
 * the backward array is wb[64 f + lane]: layers 1..3 4 x 2 each, layer 4
 * 4 x 1): dz[L - 1] = dZ_L fragments for L = 1..4 from dZ_5. */
#define NRC_BWD_FRAGS (NRC_BWD_ELEMS / 512u)
__device__ static inline void backward_tile(const bf16x8 (&w)[NRC_BWD_FRAGS], const bf16x8 h[4][2], bf16x8 dz5,
                                            bf16x8 dz[4][2])
{
    f32x4 acc[4];
#pragma unroll
    for (uint32_t t = 0; t < 4u; t++) acc[t] = mfma16(w[24u + t], dz5, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    mask_pack(acc, h[3], dz[3]);
#pragma unroll
    for (uint32_t L = 3; L >= 1u; L--) {
#pragma unroll
        for (uint32_t t = 0; t < 4u; t++) {
            acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (uint32_t s = 0; s < 2u; s++) acc[t] = mfma16(w[(L - 1u) * 8u + t * 2u + s], dz[L][s], acc[t]);
        }
        mask_pack(acc, h[L - 1u], dz[L - 1u]);
    }
}

/* Error-estimate cell of a surface: position quantised by 2^cell_shift
 * voxels and the dominant normal direction. */
__device__ static inline uint32_t cell_of(const gi_surf_rec& r, uint32_t cell_shift)
{
    const int x = (int)floorf(r.p[0]) >> cell_shift, y = (int)floorf(r.p[1]) >> cell_shift,
              z = (int)floorf(r.p[2]) >> cell_shift;
    const float ax = fabsf(r.n[0]), ay = fabsf(r.n[1]), az = fabsf(r.n[2]);
    const uint32_t a = ax >= ay && ax >= az ? (r.n[0] > 0.0f ? 0u : 1u)
                     : ay >= az ? (r.n[1] > 0.0f ? 2u : 3u) : (r.n[2] > 0.0f ? 4u : 5u);
    const uint32_t h = gi_fmix32((uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u ^
                                 a * 2654435761u);
    return h & (NRC_CELLS - 1u);
}

__device__ static inline uint32_t to_fix(float v)
{
    const float c = fminf(fmaxf(v, 0.0f), 16.0f);
    return (uint32_t)(c * NRC_FIX);
}

/* ------------------------------------------------------------------------ */
/* Kernels                                                                  */
/* ------------------------------------------------------------------------ */

/* Forward + backward of one 16-sample tile per wave. Sample b uses
 * surf[list ? list[b] : b] and targets[3 b]; samples b >= *count get a
 * zero loss gradient. inv_n = 1 / (3 n). */
__global__ __launch_bounds__(NRC_BLOCK) __attribute__((amdgpu_waves_per_eu(1, 2))) void nrc_fwdbwd_k(
                                                          const bf16x8* __restrict__ wf, const bf16x8* __restrict__ wb,
                                                          const gi_surf_rec* __restrict__ surf,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ count_ptr, uint32_t first,
                                                          const float* __restrict__ targets,
                                                          const float* __restrict__ tvar, nrc_enc enc, float two_inv_n,
                                                          uint32_t Bc, __bf16* __restrict__ train,
                                                          float* __restrict__ loss, nrc_cell* __restrict__ cells,
                                                          uint32_t cell_shift, nrc_segs sg)
{
    const uint32_t lane = threadIdx.x & 63u, g = lane >> 4, n = lane & 15u;
    const uint32_t tile = (blockIdx.x * NRC_BLOCK + threadIdx.x) >> 6;
    const uint32_t b = tile * 16u + n;          /* sample within the batch */
    if (tile * 16u >= Bc) return;               /* whole wave */
    /* every fragment is loaded up front, so the layers wait on no load */
    bf16x8 wfr[NRC_FWD_FRAGS], wbr[NRC_BWD_FRAGS];
#pragma unroll
    for (uint32_t f = 0; f < NRC_FWD_FRAGS; f++) wfr[f] = wf[64u * f + lane];
#pragma unroll
    for (uint32_t f = 0; f < NRC_BWD_FRAGS; f++) wbr[f] = wb[64u * f + lane];
    /* the batch is sg.len samples from each of Bc / sg.len cameras */
    const uint32_t k = b / sg.len, bl = b - k * sg.len;
    surf += (size_t)k * sg.surf;
    if (list) list += (size_t)k * sg.list;
    targets += 3ull * k * sg.tgt;
    if (tvar) tvar += (size_t)k * sg.tgt;
    const uint32_t count = count_ptr[(size_t)k * sg.count];
    const bool valid = first + bl < count;
    gi_surf_rec r = {};
    if (valid) r = surf[list ? list[first + bl] : first + bl];
    bf16x8 x[3], h[4][2];
    encode(r, enc, g, x);
    f32x4 y;
    forward_tile(wfr, x, h, y);
    /* loss gradient in lanes g == 0 (features 0..2 of the output tile) */
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
    if (valid && g == 0u) { t0 = targets[3ull * (first + bl)]; t1 = targets[3ull * (first + bl) + 1]; t2 = targets[3ull * (first + bl) + 2]; }
    const float d0 = y[0] - t0, d1 = y[1] - t1, d2 = y[2] - t2;
    bf16x8 dz5;
#pragma unroll
    for (uint32_t j = 0; j < 8u; j++) dz5[j] = (__bf16)0.0f;
    if (valid && g == 0u) {
        dz5[0] = (__bf16)(two_inv_n * d0);
        dz5[1] = (__bf16)(two_inv_n * d1);
        dz5[2] = (__bf16)(two_inv_n * d2);
        const float se = d0 * d0 + d1 * d1 + d2 * d2;
        loss[b] = se;
        if (cells) {
            const uint32_t c = cell_of(r, cell_shift);
            const float tv = tvar ? tvar[first + bl] : 0.0f;
            const float bias2 = se - tv;
            const float cl = fminf(fmaxf(bias2, -16.0f), 16.0f);
            atomicAdd(&cells[c].err, (int32_t)(cl * NRC_FIX));
            atomicAdd(&cells[c].ref, to_fix(t0 * t0 + t1 * t1 + t2 * t2));
            atomicAdd(&cells[c].n, 256u);
            atomicAdd(&cells[c].var, to_fix(tv));
        }
    } else if (g == 0u) {
        loss[b] = 0.0f;
    }
    bf16x8 dz[4][2];
    backward_tile(wbr, h, dz5, dz);
    /* store activations and gradients [feature][batch] */
#pragma unroll
    for (uint32_t s = 0; s < 3u; s++)
#pragma unroll
        for (uint32_t j = 0; j < 8u; j++) train[(size_t)(R_ACT0 + 32u * s + 8u * g + j) * Bc + b] = x[s][j];
#pragma unroll
    for (uint32_t L = 0; L < 4u; L++)
#pragma unroll
        for (uint32_t s = 0; s < 2u; s++)
#pragma unroll
            for (uint32_t j = 0; j < 8u; j++) {
                const uint32_t f = kfeat_hidden(s, g, j);
                train[(size_t)(R_ACT1 + 64u * L + f) * Bc + b] = h[L][s][j];
                train[(size_t)(R_DZ1 + 64u * L + f) * Bc + b] = dz[L][s][j];
            }
    if (g == 0u)
#pragma unroll
        for (uint32_t c = 0; c < 3u; c++) train[(size_t)(R_DZ5 + c) * Bc + b] = dz5[c];
}

/* One 16 x 16 tile of dW_L = A_L^T dZ_(L+1) over `per` consecutive
 * chunks of NRC_CHUNK samples per wave (in sample order); partial[group]
 * [param] for the n_chunks groups. */
__global__ __launch_bounds__(NRC_BLOCK) void nrc_wgrad_k(const __bf16* __restrict__ train, uint32_t Bc,
                                                         uint32_t n_chunks, uint32_t per, float* __restrict__ partial)
{
    const uint32_t lane = threadIdx.x & 63u, g = lane >> 4, n = lane & 15u;
    const uint32_t wave = (blockIdx.x * NRC_BLOCK + threadIdx.x) >> 6;
    if (wave >= NRC_WTILES * n_chunks) return;
    const uint32_t tile = wave % NRC_WTILES, chunk = wave / NRC_WTILES;
    uint32_t L, ti, to;
    if (tile < 24u) { L = 0; ti = tile / 4u; to = tile % 4u; }
    else if (tile < 72u) { L = 1u + (tile - 24u) / 16u; ti = ((tile - 24u) % 16u) / 4u; to = (tile - 24u) % 4u; }
    else { L = 4; ti = tile - 72u; to = 0; }
    const uint32_t arow = (L == 0u ? R_ACT0 : R_ACT1 + 64u * (L - 1u)) + 16u * ti + n;
    const uint32_t brow = (L == 4u ? R_DZ5 : R_DZ1 + 64u * L) + 16u * to + n;
    const bool bvalid = L != 4u || n < 3u;
    const __bf16* ap = train + (size_t)arow * Bc + (size_t)chunk * per * NRC_CHUNK + 8u * g;
    const __bf16* bp = train + (size_t)brow * Bc + (size_t)chunk * per * NRC_CHUNK + 8u * g;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 zero;
#pragma unroll
    for (uint32_t j = 0; j < 8u; j++) zero[j] = (__bf16)0.0f;
    for (uint32_t k = 0; k < per * NRC_CHUNK; k += 32u) {
        const bf16x8 a = *(const bf16x8*)(ap + k);
        const bf16x8 bq = bvalid ? *(const bf16x8*)(bp + k) : zero;
        acc = mfma16(a, bq, acc);
    }
    const uint32_t out_dim = L == 4u ? 3u : 64u;
    const uint32_t off = L == 0u ? P_W0 : (L == 4u ? P_W4 : P_W1 + 4096u * (L - 1u));
    const uint32_t out = 16u * to + n;
    if (out >= out_dim) return;
#pragma unroll
    for (uint32_t i = 0; i < 4u; i++) {
        const uint32_t in = 16u * ti + 4u * g + i;
        partial[(size_t)chunk * NRC_N + off + in * out_dim + out] = acc[i];
    }
}

/* Fragment positions of parameter i: forward (always) and backward
 * (layers 1..4; -1 for layer 0). */
__device__ static inline void param_slots(uint32_t i, int& fwd, int& bwd)
{
    uint32_t L, in, out, off, od;
    if (i < P_W1) { L = 0; off = P_W0; od = 64u; }
    else if (i < P_W4) { L = 1u + (i - P_W1) / 4096u; off = P_W1 + 4096u * (L - 1u); od = 64u; }
    else { L = 4; off = P_W4; od = 3u; }
    in = (i - off) / od;
    out = (i - off) % od;
    /* forward: A[m = out % 16 of tile out / 16][k slot of in] */
    uint32_t s, g, j;
    if (L == 0u) { s = in / 32u; g = (in % 32u) / 8u; j = in % 8u; }
    else { s = in / 32u; const uint32_t r = in % 32u; g = (r % 16u) / 4u; j = 4u * (r / 16u) + (r % 4u); }
    const uint32_t t = out / 16u, m = out % 16u;
    if (L == 0u) fwd = (int)(F_L0 + (t * 3u + s) * 512u + (g * 16u + m) * 8u + j);
    else if (L < 4u) fwd = (int)(F_L1 + (L - 1u) * 4096u + (t * 2u + s) * 512u + (g * 16u + m) * 8u + j);
    else fwd = (int)(F_L4 + s * 512u + (g * 16u + m) * 8u + j);
    /* backward: A[m = in % 16 of tile in / 16][k slot of out] */
    if (L == 0u) { bwd = -1; return; }
    const uint32_t tb = in / 16u, mb = in % 16u;
    if (L == 4u) { bwd = (int)(B_L4 + tb * 512u + mb * 8u + out); return; }
    const uint32_t sb = out / 32u, rb = out % 32u, gb = (rb % 16u) / 4u, jb = 4u * (rb / 16u) + (rb % 4u);
    bwd = (int)(B_L1 + (L - 1u) * 4096u + (tb * 2u + sb) * 512u + (gb * 16u + mb) * 8u + jb);
}

__global__ void nrc_pack_k(const float* __restrict__ master, __bf16* __restrict__ fwd, __bf16* __restrict__ bwd)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NRC_N) return;
    int f, b;
    param_slots(i, f, b);
    const __bf16 v = (__bf16)master[i];
    fwd[f] = v;
    if (b >= 0 && bwd) bwd[b] = v;
}

struct nrc_adam_args {
    float lr, beta1, beta2, eps;
};

/* device counters (nrc_dev): 0 frames run, 1 Adam steps taken */
#define NRC_DEV_FRAME 0
#define NRC_DEV_STEP  1
#define NRC_DEV_DONE  2   /* blocks of the running Adam step that are done */

__global__ void nrc_tick_k(uint32_t* counter) { *counter += 1u; }

/* One Adam step with step number t = the device step counter + 1; with ema,
 * the moving average e = ema_d e + (1 - ema_d) w of the new weights and its
 * forward fragments follow. The last
 * block to finish advances the counter (replayed graphs keep counting) and,
 * with loss_out, adds the batch's per-sample squared errors to *loss_out in
 * a fixed order. dev[NRC_DEV_DONE] counts finished blocks; it is 0 between
 * launches. */
__global__ __launch_bounds__(256) void nrc_adam_k(const float* __restrict__ partial, uint32_t n_chunks, nrc_adam_args a,
                                                  uint32_t* __restrict__ dev, float* __restrict__ master,
                                                  float* __restrict__ m, float* __restrict__ v, float* __restrict__ grad,
                                                  __bf16* __restrict__ fwd, __bf16* __restrict__ bwd,
                                                  const float* __restrict__ loss, uint32_t n_loss,
                                                  float* __restrict__ loss_part, float* __restrict__ loss_out,
                                                  float ema_d, float* __restrict__ ema, __bf16* __restrict__ efwd)
{
    __shared__ float s[256];
    __shared__ uint32_t last;
    __shared__ float bc[2];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t_u = dev[NRC_DEV_STEP] + 1u;
    if (threadIdx.x == 0) {
        /* bias corrections, once per block */
        const double t = (double)t_u;
        bc[0] = (float)(1.0 / (1.0 - pow((double)a.beta1, t)));
        bc[1] = (float)(1.0 / (1.0 - pow((double)a.beta2, t)));
    }
    __syncthreads();
    if (i < NRC_N) {
        const float inv_bc1 = bc[0], inv_bc2 = bc[1];
        float gsum = 0.0f;
        for (uint32_t c = 0; c < n_chunks; c++) gsum += partial[(size_t)c * NRC_N + i];
        grad[i] = gsum;
        const float mi = a.beta1 * m[i] + (1.0f - a.beta1) * gsum;
        const float vi = a.beta2 * v[i] + (1.0f - a.beta2) * gsum * gsum;
        m[i] = mi;
        v[i] = vi;
        const float w = master[i] - a.lr * (mi * inv_bc1) / (sqrtf(vi * inv_bc2) + a.eps);
        master[i] = w;
        int f, b;
        param_slots(i, f, b);
        const __bf16 h = (__bf16)w;
        fwd[f] = h;
        if (b >= 0) bwd[b] = h;
        if (ema) {
            /* the inference weights: moving average of the trained ones */
            const float e = ema_d * ema[i] + (1.0f - ema_d) * w;
            ema[i] = e;
            efwd[f] = (__bf16)e;
        }
    }
    /* the batch's loss: every block sums a strided slice in a fixed order,
     * the last block adds the block sums in block order */
    if (loss_out) {
        float acc = 0.0f;
        for (uint32_t k = i; k < n_loss; k += gridDim.x * blockDim.x) acc += loss[k];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (uint32_t w = 128u; w > 0u; w >>= 1) {
            if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) loss_part[blockIdx.x] = s[0];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&dev[NRC_DEV_DONE], 1u) == gridDim.x - 1u ? 1u : 0u;
    }
    __syncthreads();
    if (!last) return;
    if (loss_out) {
        /* the block sums (gridDim.x <= 256) in a fixed tree */
        __threadfence();
        s[threadIdx.x] = threadIdx.x < gridDim.x ? ((volatile const float*)loss_part)[threadIdx.x] : 0.0f;
        __syncthreads();
        for (uint32_t w = 128u; w > 0u; w >>= 1) {
            if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) *loss_out += s[0];
    }
    if (threadIdx.x == 0) {
        dev[NRC_DEV_STEP] = t_u;
        dev[NRC_DEV_DONE] = 0u;
    }
}

/* Inference over list entries b < *count (persistent waves, 16 samples a
 * tile, each wave's forward fragments in its registers, two waves per SIMD;
 * the next tile's surface record is loaded while the current one is
 * evaluated). With cells, an entry whose cell is
 * unreliable gets fallback flag 1 instead of an output: fewer than min_n
 * samples, or a bias^2 sum above w_ref times the targets' t^2 sum plus
 * w_var times their variance sum. */
__global__ __launch_bounds__(NRC_BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2))) void nrc_infer_k(
    const bf16x8* __restrict__ wf_g, const gi_surf_rec* __restrict__ surf, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ count_ptr, nrc_enc enc, float* __restrict__ out_entry, float* __restrict__ radiance,
    float* __restrict__ irradiance, const nrc_cell* __restrict__ cells, uint32_t cell_shift, float w_ref, float w_var,
    uint32_t min_n, uint32_t* __restrict__ fb_flag)
{
    const uint32_t lane = threadIdx.x & 63u, g = lane >> 4, n = lane & 15u;
    bf16x8 w[NRC_FWD_FRAGS];
#pragma unroll
    for (uint32_t f = 0; f < NRC_FWD_FRAGS; f++) w[f] = wf_g[64u * f + lane];
    const uint32_t count = *count_ptr;
    const uint32_t n_tiles = (count + 15u) / 16u;
    const uint32_t waves = gridDim.x * (NRC_BLOCK / 64u);
    uint32_t tile = (blockIdx.x * NRC_BLOCK + threadIdx.x) >> 6;
    gi_surf_rec rn = {};
    {
        const uint32_t b = tile * 16u + n;
        if (tile < n_tiles && b < count) rn = surf[list ? list[b] : b];
    }
    for (; tile < n_tiles; tile += waves) {
        const uint32_t b = tile * 16u + n;
        const bool valid = b < count;
        const gi_surf_rec r = rn;
        {
            const uint32_t t2 = tile + waves, b2 = t2 * 16u + n;
            rn = gi_surf_rec{};
            if (t2 < n_tiles && b2 < count) rn = surf[list ? list[b2] : b2];
        }
        bf16x8 x[3];
        encode(r, enc, g, x);
        const f32x4 y = infer_tile(w, x);
        if (!valid || g != 0u) continue;
        const float e0 = fmaxf(y[0], 0.0f), e1 = fmaxf(y[1], 0.0f), e2 = fmaxf(y[2], 0.0f);
        if (out_entry) { out_entry[3ull * b] = e0; out_entry[3ull * b + 1] = e1; out_entry[3ull * b + 2] = e2; }
        bool fb = false;
        if (cells) {
            const nrc_cell c = cells[cell_of(r, cell_shift)];
            fb = c.n < min_n || (float)c.err > w_ref * (float)c.ref + w_var * (float)c.var;
            fb_flag[b] = fb ? 1u : 0u;
        }
        if (fb) continue;
        const uint32_t pix = list ? list[b] : b;
        if (irradiance) { irradiance[3ull * pix] = e0; irradiance[3ull * pix + 1] = e1; irradiance[3ull * pix + 2] = e2; }
        if (radiance) {
            radiance[3ull * pix] = r.albedo[0] * e0;
            radiance[3ull * pix + 1] = r.albedo[1] * e1;
            radiance[3ull * pix + 2] = r.albedo[2] * e2;
        }
    }
}

/* Training pixels: entry j takes visible pixel hash(j, frame) % n_visible. */
__global__ void nrc_pick_k(const uint32_t* __restrict__ vis_list, const uint32_t* __restrict__ n_vis_ptr,
                           uint32_t n_samples, const uint32_t* __restrict__ dev, uint32_t* __restrict__ out,
                           uint32_t* __restrict__ count, uint32_t frame_add)
{
    const uint32_t frame = dev[NRC_DEV_FRAME] + frame_add;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nv = *n_vis_ptr;
    if (j == 0) *count = nv ? n_samples : 0u;
    if (j >= n_samples || nv == 0u) return;
    const uint32_t h = gi_fmix32(j * 0x9E3779B9u ^ gi_fmix32(frame + 0x632BE5ABu));
    out[j] = vis_list[h % nv];
}

/* Scales every cell of the error estimate by 7/8 (older frames fade out
 * over about eight frames: halving each frame left too few samples per cell
 * and its noise sent many pixels to the fallback needlessly). */
__global__ void nrc_decay_k(nrc_cell* __restrict__ cells)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NRC_CELLS) return;
    nrc_cell c = cells[i];
    c.err -= c.err >> 3;
    c.ref -= c.ref >> 3;
    c.n -= c.n >> 3;
    c.var -= c.var >> 3;
    cells[i] = c;
}

/* Fallback list from the flags and their exclusive scan. */
__global__ void nrc_fallback_list_k(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off,
                                    const uint32_t* __restrict__ vis_list, const uint32_t* __restrict__ count_ptr,
                                    uint32_t* __restrict__ out, uint32_t* __restrict__ out_count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *count_ptr;
    if (i == 0 && n == 0) *out_count = 0;
    if (i >= n) return;
    if (flag[i]) out[off[i]] = vis_list[i];
    if (i == n - 1) *out_count = off[i] + flag[i];
}

__global__ void nrc_fill_k(uint32_t* p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* ------------------------------------------------------------------------ */
/* Host side                                                                */
/* ------------------------------------------------------------------------ */

static inline uint32_t nrc_blocks(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

void tgx_nrc_release(tgx_context* ctx)
{
    tgx_buffer* bufs[] = {&ctx->nrc_master, &ctx->nrc_m, &ctx->nrc_v, &ctx->nrc_fwd, &ctx->nrc_bwd, &ctx->nrc_ema,
                          &ctx->nrc_efwd,
                          &ctx->nrc_partials, &ctx->nrc_grad, &ctx->nrc_train, &ctx->nrc_loss,
                          &ctx->nrc_samples, &ctx->nrc_targets, &ctx->nrc_out, &ctx->nrc_cells, &ctx->nrc_queue,
                          &ctx->nrc_dev};
    for (tgx_buffer* b : bufs) tgx_buffer_release(b);
    ctx->nrc_ready = false;
}

namespace {

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* He-uniform initial weights (rows of W0 beyond the 79 used features 0). */
void init_params(uint32_t seed, std::vector<float>& p)
{
    p.assign(NRC_N, 0.0f);
    uint64_t st = 0x243F6A8885A308D3ull ^ seed;
    const uint32_t offs[6] = {P_W0, P_W1, P_W2, P_W3, P_W4, NRC_N};
    for (int L = 0; L < 5; L++) {
        const uint32_t od = L == 4 ? 3u : 64u;
        const double fan_in = L == 0 ? 79.0 : 64.0;
        const double lim = std::sqrt(6.0 / fan_in) * (L == 4 ? 0.5 : 1.0);
        for (uint32_t i = offs[L]; i < offs[L + 1]; i++) {
            const uint32_t in = (i - offs[L]) / od;
            const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
            p[i] = (L == 0 && in >= 79u) ? 0.0f : (float)((2.0 * u - 1.0) * lim);
        }
    }
}

nrc_enc enc_of(const tgx_nrc_config& c)
{
    nrc_enc e = {c.origin[0], c.origin[1], c.origin[2], c.scale};
    return e;
}

int check_ready(tgx_context* ctx)
{
    if (!ctx->nrc_ready) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no cache: call tgx_nrc_create first");
    return TGX_OK;
}

int pack(tgx_context* ctx)
{
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_fwd.ptr, 0, 2ull * NRC_FWD_ELEMS, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_bwd.ptr, 0, 2ull * NRC_BWD_ELEMS, ctx->stream));
    hipLaunchKernelGGL(nrc_pack_k, dim3(nrc_blocks(NRC_N, 256)), dim3(256), 0, ctx->stream,
                       (const float*)ctx->nrc_master.ptr, (__bf16*)ctx->nrc_fwd.ptr, (__bf16*)ctx->nrc_bwd.ptr);
    TGX_LAUNCH_CHECK(ctx, "nrc_pack_k");
    /* the moving average starts from the same weights */
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->nrc_ema.ptr, ctx->nrc_master.ptr, 4ull * NRC_N, hipMemcpyDeviceToDevice,
                                      ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_efwd.ptr, 0, 2ull * NRC_FWD_ELEMS, ctx->stream));
    hipLaunchKernelGGL(nrc_pack_k, dim3(nrc_blocks(NRC_N, 256)), dim3(256), 0, ctx->stream,
                       (const float*)ctx->nrc_ema.ptr, (__bf16*)ctx->nrc_efwd.ptr, (__bf16*)nullptr);
    TGX_LAUNCH_CHECK(ctx, "nrc_pack_k");
    return TGX_OK;
}

/* One Adam step on samples [first, first + Bc) of (surf, list) with the
 * given targets; *count_dev bounds the valid samples, n_norm sets the loss
 * normalisation (3 n_norm). loss_acc (device float) accumulates the sum of
 * squared errors. */
int train_step(tgx_context* ctx, const gi_surf_rec* surf, const uint32_t* list, const uint32_t* count_dev,
               uint32_t first, uint32_t Bc, uint32_t n_norm, const float* targets, const float* tvar,
               nrc_cell* cells, float* loss_acc, const nrc_segs* segs = nullptr)
{
    const nrc_segs sg = segs ? *segs : nrc_segs{Bc, 0u, 0u, 0u, 0u};
    const tgx_nrc_config& c = ctx->nrc_cfg;
    /* weight-gradient partials: one per group of `per` chunks, at most 16
     * groups (a large batch is summed in fewer, longer MFMA runs) */
    uint32_t per = 1;
    while (Bc / NRC_CHUNK / per > 16u && (Bc / NRC_CHUNK) % (2u * per) == 0u) per *= 2u;
    const uint32_t n_chunks = Bc / NRC_CHUNK / per;
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->nrc_train, 2ull * NRC_ROWS * Bc, "nrc training scratch")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_partials, 4ull * NRC_N * n_chunks, "nrc gradient partials")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_loss, 4ull * Bc + 4ull * NRC_ADAM_BLOCKS, "nrc losses")))
        return rc;
    uint32_t shift = 0;
    while ((1u << shift) < c.cell_size) shift++;
    {
        tgx_launch_scope s(ctx, "nrc_fwdbwd_k");
        hipLaunchKernelGGL(nrc_fwdbwd_k, dim3(nrc_blocks(Bc / 16u, NRC_BLOCK / 64u)), dim3(NRC_BLOCK), 0, ctx->stream,
                           (const bf16x8*)ctx->nrc_fwd.ptr, (const bf16x8*)ctx->nrc_bwd.ptr, surf, list, count_dev,
                           first, targets, tvar, enc_of(c), 2.0f / (3.0f * (float)n_norm), Bc,
                           (__bf16*)ctx->nrc_train.ptr, (float*)ctx->nrc_loss.ptr, cells, shift, sg);
    }
    TGX_LAUNCH_CHECK(ctx, "nrc_fwdbwd_k");
    {
        tgx_launch_scope s(ctx, "nrc_wgrad_k");
        hipLaunchKernelGGL(nrc_wgrad_k, dim3(nrc_blocks((uint64_t)NRC_WTILES * n_chunks, NRC_BLOCK / 64u)),
                           dim3(NRC_BLOCK), 0, ctx->stream, (const __bf16*)ctx->nrc_train.ptr, Bc, n_chunks, per,
                           (float*)ctx->nrc_partials.ptr);
    }
    TGX_LAUNCH_CHECK(ctx, "nrc_wgrad_k");
    nrc_adam_args a;
    a.lr = c.learning_rate; a.beta1 = c.beta1; a.beta2 = c.beta2; a.eps = c.epsilon;
    uint32_t* dev = (uint32_t*)ctx->nrc_dev.ptr;
    {
        tgx_launch_scope s(ctx, "nrc_adam_k");
        hipLaunchKernelGGL(nrc_adam_k, dim3(nrc_blocks(NRC_N, 256)), dim3(256), 0, ctx->stream,
                           (const float*)ctx->nrc_partials.ptr, n_chunks, a, dev, (float*)ctx->nrc_master.ptr,
                           (float*)ctx->nrc_m.ptr, (float*)ctx->nrc_v.ptr, (float*)ctx->nrc_grad.ptr,
                           (__bf16*)ctx->nrc_fwd.ptr, (__bf16*)ctx->nrc_bwd.ptr, (const float*)ctx->nrc_loss.ptr, Bc,
                           (float*)ctx->nrc_loss.ptr + Bc, loss_acc, c.ema_decay, c.ema_decay > 0.0f ? (float*)ctx->nrc_ema.ptr : nullptr,
                           (__bf16*)ctx->nrc_efwd.ptr);
    }
    TGX_LAUNCH_CHECK(ctx, "nrc_adam_k");
    return TGX_OK;
}

/* Host samples (9 floats) -> surface records in ctx->nrc_samples; also a
 * device copy of n at the end of the buffer. */
int upload_samples(tgx_context* ctx, const float* samples, uint32_t n, uint32_t padded, uint32_t** count_dev)
{
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->nrc_samples, sizeof(gi_surf_rec) * (size_t)padded + 256, "nrc samples")))
        return rc;
    std::vector<gi_surf_rec> h(padded);
    memset(h.data(), 0, sizeof(gi_surf_rec) * padded);
    for (uint32_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            h[i].p[k] = samples[9ull * i + k];
            h[i].n[k] = samples[9ull * i + 3 + k];
            h[i].albedo[k] = samples[9ull * i + 6 + k];
        }
        h[i].valid = 1u;
    }
    char* base = (char*)ctx->nrc_samples.ptr;
    *count_dev = (uint32_t*)(base + sizeof(gi_surf_rec) * (size_t)padded);
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(base, h.data(), sizeof(gi_surf_rec) * (size_t)padded, hipMemcpyHostToDevice, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(*count_dev, &n, 4, hipMemcpyHostToDevice, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int launch_infer(tgx_context* ctx, const gi_surf_rec* surf, const uint32_t* list, const uint32_t* count_dev,
                 uint32_t max_count, float* out_entry, float* radiance, float* irradiance, const nrc_cell* cells,
                 uint32_t* fb_flag, float w_var = -1.0f)
{
    /* w_var < 0: the t^2 rule with the configured threshold */
    if (max_count == 0) return TGX_OK;
    const tgx_nrc_config& c = ctx->nrc_cfg;
    uint32_t shift = 0;
    while ((1u << shift) < c.cell_size) shift++;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    const uint32_t want = nrc_blocks((max_count + 15u) / 16u, NRC_BLOCK / 64u);
    const uint32_t cap = (uint32_t)cus * 2u;   /* two waves per SIMD */
    {
        tgx_launch_scope s(ctx, "nrc_infer_k");
        hipLaunchKernelGGL(nrc_infer_k, dim3(want < cap ? want : cap), dim3(NRC_BLOCK), 0, ctx->stream,
                           (const bf16x8*)(c.ema_decay > 0.0f ? ctx->nrc_efwd.ptr : ctx->nrc_fwd.ptr), surf, list,
                           count_dev, enc_of(c), out_entry, radiance,
                           irradiance, cells, shift, w_var < 0.0f ? c.fallback_threshold : 0.0f,
                           w_var < 0.0f ? 0.0f : w_var, c.fallback_min_samples * 256u, fb_flag);
    }
    TGX_LAUNCH_CHECK(ctx, "nrc_infer_k");
    return TGX_OK;
}

} /* namespace */

extern "C" {

int tgx_nrc_create(tgx_context* ctx, const tgx_nrc_config* cfg)
{
    if (!ctx || !cfg) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    const tgx_nrc_config& c = *cfg;
    if (!(c.scale > 0.0f) || !(c.learning_rate > 0.0f) || !(c.beta1 >= 0.0f && c.beta1 < 1.0f) ||
        !(c.beta2 >= 0.0f && c.beta2 < 1.0f) || !(c.epsilon >= 0.0f) || !(c.ema_decay >= 0.0f && c.ema_decay < 1.0f))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "bad encoding scale, Adam parameters or ema_decay");
    if (c.train_batch == 0 || c.train_batch % TGX_NRC_BATCH_ALIGN || c.train_samples % c.train_batch ||
        c.train_samples > (1u << 24))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG,
                        "train_batch must be a positive multiple of %u dividing train_samples (<= 2^24)",
                        TGX_NRC_BATCH_ALIGN);
    if (c.train_spp < 1 || c.train_spp > 64 || c.cell_size == 0 || (c.cell_size & (c.cell_size - 1)) ||
        c.cell_size > (1u << 16) || c.fallback_min_samples > (1u << 20))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "train_spp in [1, 64], cell_size a power of two");
    if (ctx->n_graphs > 0)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "a frame graph uses the cache: destroy it before re-creating");
    tgx_nrc_release(ctx);
    int rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->nrc_master, 4ull * NRC_N, "nrc weights")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_m, 4ull * NRC_N, "nrc adam m")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_v, 4ull * NRC_N, "nrc adam v")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_grad, 4ull * NRC_N, "nrc gradient")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_fwd, 2ull * NRC_FWD_ELEMS, "nrc forward fragments")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_bwd, 2ull * NRC_BWD_ELEMS, "nrc backward fragments")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_ema, 4ull * NRC_N, "nrc weight average")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_efwd, 2ull * NRC_FWD_ELEMS, "nrc average fragments")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_cells, sizeof(nrc_cell) * (size_t)NRC_CELLS, "nrc error cells")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_dev, 64, "nrc counters")))
        return rc;
    ctx->nrc_cfg = c;
    ctx->nrc_stats = {};
    std::vector<float> p;
    init_params(c.seed, p);
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->nrc_master.ptr, p.data(), 4ull * NRC_N, hipMemcpyHostToDevice, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_m.ptr, 0, 4ull * NRC_N, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_v.ptr, 0, 4ull * NRC_N, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_grad.ptr, 0, 4ull * NRC_N, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_cells.ptr, 0, sizeof(nrc_cell) * (size_t)NRC_CELLS, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_dev.ptr, 0, 64, ctx->stream));
    if ((rc = pack(ctx))) return rc;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->nrc_ready = true;
    return TGX_OK;
}

int tgx_nrc_get_params(tgx_context* ctx, float* params)
{
    if (!ctx || !params) return TGX_ERR_INVALID_ARG;
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemcpy(params, ctx->nrc_master.ptr, 4ull * NRC_N, hipMemcpyDeviceToHost));
    return TGX_OK;
}

int tgx_nrc_get_inference_params(tgx_context* ctx, float* params)
{
    if (!ctx || !params) return TGX_ERR_INVALID_ARG;
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const tgx_buffer& src = ctx->nrc_cfg.ema_decay > 0.0f ? ctx->nrc_ema : ctx->nrc_master;
    TGX_HIP_CHECK(ctx, hipMemcpy(params, src.ptr, 4ull * NRC_N, hipMemcpyDeviceToHost));
    return TGX_OK;
}

int tgx_nrc_set_params(tgx_context* ctx, const float* params)
{
    if (!ctx || !params) return TGX_ERR_INVALID_ARG;
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemcpy(ctx->nrc_master.ptr, params, 4ull * NRC_N, hipMemcpyHostToDevice));
    if ((rc = pack(ctx))) return rc;
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_nrc_infer(tgx_context* ctx, const float* samples, uint32_t n, float* out)
{
    if (!ctx || (n && (!samples || !out))) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    if (n == 0) return TGX_OK;
    if (n > (1u << 26)) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "at most 2^26 samples per call");
    uint32_t* count_dev = nullptr;
    if ((rc = upload_samples(ctx, samples, n, n, &count_dev))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->nrc_out, 12ull * n, "nrc outputs"))) return rc;
    if ((rc = launch_infer(ctx, (const gi_surf_rec*)ctx->nrc_samples.ptr, nullptr, count_dev, n,
                           (float*)ctx->nrc_out.ptr, nullptr, nullptr, nullptr, nullptr)))
        return rc;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(out, ctx->nrc_out.ptr, 12ull * n, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

int tgx_nrc_train_batch(tgx_context* ctx, const float* samples, const float* targets, uint32_t n, float* loss,
                        float* grad)
{
    if (!ctx || !samples || !targets || n == 0) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    if (n > (1u << 22)) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "at most 2^22 samples per batch");
    const uint32_t Bc = (n + NRC_CHUNK - 1) / NRC_CHUNK * NRC_CHUNK;
    uint32_t* count_dev = nullptr;
    if ((rc = upload_samples(ctx, samples, n, Bc, &count_dev))) return rc;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->nrc_targets, 12ull * Bc + 16, "nrc targets"))) return rc;
    float* loss_acc = (float*)((char*)ctx->nrc_targets.ptr + 12ull * Bc);
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(ctx->nrc_targets.ptr, targets, 12ull * n, hipMemcpyHostToDevice, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(loss_acc, 0, 4, ctx->stream));
    if ((rc = train_step(ctx, (const gi_surf_rec*)ctx->nrc_samples.ptr, nullptr, count_dev, 0, Bc, n,
                         (const float*)ctx->nrc_targets.ptr, nullptr, nullptr, loss_acc)))
        return rc;
    float l = 0.0f;
    TGX_HIP_CHECK(ctx, hipMemcpyAsync(&l, loss_acc, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (grad) TGX_HIP_CHECK(ctx, hipMemcpyAsync(grad, ctx->nrc_grad.ptr, 4ull * NRC_N, hipMemcpyDeviceToHost, ctx->stream));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l / (3.0f * (float)n);
    return TGX_OK;
}

int tgx_nrc_frame_from_vis(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const tgx_gi_params* gi,
                           const uint64_t* vis_dev, float* radiance_dev, float* irradiance_dev)
{
    if (!ctx || !cam || !gi || !vis_dev || !radiance_dev) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    if ((rc = tgx_gi_validate(ctx, w, h, gi))) return rc;
    const tgx_nrc_config& c = ctx->nrc_cfg;
    const uint32_t n_pix = w * h;
    const uint32_t n_train = (c.flags & TGX_NRC_FLAG_NO_TRAINING) ? 0u : c.train_samples;
    if ((uint64_t)n_train * c.train_spp >= (1ull << 31))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "train_samples * train_spp too large");
    /* queue buffer: counters (64 u32), training list, fallback flags, scan, list */
    const size_t q_words = 64ull + n_train + 3ull * n_pix;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->nrc_queue, 4ull * q_words, "nrc queues")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_targets, 16ull * (n_train ? n_train : 1) + 16, "nrc targets")))
        return rc;
    uint32_t* q = (uint32_t*)ctx->nrc_queue.ptr;
    uint32_t* qc = q;                  /* 0 train count, 1 fallback count, 8 loss (float) */
    uint32_t* train_list = q + 64;
    uint32_t* fb_flag = train_list + n_train;
    uint32_t* fb_off = fb_flag + n_pix;
    uint32_t* fb_list = fb_off + n_pix;
    float* e_train = (float*)ctx->nrc_targets.ptr;
    float* v_train = e_train + 3ull * n_train;
    nrc_cell* cells = (nrc_cell*)ctx->nrc_cells.ptr;

    if ((rc = tgx_gi_surfaces(ctx, cam, w, h, vis_dev, radiance_dev, irradiance_dev))) return rc;
    uint32_t* gcnt = (uint32_t*)ctx->gi_counts.ptr;
    const uint32_t* vis_list = (const uint32_t*)ctx->gi_list.ptr;
    const uint32_t* n_vis = gcnt + GI_CNT_VISIBLE;
    const gi_surf_rec* surf = (const gi_surf_rec*)ctx->gi_surf.ptr;
    hipLaunchKernelGGL(nrc_fill_k, dim3(1), dim3(64), 0, ctx->stream, qc, 64ull, 0u);
    TGX_LAUNCH_CHECK(ctx, "nrc_fill_k");

    /* Inference with the cache as the last frame left it; pixels of cells
     * it does not trust are listed for the SVO estimator. */
    const bool all_fb = (c.flags & TGX_NRC_FLAG_ALL_FALLBACK) != 0;
    const bool use_cells = !(c.flags & TGX_NRC_FLAG_NO_FALLBACK) && !all_fb;
    if (!all_fb) {
        /* against the tracer: a fallback pixel's error is the variance of a
         * gi->spp-sample mean, train_spp / spp times a target's */
        const float w_var = (c.flags & TGX_NRC_FLAG_FALLBACK_VS_TRACER)
                                ? c.fallback_threshold * (float)c.train_spp / (float)gi->spp : -1.0f;
        if ((rc = launch_infer(ctx, surf, vis_list, n_vis, n_pix, nullptr, radiance_dev, irradiance_dev,
                               use_cells ? cells : nullptr, fb_flag, w_var)))
            return rc;
    }
    const bool fallback = use_cells || all_fb;
    if (fallback) {
        if (all_fb) {
            hipLaunchKernelGGL(nrc_fill_k, dim3(nrc_blocks(n_pix, 256)), dim3(256), 0, ctx->stream, fb_flag,
                               (uint64_t)n_pix, 1u);
            TGX_LAUNCH_CHECK(ctx, "nrc_fill_k");
        }
        /* flags past the visible count are never read: the scan covers n_pix
         * entries, the list kernel only the first *n_vis */
        if ((rc = tgx_prim_exclusive_scan_u32(ctx, fb_flag, fb_off, n_pix))) return rc;
        hipLaunchKernelGGL(nrc_fallback_list_k, dim3(nrc_blocks(n_pix, 256)), dim3(256), 0, ctx->stream, fb_flag, fb_off,
                           vis_list, n_vis, fb_list, qc + 1);
        TGX_LAUNCH_CHECK(ctx, "nrc_fallback_list_k");
    }
    /* Training targets and fallback pixels share one bounce and one shadow
     * launch; then the cache trains on this frame's targets. */
    ctx->nrc_stats.train_samples = 0;
    gi_trace_set sets[2];
    uint32_t n_sets = 0;
    if (n_train) {
        hipLaunchKernelGGL(nrc_pick_k, dim3(nrc_blocks(n_train, 256)), dim3(256), 0, ctx->stream, vis_list, n_vis,
                           n_train, (const uint32_t*)ctx->nrc_dev.ptr, train_list, qc, 0u);
        TGX_LAUNCH_CHECK(ctx, "nrc_pick_k");
        tgx_gi_params tp = *gi;
        tp.spp = c.train_spp;
        /* fresh bounce samples every frame (device frame counter), decorrelated
         * from the displayed frame's */
        tp.frame = gi->frame ^ 0x5BD1E995u;
        sets[n_sets++] = {tp, train_list, qc, n_train, gcnt + GI_CNT_TRAIN, e_train, v_train, nullptr, nullptr,
                          (const uint32_t*)ctx->nrc_dev.ptr + NRC_DEV_FRAME, nullptr, 0u};
    }
    if (fallback)
        sets[n_sets++] = {*gi, fb_list, qc + 1, n_pix, gcnt + GI_CNT_PLAIN, nullptr, nullptr, radiance_dev,
                          irradiance_dev, nullptr, nullptr, 0u};
    if (n_sets && (rc = tgx_gi_trace_sets(ctx, sets, n_sets))) return rc;
    if (n_train) {
        hipLaunchKernelGGL(nrc_decay_k, dim3(nrc_blocks(NRC_CELLS, 256)), dim3(256), 0, ctx->stream, cells);
        TGX_LAUNCH_CHECK(ctx, "nrc_decay_k");
        for (uint32_t first = 0; first < n_train; first += c.train_batch)
            if ((rc = train_step(ctx, surf, train_list, qc, first, c.train_batch, c.train_batch, e_train, v_train,
                                 cells, (float*)(qc + 8))))
                return rc;
        ctx->nrc_stats.train_samples = n_train;
    }
    hipLaunchKernelGGL(nrc_tick_k, dim3(1), dim3(1), 0, ctx->stream, (uint32_t*)ctx->nrc_dev.ptr + NRC_DEV_FRAME);
    TGX_LAUNCH_CHECK(ctx, "nrc_tick_k");
    ctx->nrc_fb_traced = fallback;
    ctx->nrc_qc_off = 0;
    return TGX_OK;
}

/* Frame totals (tgx_frame_stats order: 0 visible, 1 bounce, 2 shadow,
 * 3 hits, 4 training rays, 5 fallback pixels). With qc: the camera's
 * visible pixels, and with mode 4 its fallback count; mode 2: the training
 * block; mode 8: the fallback block. */
__global__ void nrc_acc_k(const uint32_t* __restrict__ qc, const uint32_t* __restrict__ gcnt, uint32_t mode,
                          uint32_t* __restrict__ acc)
{
    if (qc) acc[0] += gcnt[GI_CNT_VISIBLE];
    if (mode & 8u) {
        acc[1] += gcnt[GI_CNT_PLAIN + 1];
        acc[2] += gcnt[GI_CNT_PLAIN + 3];
        acc[3] += gcnt[GI_CNT_PLAIN + 2];
    }
    if (mode & 2u) acc[4] += gcnt[GI_CNT_TRAIN + 1] + gcnt[GI_CNT_TRAIN + 3];
    if ((mode & 4u) && qc) acc[5] += qc[1];
}

/* d[0] = s[0] (one word between device buffers, capture-safe) */
__global__ void nrc_copy1_k(const uint32_t* __restrict__ s, uint32_t* __restrict__ d) { *d = *s; }

int tgx_nrc_frames_from_vis(tgx_context* ctx, const tgx_camera* cams, uint32_t n_cams, uint32_t w, uint32_t h,
                            const tgx_gi_params* gi, const uint64_t* vis_dev, float* radiance_dev,
                            float* irradiance_dev, uint32_t* acc)
{
    if (!ctx || !cams || !gi || !vis_dev || !radiance_dev || n_cams == 0) return TGX_ERR_INVALID_ARG;
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    if ((rc = tgx_gi_validate(ctx, w, h, gi))) return rc;
    const tgx_nrc_config& c = ctx->nrc_cfg;
    const uint64_t px = (uint64_t)w * h;
    const uint32_t n_train = (c.flags & TGX_NRC_FLAG_NO_TRAINING) ? 0u : c.train_samples;
    if ((uint64_t)n_train * c.train_spp * n_cams >= (1ull << 31))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "train_samples * train_spp * cameras too large");
    /* queue buffer: per camera 64 counters (0 train count, 1 fallback count,
     * 8 loss), its training list and its fallback flags, scan and list */
    const size_t q_words = (64ull + n_train + 3ull * px) * n_cams;
    if ((rc = tgx_buffer_reserve(ctx, &ctx->nrc_queue, 4ull * q_words, "nrc queues")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->nrc_targets, 16ull * n_cams * (n_train ? n_train : 1) + 16, "nrc targets")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_counts, 4ull * GI_CNT_TOTAL, "gi counters")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_bsurf, sizeof(gi_surf_rec) * n_cams * px, "gi surfaces per camera")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_blist, 12ull * n_cams * px, "gi pixel lists per camera")) ||
        (rc = tgx_buffer_reserve(ctx, &ctx->gi_bcount, 4ull * n_cams, "gi visible counts per camera")))
        return rc;
    uint32_t* q = (uint32_t*)ctx->nrc_queue.ptr;
    uint32_t* train_lists = q + 64ull * n_cams;
    uint32_t* fb_lists = train_lists + (size_t)n_train * n_cams;   /* camera k: flags, scan, list at 3 k px */
    float* e_train = (float*)ctx->nrc_targets.ptr;
    float* v_train = e_train + 3ull * n_train * n_cams;
    nrc_cell* cells = (nrc_cell*)ctx->nrc_cells.ptr;
    gi_surf_rec* surf = (gi_surf_rec*)ctx->gi_bsurf.ptr;
    uint32_t* lists = (uint32_t*)ctx->gi_blist.ptr;
    uint32_t* counts = (uint32_t*)ctx->gi_bcount.ptr;
    uint32_t* gcnt = (uint32_t*)ctx->gi_counts.ptr;
    const uint32_t* dev = (const uint32_t*)ctx->nrc_dev.ptr;
    hipLaunchKernelGGL(nrc_fill_k, dim3(nrc_blocks(64ull * n_cams, 256)), dim3(256), 0, ctx->stream, q, 64ull * n_cams,
                       0u);
    TGX_LAUNCH_CHECK(ctx, "nrc_fill_k");

    /* Every camera's surfaces, training pixels and training targets first:
     * targets need no cache state, so one trace launch covers all cameras.
     * Camera k draws what the k-th frame of a camera-by-camera sequence
     * would: the frame counter is k ticks further there. */
    std::vector<gi_trace_set> sets(n_train ? n_cams : 0u);
    for (uint32_t k = 0; k < n_cams; k++) {
        float* rad = radiance_dev + 3ull * k * px;
        float* irr = irradiance_dev ? irradiance_dev + 3ull * k * px : nullptr;
        if ((rc = tgx_gi_surfaces_at(ctx, &cams[k], w, h, vis_dev + k * px, rad, irr, surf + k * px,
                                     lists + 3ull * k * px, counts + k)))
            return rc;
        if (!n_train) continue;
        uint32_t* tl = train_lists + (size_t)n_train * k;
        hipLaunchKernelGGL(nrc_pick_k, dim3(nrc_blocks(n_train, 256)), dim3(256), 0, ctx->stream, lists + 3ull * k * px,
                           counts + k, n_train, dev, tl, q + 64ull * k, k);
        TGX_LAUNCH_CHECK(ctx, "nrc_pick_k");
        tgx_gi_params tp = *gi;
        tp.spp = c.train_spp;
        tp.frame = gi->frame ^ 0x5BD1E995u;
        sets[k] = {tp, tl, q + 64ull * k, n_train, gcnt + GI_CNT_TRAIN, e_train + 3ull * n_train * k,
                   v_train + (size_t)n_train * k, nullptr, nullptr, dev + NRC_DEV_FRAME, surf + k * px, k};
    }
    if (n_train && (rc = tgx_gi_trace_sets(ctx, sets.data(), n_cams))) return rc;
    if (acc && n_train) {
        hipLaunchKernelGGL(nrc_acc_k, dim3(1), dim3(1), 0, ctx->stream, (const uint32_t*)nullptr, (const uint32_t*)gcnt,
                           2u, acc);
        TGX_LAUNCH_CHECK(ctx, "nrc_acc_k");
    }

    /* Then camera by camera: inference with the cache as the previous
     * camera left it, the fallback list, the cache's training step. The
     * fallback pixels change no cache state, so all cameras' fallback lists
     * are traced together at the end. */
    const bool all_fb = (c.flags & TGX_NRC_FLAG_ALL_FALLBACK) != 0;
    const bool use_cells = !(c.flags & TGX_NRC_FLAG_NO_FALLBACK) && !all_fb;
    const bool fallback = use_cells || all_fb;
    const float w_var = (c.flags & TGX_NRC_FLAG_FALLBACK_VS_TRACER)
                            ? c.fallback_threshold * (float)c.train_spp / (float)gi->spp : -1.0f;
    /* joint training: every camera renders with the cache the previous frame
     * left; one training pass over all cameras' targets follows */
    const bool joint = (c.flags & TGX_NRC_FLAG_JOINT_TRAINING) && n_cams > 1;
    for (uint32_t k = 0; k < n_cams; k++) {
        float* rad = radiance_dev + 3ull * k * px;
        float* irr = irradiance_dev ? irradiance_dev + 3ull * k * px : nullptr;
        const gi_surf_rec* sk = surf + k * px;
        const uint32_t* vis_list = lists + 3ull * k * px;
        const uint32_t* n_vis = counts + k;
        uint32_t* qc = q + 64ull * k;
        uint32_t* fb_flag = fb_lists + 3ull * k * px;
        uint32_t* fb_off = fb_flag + px;
        uint32_t* fb_list = fb_off + px;
        if (!all_fb &&
            (rc = launch_infer(ctx, sk, vis_list, n_vis, (uint32_t)px, nullptr, rad, irr, use_cells ? cells : nullptr,
                               fb_flag, w_var)))
            return rc;
        if (fallback) {
            if (all_fb) {
                hipLaunchKernelGGL(nrc_fill_k, dim3(nrc_blocks(px, 256)), dim3(256), 0, ctx->stream, fb_flag, px, 1u);
                TGX_LAUNCH_CHECK(ctx, "nrc_fill_k");
            }
            if ((rc = tgx_prim_exclusive_scan_u32(ctx, fb_flag, fb_off, (uint32_t)px))) return rc;
            hipLaunchKernelGGL(nrc_fallback_list_k, dim3(nrc_blocks(px, 256)), dim3(256), 0, ctx->stream, fb_flag,
                               fb_off, vis_list, n_vis, fb_list, qc + 1);
            TGX_LAUNCH_CHECK(ctx, "nrc_fallback_list_k");
        }
        if (n_train && !joint) {
            hipLaunchKernelGGL(nrc_decay_k, dim3(nrc_blocks(NRC_CELLS, 256)), dim3(256), 0, ctx->stream, cells);
            TGX_LAUNCH_CHECK(ctx, "nrc_decay_k");
            const uint32_t* tl = train_lists + (size_t)n_train * k;
            for (uint32_t first = 0; first < n_train; first += c.train_batch)
                if ((rc = train_step(ctx, sk, tl, qc, first, c.train_batch, c.train_batch, e_train + 3ull * n_train * k,
                                     v_train + (size_t)n_train * k, cells, (float*)(qc + 8))))
                    return rc;
        }
        hipLaunchKernelGGL(nrc_tick_k, dim3(1), dim3(1), 0, ctx->stream, (uint32_t*)ctx->nrc_dev.ptr + NRC_DEV_FRAME);
        TGX_LAUNCH_CHECK(ctx, "nrc_tick_k");
        /* the camera's visible count where the statistics look for it */
        hipLaunchKernelGGL(nrc_copy1_k, dim3(1), dim3(1), 0, ctx->stream, n_vis, gcnt + GI_CNT_VISIBLE);
        TGX_LAUNCH_CHECK(ctx, "nrc_copy1_k");
        if (acc) {
            hipLaunchKernelGGL(nrc_acc_k, dim3(1), dim3(1), 0, ctx->stream, (const uint32_t*)qc, (const uint32_t*)gcnt,
                               fallback ? 4u : 0u, acc);
            TGX_LAUNCH_CHECK(ctx, "nrc_acc_k");
        }
    }
    if (n_train && joint) {
        /* one decay of the error estimate per frame, then train_samples /
         * train_batch Adam steps, each on train_batch samples of every
         * camera (camera-major within the batch) */
        hipLaunchKernelGGL(nrc_decay_k, dim3(nrc_blocks(NRC_CELLS, 256)), dim3(256), 0, ctx->stream, cells);
        TGX_LAUNCH_CHECK(ctx, "nrc_decay_k");
        const nrc_segs sg = {c.train_batch, (uint32_t)px, n_train, 64u, n_train};
        const uint32_t bj = c.train_batch * n_cams;
        for (uint32_t first = 0; first < n_train; first += c.train_batch)
            if ((rc = train_step(ctx, surf, train_lists, q, first, bj, bj, e_train, v_train, cells,
                                 (float*)(q + 64ull * (n_cams - 1u) + 8), &sg)))
                return rc;
    }
    if (fallback) {
        std::vector<gi_trace_set> fs(n_cams);
        for (uint32_t k = 0; k < n_cams; k++)
            fs[k] = {*gi, fb_lists + 3ull * k * px + 2ull * px, q + 64ull * k + 1, (uint32_t)px, gcnt + GI_CNT_PLAIN,
                     nullptr, nullptr, radiance_dev + 3ull * k * px,
                     irradiance_dev ? irradiance_dev + 3ull * k * px : nullptr, nullptr, surf + k * px, 0u};
        if ((rc = tgx_gi_trace_sets(ctx, fs.data(), n_cams))) return rc;
        if (acc) {
            hipLaunchKernelGGL(nrc_acc_k, dim3(1), dim3(1), 0, ctx->stream, (const uint32_t*)nullptr,
                               (const uint32_t*)gcnt, 8u, acc);
            TGX_LAUNCH_CHECK(ctx, "nrc_acc_k");
        }
    }
    ctx->nrc_stats.train_samples = joint ? n_train * n_cams : n_train;
    ctx->nrc_fb_traced = fallback;
    ctx->nrc_qc_off = 64u * (n_cams - 1u);
    return TGX_OK;
}

int tgx_nrc_frame_device(tgx_context* ctx, const tgx_camera* cam, uint32_t w, uint32_t h, const tgx_gi_params* gi,
                         uint64_t* vis_dev, float* radiance_dev, float* irradiance_dev)
{
    if (!ctx || !cam || !gi || !vis_dev || !radiance_dev) return TGX_ERR_INVALID_ARG;
    int rc;
    if ((rc = tgx_visibility_device(ctx, cam, 1, w, h, vis_dev))) return rc;
    return tgx_nrc_frame_from_vis(ctx, cam, w, h, gi, vis_dev, radiance_dev, irradiance_dev);
}

int tgx_nrc_last_stats(tgx_context* ctx, tgx_nrc_stats* out)
{
    if (!ctx || !out) return TGX_ERR_INVALID_ARG;
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    if (!ctx->nrc_queue.ptr || !ctx->gi_counts.ptr) return tgx_fail(ctx, TGX_ERR_NOT_BUILT, "no cache frame yet");
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t qc[64], gc[GI_CNT_TOTAL], dv[2];
    TGX_HIP_CHECK(ctx, hipMemcpy(qc, (const uint32_t*)ctx->nrc_queue.ptr + ctx->nrc_qc_off, sizeof(qc),
                                 hipMemcpyDeviceToHost));
    TGX_HIP_CHECK(ctx, hipMemcpy(dv, ctx->nrc_dev.ptr, sizeof(dv), hipMemcpyDeviceToHost));
    TGX_HIP_CHECK(ctx, hipMemcpy(gc, ctx->gi_counts.ptr, sizeof(gc), hipMemcpyDeviceToHost));
    tgx_nrc_stats s = {};
    s.visible_pixels = gc[GI_CNT_VISIBLE];
    const bool fb = ctx->nrc_fb_traced;
    s.fallback_pixels = fb ? qc[1] : 0u;
    s.train_samples = ctx->nrc_stats.train_samples;
    s.train_rays = s.train_samples ? (uint64_t)gc[GI_CNT_TRAIN + 1] + gc[GI_CNT_TRAIN + 3] : 0u;
    s.fallback_rays = fb ? (uint64_t)gc[GI_CNT_PLAIN + 1] + gc[GI_CNT_PLAIN + 3] : 0u;
    s.adam_steps = dv[NRC_DEV_STEP];
    float lsum;
    memcpy(&lsum, &qc[8], 4);
    s.train_loss = s.train_samples ? lsum / (3.0f * (float)s.train_samples) : 0.0f;
    *out = s;
    return TGX_OK;
}

} /* extern "C" */

/* ------------------------------------------------------------------------ */
/* Training-state files                                                     */
/* ------------------------------------------------------------------------ */

/* "TGXNRC01" | nrc_file_header | config | device counters (64 B) | master,
 * Adam m, Adam v, moving average (fp32 N each) | error-estimate cells |
 * u64 FNV-1a 64 of every byte before it. */
struct nrc_file_header {
    char     magic[8];
    uint32_t version;
    uint32_t n_params;
    uint32_t n_cells;
    uint32_t cell_bytes;
    uint32_t config_bytes;
    uint32_t pad;
};
static const char k_nrc_magic[8] = {'T', 'G', 'X', 'N', 'R', 'C', '0', '1'};

static uint64_t nrc_fnv(const unsigned char* p, size_t n)
{
    uint64_t h = 0xCBF29CE484222325ull;
    for (size_t i = 0; i < n; i++) { h ^= p[i]; h *= 0x100000001B3ull; }
    return h;
}

static size_t nrc_file_bytes()
{
    return sizeof(nrc_file_header) + sizeof(tgx_nrc_config) + 64u + 4u * 4ull * NRC_N +
           sizeof(nrc_cell) * (size_t)NRC_CELLS + 8u;
}

extern "C" {

int tgx_nrc_save(tgx_context* ctx, const char* path)
{
    if (!ctx || !path) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    int rc;
    if ((rc = check_ready(ctx))) return rc;
    TGX_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<unsigned char> buf(nrc_file_bytes());
    unsigned char* q = buf.data();
    nrc_file_header hd = {};
    memcpy(hd.magic, k_nrc_magic, 8);
    hd.version = 1u;
    hd.n_params = NRC_N;
    hd.n_cells = NRC_CELLS;
    hd.cell_bytes = (uint32_t)sizeof(nrc_cell);
    hd.config_bytes = (uint32_t)sizeof(tgx_nrc_config);
    memcpy(q, &hd, sizeof(hd));
    q += sizeof(hd);
    memcpy(q, &ctx->nrc_cfg, sizeof(tgx_nrc_config));
    q += sizeof(tgx_nrc_config);
    TGX_HIP_CHECK(ctx, hipMemcpy(q, ctx->nrc_dev.ptr, 64, hipMemcpyDeviceToHost));
    q += 64;
    const tgx_buffer* arrays[4] = {&ctx->nrc_master, &ctx->nrc_m, &ctx->nrc_v, &ctx->nrc_ema};
    for (const tgx_buffer* b : arrays) {
        TGX_HIP_CHECK(ctx, hipMemcpy(q, b->ptr, 4ull * NRC_N, hipMemcpyDeviceToHost));
        q += 4ull * NRC_N;
    }
    TGX_HIP_CHECK(ctx, hipMemcpy(q, ctx->nrc_cells.ptr, sizeof(nrc_cell) * (size_t)NRC_CELLS, hipMemcpyDeviceToHost));
    q += sizeof(nrc_cell) * (size_t)NRC_CELLS;
    const uint64_t h = nrc_fnv(buf.data(), (size_t)(q - buf.data()));
    memcpy(q, &h, 8);
    FILE* f = fopen(path, "wb");
    if (!f) return tgx_fail(ctx, TGX_ERR_IO, "cannot open %s for writing", path);
    const bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (fclose(f) != 0 || !ok) return tgx_fail(ctx, TGX_ERR_IO, "write of %s failed", path);
    return TGX_OK;
}

int tgx_nrc_load(tgx_context* ctx, const char* path)
{
    if (!ctx || !path) return TGX_ERR_INVALID_ARG;
    ctx->last_error.clear();
    if (ctx->n_graphs > 0)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "a frame graph uses the cache: destroy it before loading");
    FILE* f = fopen(path, "rb");
    if (!f) return tgx_fail(ctx, TGX_ERR_IO, "cannot open %s", path);
    const size_t want = nrc_file_bytes();
    std::vector<unsigned char> buf(want + 1);
    const size_t got = fread(buf.data(), 1, want + 1, f);
    fclose(f);
    if (got != want) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: %zu bytes, a cache state file has %zu", path, got, want);
    nrc_file_header hd;
    memcpy(&hd, buf.data(), sizeof(hd));
    if (memcmp(hd.magic, k_nrc_magic, 8) != 0 || hd.version != 1u)
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: not a cache state file", path);
    if (hd.n_params != NRC_N || hd.n_cells != NRC_CELLS || hd.cell_bytes != sizeof(nrc_cell) ||
        hd.config_bytes != sizeof(tgx_nrc_config))
        return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: network or table sizes differ from this library's", path);
    uint64_t h;
    memcpy(&h, buf.data() + want - 8, 8);
    if (nrc_fnv(buf.data(), want - 8) != h) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: checksum mismatch", path);
    const unsigned char* q = buf.data() + sizeof(hd);
    tgx_nrc_config cfg;
    memcpy(&cfg, q, sizeof(cfg));
    q += sizeof(cfg);
    /* the stored configuration passes the same checks as a new cache's;
     * a rejected file leaves the current cache as it was */
    std::vector<float> arr(4ull * NRC_N);
    memcpy(arr.data(), q + 64, 4ull * 4 * NRC_N);
    for (float v : arr)
        if (!std::isfinite(v)) return tgx_fail(ctx, TGX_ERR_INVALID_ARG, "%s: non-finite weights or moments", path);
    int rc = tgx_nrc_create(ctx, &cfg);
    if (rc) return rc;
    TGX_HIP_CHECK(ctx, hipMemcpy(ctx->nrc_dev.ptr, q, 64, hipMemcpyHostToDevice));
    q += 64;
    tgx_buffer* arrays[4] = {&ctx->nrc_master, &ctx->nrc_m, &ctx->nrc_v, &ctx->nrc_ema};
    for (tgx_buffer* b : arrays) {
        TGX_HIP_CHECK(ctx, hipMemcpy(b->ptr, q, 4ull * NRC_N, hipMemcpyHostToDevice));
        q += 4ull * NRC_N;
    }
    TGX_HIP_CHECK(ctx, hipMemcpy(ctx->nrc_cells.ptr, q, sizeof(nrc_cell) * (size_t)NRC_CELLS, hipMemcpyHostToDevice));
    /* BF16 fragments of the master weights and of the moving average */
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_fwd.ptr, 0, 2ull * NRC_FWD_ELEMS, ctx->stream));
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_bwd.ptr, 0, 2ull * NRC_BWD_ELEMS, ctx->stream));
    hipLaunchKernelGGL(nrc_pack_k, dim3(nrc_blocks(NRC_N, 256)), dim3(256), 0, ctx->stream,
                       (const float*)ctx->nrc_master.ptr, (__bf16*)ctx->nrc_fwd.ptr, (__bf16*)ctx->nrc_bwd.ptr);
    TGX_LAUNCH_CHECK(ctx, "nrc_pack_k");
    TGX_HIP_CHECK(ctx, hipMemsetAsync(ctx->nrc_efwd.ptr, 0, 2ull * NRC_FWD_ELEMS, ctx->stream));
    hipLaunchKernelGGL(nrc_pack_k, dim3(nrc_blocks(NRC_N, 256)), dim3(256), 0, ctx->stream,
                       (const float*)ctx->nrc_ema.ptr, (__bf16*)ctx->nrc_efwd.ptr, (__bf16*)nullptr);
    TGX_LAUNCH_CHECK(ctx, "nrc_pack_k");
    TGX_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TGX_OK;
}

} /* extern "C" */
