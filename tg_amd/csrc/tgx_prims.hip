/*
 * tgx_prims.hip — device-wide scan and stable radix sort (rocPRIM) used by
 * the SVO build. Kept in one translation unit: rocPRIM's templates are
 * slow to compile.
 */
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "tgx_internal.h"

int tgx_prim_exclusive_scan_u32(tgx_context* ctx, const uint32_t* in, uint32_t* out, uint32_t n)
{
    if (n == 0) return TGX_OK;
    size_t bytes = 0;
    TGX_HIP_CHECK(ctx, rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), ctx->stream));
    int rc = tgx_buffer_reserve(ctx, &ctx->prim_tmp, bytes, "scan scratch");
    if (rc) return rc;
    TGX_HIP_CHECK(ctx, rocprim::exclusive_scan(ctx->prim_tmp.ptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(),
                                               ctx->stream));
    return TGX_OK;
}

/* Stable ascending sort of (key, value) by the low end_bit bits of key. */
int tgx_prim_sort_pairs_u32(tgx_context* ctx, const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in,
                            uint32_t* vals_out, uint32_t n, unsigned end_bit)
{
    if (n == 0) return TGX_OK;
    size_t bytes = 0;
    TGX_HIP_CHECK(ctx, rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, n, 0u, end_bit,
                                                 ctx->stream));
    int rc = tgx_buffer_reserve(ctx, &ctx->prim_tmp, bytes, "sort scratch");
    if (rc) return rc;
    TGX_HIP_CHECK(ctx, rocprim::radix_sort_pairs(ctx->prim_tmp.ptr, bytes, keys_in, keys_out, vals_in, vals_out, n, 0u,
                                                 end_bit, ctx->stream));
    return TGX_OK;
}
