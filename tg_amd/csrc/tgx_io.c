/*
 * tgx_io.c — on-disk scene format (host only). tg builds its scene in
 * memory on every start (tgvk_raytracer.c:805-992) and has no file format;
 * this one stores tg's data model byte for byte so a scene can be saved
 * once and uploaded again.
 *
 * Layout (little-endian):
 *   "TGXSCN01"                      8 B magic
 *   u32 version (1), u32 flags      bit 0: materials present
 *   u32 n_objects, n_cluster_pointers, n_clusters, n_luts
 *   objects[n_objects]              96 B each (tgx_object)
 *   cluster_pointers[n_cptr]        u32
 *   cluster_idx_to_object[n_c]      u32
 *   bits[16 n_c]                    u32
 *   materials[512 n_c]              u8, if flag bit 0
 *   luts[256 n_luts]                16 B each (tgx_lut_entry)
 *   u64 FNV-1a 64 of every byte above
 */
#include "tgx.h"

#include <stdio.h>
#include <string.h>

static const char k_magic[8] = {'T', 'G', 'X', 'S', 'C', 'N', '0', '1'};

typedef struct {
    FILE*    f;
    uint64_t h;
    int      bad;
} io_stream;

static void fnv(uint64_t* h, const void* p, size_t n)
{
    const unsigned char* b = (const unsigned char*)p;
    uint64_t x = *h;
    for (size_t i = 0; i < n; i++) {
        x ^= b[i];
        x *= 0x100000001B3ull;
    }
    *h = x;
}

static void put(io_stream* s, const void* p, size_t n)
{
    if (s->bad || n == 0) return;
    if (fwrite(p, 1, n, s->f) != n) s->bad = 1;
    fnv(&s->h, p, n);
}

static void get(io_stream* s, void* p, size_t n)
{
    if (s->bad || n == 0) return;
    if (fread(p, 1, n, s->f) != n) { s->bad = 1; return; }
    fnv(&s->h, p, n);
}

/* skip n bytes while hashing them */
static void skip(io_stream* s, size_t n)
{
    unsigned char buf[1 << 16];
    while (!s->bad && n) {
        const size_t k = n < sizeof(buf) ? n : sizeof(buf);
        get(s, buf, k);
        n -= k;
    }
}

typedef struct {
    uint32_t version, flags, n_objects, n_cluster_pointers, n_clusters, n_luts;
} file_header;

static int read_header(io_stream* s, file_header* hd)
{
    char magic[8];
    get(s, magic, 8);
    get(s, hd, sizeof(*hd));
    if (s->bad || memcmp(magic, k_magic, 8) != 0 || hd->version != 1u || (hd->flags & ~1u) || hd->n_luts == 0)
        return TGX_ERR_INVALID_ARG;
    return TGX_OK;
}

int tgx_scene_save(const char* path, const tgx_scene_desc* d)
{
    if (!path || !d || !d->objects || !d->cluster_pointers || !d->cluster_idx_to_object || !d->bits || !d->luts ||
        d->n_luts == 0)
        return TGX_ERR_INVALID_ARG;
    io_stream s = {fopen(path, "wb"), 0xCBF29CE484222325ull, 0};
    if (!s.f) return TGX_ERR_IO;
    const file_header hd = {1u, d->materials ? 1u : 0u, d->n_objects, d->n_cluster_pointers, d->n_clusters, d->n_luts};
    put(&s, k_magic, 8);
    put(&s, &hd, sizeof(hd));
    put(&s, d->objects, sizeof(tgx_object) * (size_t)d->n_objects);
    put(&s, d->cluster_pointers, 4u * (size_t)d->n_cluster_pointers);
    put(&s, d->cluster_idx_to_object, 4u * (size_t)d->n_clusters);
    put(&s, d->bits, 64u * (size_t)d->n_clusters);
    if (d->materials) put(&s, d->materials, 512u * (size_t)d->n_clusters);
    put(&s, d->luts, sizeof(tgx_lut_entry) * TGX_LUT_ENTRIES * (size_t)d->n_luts);
    const uint64_t h = s.h;
    if (!s.bad && fwrite(&h, 1, 8, s.f) != 8) s.bad = 1;
    if (fclose(s.f) != 0) s.bad = 1;
    return s.bad ? TGX_ERR_IO : TGX_OK;
}

int tgx_scene_file_counts(const char* path, tgx_scene_counts* out)
{
    if (!path || !out) return TGX_ERR_INVALID_ARG;
    io_stream s = {fopen(path, "rb"), 0xCBF29CE484222325ull, 0};
    if (!s.f) return TGX_ERR_IO;
    file_header hd;
    const int rc = read_header(&s, &hd);
    fclose(s.f);
    if (rc) return rc;
    out->n_objects = hd.n_objects;
    out->n_cluster_pointers = hd.n_cluster_pointers;
    out->n_clusters = hd.n_clusters;
    out->n_luts = hd.n_luts;
    out->has_materials = hd.flags & 1u;
    return TGX_OK;
}

int tgx_scene_load(const char* path, const tgx_scene_counts* counts, tgx_object* objects, uint32_t* cluster_pointers,
                   uint32_t* cluster_idx_to_object, uint32_t* bits, uint8_t* materials, tgx_lut_entry* luts)
{
    if (!path || !counts || !objects || !cluster_pointers || !cluster_idx_to_object || !bits || !luts)
        return TGX_ERR_INVALID_ARG;
    io_stream s = {fopen(path, "rb"), 0xCBF29CE484222325ull, 0};
    if (!s.f) return TGX_ERR_IO;
    file_header hd;
    int rc = read_header(&s, &hd);
    if (rc) { fclose(s.f); return rc; }
    /* the caller's arrays hold exactly `counts`: a header that differs (the
     * file was replaced since tgx_scene_file_counts) is refused before any
     * array is written */
    if (hd.n_objects != counts->n_objects || hd.n_cluster_pointers != counts->n_cluster_pointers ||
        hd.n_clusters != counts->n_clusters || hd.n_luts != counts->n_luts) {
        fclose(s.f);
        return TGX_ERR_INVALID_ARG;
    }
    get(&s, objects, sizeof(tgx_object) * (size_t)hd.n_objects);
    get(&s, cluster_pointers, 4u * (size_t)hd.n_cluster_pointers);
    get(&s, cluster_idx_to_object, 4u * (size_t)hd.n_clusters);
    get(&s, bits, 64u * (size_t)hd.n_clusters);
    if (hd.flags & 1u) {
        if (materials) get(&s, materials, 512u * (size_t)hd.n_clusters);
        else skip(&s, 512u * (size_t)hd.n_clusters);
    } else if (materials) {
        memset(materials, 0, 512u * (size_t)hd.n_clusters);
    }
    get(&s, luts, sizeof(tgx_lut_entry) * TGX_LUT_ENTRIES * (size_t)hd.n_luts);
    uint64_t stored = 0;
    const uint64_t h = s.h;
    if (!s.bad && fread(&stored, 1, 8, s.f) != 8) s.bad = 1;
    const int trailing = !s.bad && fgetc(s.f) != EOF;
    fclose(s.f);
    if (s.bad || trailing || stored != h) return TGX_ERR_INVALID_ARG;
    return TGX_OK;
}
