/*
 * PNG decoding for vkvg_surface_create_from_image (src/vkvg_surface.c:181-194).
 *
 * The reference decodes with stb_image (stbi_load with 4 requested components; the header is vendored
 * as src/stb_image.h). This file decodes the PNG format (ISO/IEC 15948) with zlib's inflate
 * and returns what stbi_load returns for a PNG: 8-bit RGBA rows, top row first, not premultiplied.
 * Conversions follow that decoder's documented results: grey g -> (g, g, g, 255), grey + alpha ->
 * (g, g, g, a), RGB -> (r, g, b, 255), palette entries with their tRNS alpha (255 without), a tRNS
 * colour key of a grey or RGB image makes the matching pixels alpha 0, grey of 1, 2 or 4 bits is
 * scaled by 255, 85 or 17, 16-bit samples keep their high byte. Adam7-interlaced images are
 * supported. Chunk CRCs are not checked (stb_image does not check them either). Other image formats
 * that stb_image reads (JPEG, GIF, BMP, ...) are not decoded.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <zlib.h>

#include "vkvg_amd.h"

#define PNG_MAX_SIDE 32768u /* the largest surface side of this backend */

static uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

static int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc)
        return a;
    return pb <= pc ? b : c;
}

/* undoes the per-row filters of one (sub-)image in place: rows of 1 + rowbytes bytes */
static int unfilter(uint8_t *d, uint32_t rows, size_t rowbytes, uint32_t bpp) {
    const uint8_t *prev = NULL;
    for (uint32_t y = 0; y < rows; y++) {
        uint8_t *row = d + (size_t)y * (rowbytes + 1);
        const uint8_t f = row[0];
        uint8_t *r = row + 1;
        for (size_t i = 0; i < rowbytes; i++) {
            const int a = i >= bpp ? r[i - bpp] : 0, b = prev ? prev[i] : 0, c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            switch (f) {
            case 0: break;
            case 1: r[i] = (uint8_t)(r[i] + a); break;
            case 2: r[i] = (uint8_t)(r[i] + b); break;
            case 3: r[i] = (uint8_t)(r[i] + ((a + b) >> 1)); break;
            case 4: r[i] = (uint8_t)(r[i] + paeth(a, b, c)); break;
            default: return -1;
            }
        }
        prev = r;
    }
    return 0;
}

typedef struct {
    uint32_t w, h, depth, ctype, channels;
    uint8_t  pal[256][4];
    uint32_t n_pal;
    int      has_key;
    uint16_t key[3]; /* tRNS colour key of a grey (key[0]) or RGB image: raw 16-bit samples for
                      * 16-bit images, else the key's low byte scaled to 8 bits as stb_image does */
} png_info;

static uint32_t sample(const uint8_t *row, uint32_t depth, size_t i) {
    switch (depth) {
    case 16: return (uint32_t)row[2 * i] << 8 | row[2 * i + 1];
    case 8: return row[i];
    default: { /* 1, 2, 4 bits, most significant first */
        const size_t bit = i * depth;
        return (uint32_t)(row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1u);
    }
    }
}

/* converts one unfiltered row of `n` pixels into RGBA8 at out, every `step`-th pixel */
static void expand_row(const png_info *in, const uint8_t *row, uint32_t n, uint8_t *out, uint32_t step) {
    static const uint32_t grey_scale[17] = {0, 255, 85, 0, 17, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1};
    for (uint32_t x = 0; x < n; x++) {
        uint8_t *o = out + (size_t)x * step * 4;
        const size_t s = (size_t)x * in->channels;
        uint32_t v[4];
        for (uint32_t c = 0; c < in->channels; c++)
            v[c] = sample(row, in->depth, s + c);
        const uint32_t sh = in->depth == 16 ? 8 : 0; /* 16 bits: the high byte */
        switch (in->ctype) {
        case 0: { /* grey */
            const uint8_t g = (uint8_t)(in->depth < 8 ? v[0] * grey_scale[in->depth] : v[0] >> sh);
            o[0] = o[1] = o[2] = g;
            o[3] = (in->has_key && (in->depth == 16 ? v[0] : g) == in->key[0]) ? 0 : 255;
            break;
        }
        case 2: /* RGB */
            o[0] = (uint8_t)(v[0] >> sh);
            o[1] = (uint8_t)(v[1] >> sh);
            o[2] = (uint8_t)(v[2] >> sh);
            if (in->depth == 16)
                o[3] = (in->has_key && v[0] == in->key[0] && v[1] == in->key[1] && v[2] == in->key[2]) ? 0 : 255;
            else
                o[3] = (in->has_key && o[0] == in->key[0] && o[1] == in->key[1] && o[2] == in->key[2]) ? 0 : 255;
            break;
        case 3: /* palette (an index beyond the palette was rejected before) */
            memcpy(o, in->pal[v[0]], 4);
            break;
        case 4: /* grey + alpha */
            o[0] = o[1] = o[2] = (uint8_t)(v[0] >> sh);
            o[3] = (uint8_t)(v[1] >> sh);
            break;
        default: /* 6: RGBA */
            o[0] = (uint8_t)(v[0] >> sh);
            o[1] = (uint8_t)(v[1] >> sh);
            o[2] = (uint8_t)(v[2] >> sh);
            o[3] = (uint8_t)(v[3] >> sh);
            break;
        }
    }
}

static int check_indices(const png_info *in, const uint8_t *row, uint32_t n) {
    if (in->ctype != 3)
        return 0;
    for (uint32_t x = 0; x < n; x++)
        if (sample(row, in->depth, x) >= in->n_pal)
            return -1;
    return 0;
}

uint8_t *vkvg_amd_png_decode(const uint8_t *buf, size_t n, uint32_t *width, uint32_t *height) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (!buf || n < 8 || memcmp(buf, sig, 8))
        return NULL;
    png_info in;
    memset(&in, 0, sizeof(in));
    uint32_t interlace = 0;
    int have_hdr = 0;
    uint8_t *idat = NULL, *raw = NULL, *rgba = NULL;
    size_t idat_n = 0, idat_cap = 0;
    size_t pos = 8;
    for (uint32_t k = 0; k < 256; k++)
        in.pal[k][3] = 255;
    while (pos + 12 <= n) {
        const uint32_t len = be32(buf + pos);
        const uint8_t *type = buf + pos + 4, *data = buf + pos + 8;
        if ((size_t)len > n - pos - 12)
            goto fail;
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13 || have_hdr)
                goto fail;
            in.w = be32(data);
            in.h = be32(data + 4);
            in.depth = data[8];
            in.ctype = data[9];
            interlace = data[12];
            if (!in.w || !in.h || in.w > PNG_MAX_SIDE || in.h > PNG_MAX_SIDE || data[10] || data[11] || interlace > 1)
                goto fail;
            switch (in.ctype) {
            case 0: in.channels = 1; if (in.depth != 1 && in.depth != 2 && in.depth != 4 && in.depth != 8 && in.depth != 16) goto fail; break;
            case 2: in.channels = 3; if (in.depth != 8 && in.depth != 16) goto fail; break;
            case 3: in.channels = 1; if (in.depth != 1 && in.depth != 2 && in.depth != 4 && in.depth != 8) goto fail; break;
            case 4: in.channels = 2; if (in.depth != 8 && in.depth != 16) goto fail; break;
            case 6: in.channels = 4; if (in.depth != 8 && in.depth != 16) goto fail; break;
            default: goto fail;
            }
            have_hdr = 1;
        } else if (!memcmp(type, "PLTE", 4)) {
            if (!have_hdr || len % 3 || len / 3 > 256 || len == 0)
                goto fail;
            in.n_pal = len / 3;
            for (uint32_t k = 0; k < in.n_pal; k++) {
                in.pal[k][0] = data[3 * k];
                in.pal[k][1] = data[3 * k + 1];
                in.pal[k][2] = data[3 * k + 2];
            }
        } else if (!memcmp(type, "tRNS", 4)) {
            /* stb_image (src/stb_image.h:5006-5039): rejected after IDAT, before PLTE, longer than
             * the palette, or on an image with alpha; a key of an image below 16 bits keeps its
             * low byte, scaled like the samples (x255 / x85 / x17 for 1 / 2 / 4-bit grey) */
            static const uint32_t key_scale[9] = {0, 255, 85, 0, 17, 0, 0, 0, 1};
            if (!have_hdr || idat_n)
                goto fail;
            if (in.ctype == 3) {
                if (!in.n_pal || len > in.n_pal)
                    goto fail;
                for (uint32_t k = 0; k < len; k++)
                    in.pal[k][3] = data[k];
            } else if ((in.ctype == 0 && len == 2) || (in.ctype == 2 && len == 6)) {
                in.has_key = 1;
                for (uint32_t c = 0; c < len / 2; c++) {
                    const uint32_t raw = (uint32_t)(data[2 * c] << 8 | data[2 * c + 1]);
                    in.key[c] = (uint16_t)(in.depth == 16 ? raw : (uint8_t)((raw & 255) * key_scale[in.depth]));
                }
            } else {
                goto fail;
            }
        } else if (!memcmp(type, "IDAT", 4)) {
            if (!have_hdr)
                goto fail;
            if (idat_n + len > idat_cap) {
                size_t cap = idat_cap ? idat_cap : 65536;
                while (cap < idat_n + len)
                    cap *= 2;
                uint8_t *p = (uint8_t *)realloc(idat, cap);
                if (!p)
                    goto fail;
                idat = p;
                idat_cap = cap;
            }
            memcpy(idat + idat_n, data, len);
            idat_n += len;
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        } else if (!(type[0] & 0x20)) {
            goto fail; /* an unknown critical chunk */
        }
        pos += 12 + (size_t)len;
    }
    if (!have_hdr || !idat_n || (in.ctype == 3 && !in.n_pal))
        goto fail;
    {
        /* the (sub-)images: one, or the seven Adam7 passes */
        static const uint32_t x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
        static const uint32_t dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
        const uint32_t passes = interlace ? 7 : 1;
        const uint32_t bits = in.depth * in.channels, bpp = bits < 8 ? 1 : bits / 8;
        uint32_t pw[7], ph[7];
        size_t total = 0;
        for (uint32_t p = 0; p < passes; p++) {
            pw[p] = interlace ? (in.w > x0[p] ? (in.w - x0[p] + dx[p] - 1) / dx[p] : 0) : in.w;
            ph[p] = interlace ? (in.h > y0[p] ? (in.h - y0[p] + dy[p] - 1) / dy[p] : 0) : in.h;
            if (pw[p] && ph[p])
                total += (size_t)ph[p] * (1 + ((size_t)pw[p] * bits + 7) / 8);
        }
        raw = (uint8_t *)malloc(total ? total : 1);
        rgba = (uint8_t *)malloc((size_t)in.w * in.h * 4);
        if (!raw || !rgba)
            goto fail;
        /* inflate exactly the rows the image needs; data after them is ignored */
        if (total > 0xFFFFFFFFu || idat_n > 0xFFFFFFFFu)
            goto fail;
        z_stream zs;
        memset(&zs, 0, sizeof(zs));
        if (inflateInit(&zs) != Z_OK)
            goto fail;
        zs.next_in = idat;
        zs.avail_in = (uInt)idat_n;
        zs.next_out = raw;
        zs.avail_out = (uInt)total;
        const int zr = inflate(&zs, Z_FINISH);
        const size_t got = total - zs.avail_out;
        inflateEnd(&zs);
        if (got != total || (zr != Z_STREAM_END && zr != Z_OK && zr != Z_BUF_ERROR))
            goto fail;
        size_t off = 0;
        for (uint32_t p = 0; p < passes; p++) {
            if (!pw[p] || !ph[p])
                continue;
            const size_t rowbytes = ((size_t)pw[p] * bits + 7) / 8;
            if (unfilter(raw + off, ph[p], rowbytes, bpp))
                goto fail;
            for (uint32_t y = 0; y < ph[p]; y++) {
                const uint8_t *row = raw + off + (size_t)y * (rowbytes + 1) + 1;
                if (check_indices(&in, row, pw[p]))
                    goto fail;
                const uint32_t ty = interlace ? y0[p] + y * dy[p] : y;
                uint8_t *o = rgba + ((size_t)ty * in.w + (interlace ? x0[p] : 0)) * 4;
                expand_row(&in, row, pw[p], o, interlace ? dx[p] : 1);
            }
            off += (size_t)ph[p] * (rowbytes + 1);
        }
    }
    free(idat);
    free(raw);
    *width = in.w;
    *height = in.h;
    return rgba;
fail:
    free(idat);
    free(raw);
    free(rgba);
    return NULL;
}

void vkvg_amd_png_free(uint8_t *rgba) { free(rgba); }
