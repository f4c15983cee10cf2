/*
 * Texture images: BMP and PNG to RGBA8 texels, as the reference's own
 * readers lay them out (get_texture_image, src/parsing/get_texture.c:16-43).
 *
 * BMP (src/utils/bmp): BITMAPINFOHEADER images of 8 (palette), 16
 * (5-6-5), 24 or 32 bits per pixel, uncompressed; rows kept in file order
 * (extract_raw_pixels, bmp_extract_pixel.c:57-84), each padded to 4 bytes;
 * alpha 255 except for 32-bit pixels.  The reference reads 1- and 4-bit
 * pixels as whole bytes; those are refused here.
 *
 * PNG (src/utils/png): 8-bit, non-interlaced images of colour types 0,
 * 2, 3, 4 and 6; inflated with zlib (the reference has its own inflate),
 * unfiltered (None / Sub / Up / Average / Paeth), and stored with the last
 * scanline first (extract_rgba, decode_deflate_utils.c:79-101).  Pixels
 * follow get_pixel_rgva (:59-77): one channel is grey (the palette index
 * for type 3, which the reference does not look up), two channels take the
 * grey from byte x (not 2x) and the alpha from byte 2x + 1.
 */
#include "mrt_host_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static uint32_t rd32le(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16)
           | ((uint32_t)p[3] << 24);
}

static uint32_t rd32be(const uint8_t *p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16)
           | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

static int decode_bmp(const uint8_t *d, size_t n, uint8_t **out, uint32_t *ow,
                      uint32_t *oh)
{
    if (n < 54)
        return MRT_E_INVALID;
    uint32_t off = rd32le(d + 10), hsz = rd32le(d + 14);
    int32_t w = (int32_t)rd32le(d + 18), h = (int32_t)rd32le(d + 22);
    uint32_t bpp = (uint32_t)d[28] | ((uint32_t)d[29] << 8);
    uint32_t comp = rd32le(d + 30), ncol = rd32le(d + 46);
    if (hsz < 40 || w <= 0 || h == 0 || w > 65536 || h > 65536 || h < -65536)
        return MRT_E_INVALID;
    if (comp != 0 && !(comp == 3 && (bpp == 16 || bpp == 32)))
        return MRT_E_UNSUPPORTED;
    if (bpp != 8 && bpp != 16 && bpp != 24 && bpp != 32)
        return MRT_E_UNSUPPORTED;
    uint32_t W = (uint32_t)w, H = (uint32_t)(h < 0 ? -h : h);
    uint32_t pal[256];
    if (bpp == 8) {
        if (ncol == 0 || ncol > 256)
            ncol = 256;
        size_t p0 = 14 + (size_t)hsz;
        for (uint32_t i = 0; i < 256; i++) {
            pal[i] = 255u << 24;            /* out of range: black */
            if (i < ncol && p0 + 4 * (size_t)i + 4 <= n) {
                const uint8_t *q = d + p0 + 4 * (size_t)i;
                pal[i] = (255u << 24) | ((uint32_t)q[2] << 16)
                         | ((uint32_t)q[1] << 8) | q[0];
            }
        }
    }
    size_t row = ((size_t)W * bpp + 7) / 8;
    size_t pad = (4 - row % 4) % 4;
    if ((size_t)off + (row + pad) * H - pad > n)
        return MRT_E_INVALID;
    uint8_t *rgba = malloc((size_t)W * H * 4);
    if (!rgba)
        return MRT_E_NOMEM;
    const uint8_t *p = d + off;
    for (uint32_t y = 0; y < H; y++, p += row + pad) {
        for (uint32_t x = 0; x < W; x++) {
            uint32_t c;
            if (bpp == 8) {
                c = pal[p[x]];
            } else if (bpp == 16) {
                uint32_t r16 = (uint32_t)p[2 * x] | ((uint32_t)p[2 * x + 1] << 8);
                c = (255u << 24) | ((((r16 >> 11) & 0x1F) << 3) << 16)
                    | ((((r16 >> 5) & 0x3F) << 2) << 8) | ((r16 & 0x1F) << 3);
            } else if (bpp == 24) {
                const uint8_t *q = p + 3 * (size_t)x;
                c = (255u << 24) | ((uint32_t)q[2] << 16) | ((uint32_t)q[1] << 8) | q[0];
            } else {
                const uint8_t *q = p + 4 * (size_t)x;
                c = ((uint32_t)q[3] << 24) | ((uint32_t)q[2] << 16)
                    | ((uint32_t)q[1] << 8) | q[0];
            }
            uint8_t *o = rgba + ((size_t)y * W + x) * 4;
            o[0] = (uint8_t)(c >> 16);
            o[1] = (uint8_t)(c >> 8);
            o[2] = (uint8_t)c;
            o[3] = (uint8_t)(c >> 24);
        }
    }
    *out = rgba;
    *ow = W;
    *oh = H;
    return MRT_OK;
}

static int paeth(int a, int b, int c)
{
    int p = a + b - c;
    int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc)
        return a;
    return pb <= pc ? b : c;
}

static int decode_png(const uint8_t *d, size_t n, uint8_t **out, uint32_t *ow,
                      uint32_t *oh)
{
    size_t i = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, inter = 0, have_hdr = 0;
    uint8_t *idat = NULL;
    size_t nidat = 0;
    int rc = MRT_OK;
    while (i + 12 <= n) {
        uint32_t len = rd32be(d + i);
        const uint8_t *type = d + i + 4, *body = d + i + 8;
        if ((size_t)len > n - i - 12) {
            rc = MRT_E_INVALID;
            break;
        }
        if (!memcmp(type, "IHDR", 4) && len >= 13) {
            W = rd32be(body);
            H = rd32be(body + 4);
            depth = body[8];
            ctype = body[9];
            inter = body[12];
            have_hdr = 1;
        } else if (!memcmp(type, "IDAT", 4)) {
            uint8_t *p = realloc(idat, nidat + len);
            if (!p) {
                rc = MRT_E_NOMEM;
                break;
            }
            idat = p;
            memcpy(idat + nidat, body, len);
            nidat += len;
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        i += 12 + (size_t)len;
    }
    int bpp = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2
              : ctype == 6 ? 4 : -1;
    if (rc == MRT_OK && (!have_hdr || !idat || W == 0 || H == 0 || W > 65536
                         || H > 65536))
        rc = MRT_E_INVALID;
    if (rc == MRT_OK && (depth != 8 || bpp < 0 || inter != 0))
        rc = MRT_E_UNSUPPORTED;
    if (rc != MRT_OK) {
        free(idat);
        return rc;
    }
    size_t stride = 1 + (size_t)W * bpp;
    uLongf raw_n = (uLongf)(stride * H);
    uint8_t *raw = malloc(raw_n);
    uint8_t *rgba = malloc((size_t)W * H * 4);
    if (!raw || !rgba) {
        free(idat);
        free(raw);
        free(rgba);
        return MRT_E_NOMEM;
    }
    uLongf got = raw_n;
    int z = uncompress(raw, &got, idat, (uLong)nidat);
    free(idat);
    if (z != Z_OK || got != raw_n) {
        free(raw);
        free(rgba);
        return MRT_E_INVALID;
    }
    for (uint32_t y = 0; y < H; y++) {
        uint8_t *r = raw + (size_t)y * stride, *s = r + 1;
        const uint8_t *up = y ? raw + (size_t)(y - 1) * stride + 1 : NULL;
        int f = r[0];
        if (f > 4) {
            free(raw);
            free(rgba);
            return MRT_E_INVALID;
        }
        for (size_t x = 0; x < (size_t)W * bpp; x++) {
            int a = x >= (size_t)bpp ? s[x - bpp] : 0;
            int b = up ? up[x] : 0;
            int c = (up && x >= (size_t)bpp) ? up[x - bpp] : 0;
            int v = s[x];
            if (f == 1)
                v += a;
            else if (f == 2)
                v += b;
            else if (f == 3)
                v += (a + b) >> 1;
            else if (f == 4)
                v += paeth(a, b, c);
            s[x] = (uint8_t)v;
        }
    }
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t *px = raw + (size_t)(H - 1 - y) * stride + 1;
        for (uint32_t x = 0; x < W; x++) {
            uint8_t *o = rgba + ((size_t)y * W + x) * 4;
            if (bpp == 1) {
                o[0] = o[1] = o[2] = px[x];
                o[3] = 255;
            } else if (bpp == 2) {
                o[0] = o[1] = o[2] = px[x];
                o[3] = px[2 * (size_t)x + 1];
            } else {
                o[0] = px[(size_t)x * bpp];
                o[1] = px[(size_t)x * bpp + 1];
                o[2] = px[(size_t)x * bpp + 2];
                o[3] = bpp == 4 ? px[(size_t)x * 4 + 3] : 255;
            }
        }
    }
    free(raw);
    *out = rgba;
    *ow = W;
    *oh = H;
    return MRT_OK;
}

int mrt_decode_image(const uint8_t *data, size_t size, uint8_t **rgba,
                     uint32_t *w, uint32_t *h)
{
    if (!data || !rgba || !w || !h)
        return MRT_E_INVALID;
    *rgba = NULL;
    static const uint8_t png_sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (size >= 8 && !memcmp(data, png_sig, 8))
        return decode_png(data, size, rgba, w, h);
    if (size >= 2 && data[0] == 'B' && data[1] == 'M')
        return decode_bmp(data, size, rgba, w, h);
    return MRT_E_UNSUPPORTED;
}

int mrt_load_image(const char *path, uint8_t **rgba, uint32_t *w, uint32_t *h)
{
    if (!path || !rgba)
        return MRT_E_INVALID;
    FILE *f = fopen(path, "rb");
    if (!f)
        return MRT_E_INVALID;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz <= 0 || sz > (1L << 30)) {
        fclose(f);
        return MRT_E_LIMIT;
    }
    uint8_t *data = malloc((size_t)sz);
    if (!data) {
        fclose(f);
        return MRT_E_NOMEM;
    }
    size_t got = fread(data, 1, (size_t)sz, f);
    fclose(f);
    int rc = mrt_decode_image(data, got, rgba, w, h);
    free(data);
    return rc;
}
