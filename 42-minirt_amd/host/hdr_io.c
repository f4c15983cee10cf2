/*
 * Radiance .hdr (RGBE) reader: the format parse_hdr / parse_hdr_header /
 * read_hdr_data accept (src/utils/hdr/parse_hdr*.c): a "#?RADIANCE" or
 * "#?RGBE" header with FORMAT=32-bit_rle_rgbe and optional EXPOSURE= and
 * GAMMA= lines, a resolution line "-Y h +X w" (signs flip the axes,
 * hdr_set_pixel.c), then new-style run-length scanlines (0x02 0x02 + width,
 * four channel runs, parse_hdr_data.c:60-110).  Flat (unencoded) scanlines
 * are accepted as well.
 *
 * Texels are decoded as get_hdr_pixel does (get_hdr_value.c:15-40):
 * c * 2^(e - 128) / 255 per channel, and stored with row 0 at the bottom of
 * the image (the file's last scanline for "-Y"), the order the HDRI lookup
 * and the importance tables use.
 */
#include "mrt_host_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const uint8_t *p;
    size_t         n, i;
} rd;

static int read_line(rd *r, char *buf, size_t cap)
{
    size_t k = 0;
    while (r->i < r->n && r->p[r->i] != '\n') {
        if (k + 1 < cap)
            buf[k++] = (char)r->p[r->i];
        r->i++;
    }
    if (r->i >= r->n)
        return 0;
    r->i++;
    buf[k] = 0;
    return 1;
}

/* One scanline of w RGBE pixels into line[4 * w]. */
static int read_scanline(rd *r, uint32_t w, uint8_t *line)
{
    if (r->i + 4 > r->n)
        return MRT_E_INVALID;
    const uint8_t *h = r->p + r->i;
    if (w >= 8 && w < 32768 && h[0] == 2 && h[1] == 2 && !(h[2] & 0x80)) {
        uint32_t len = ((uint32_t)h[2] << 8) | h[3];
        if (len != w)
            return MRT_E_INVALID;
        r->i += 4;
        for (int c = 0; c < 4; c++) {
            uint32_t x = 0;
            while (x < w) {
                if (r->i >= r->n)
                    return MRT_E_INVALID;
                uint32_t cnt = r->p[r->i++];
                if (cnt > 128) {            /* a run of one byte */
                    cnt -= 128;
                    if (x + cnt > w || r->i >= r->n)
                        return MRT_E_INVALID;
                    uint8_t v = r->p[r->i++];
                    for (uint32_t k = 0; k < cnt; k++)
                        line[4 * (x++) + c] = v;
                } else {                    /* cnt literal bytes */
                    if (cnt == 0 || x + cnt > w || r->i + cnt > r->n)
                        return MRT_E_INVALID;
                    for (uint32_t k = 0; k < cnt; k++)
                        line[4 * (x++) + c] = r->p[r->i++];
                }
            }
        }
        return MRT_OK;
    }
    /* flat scanline */
    if (r->i + 4 * (size_t)w > r->n)
        return MRT_E_INVALID;
    memcpy(line, r->p + r->i, 4 * (size_t)w);
    r->i += 4 * (size_t)w;
    return MRT_OK;
}

int mrt_decode_hdr(const uint8_t *data, size_t size, float **rgb, uint32_t *w,
                   uint32_t *h, float *exposure, float *gamma)
{
    if (!data || !rgb || !w || !h)
        return MRT_E_INVALID;
    *rgb = NULL;
    rd r = {data, size, 0};
    char buf[512];
    if (!read_line(&r, buf, sizeof(buf))
        || (strcmp(buf, "#?RADIANCE") != 0 && strcmp(buf, "#?RGBE") != 0))
        return MRT_E_INVALID;
    int have_format = 0;
    float ex = 0.0f, ga = 1.0f;
    for (;;) {
        if (!read_line(&r, buf, sizeof(buf)))
            return MRT_E_INVALID;
        if (buf[0] == 0)
            break;
        if (strncmp(buf, "FORMAT=", 7) == 0) {
            if (strcmp(buf, "FORMAT=32-bit_rle_rgbe") != 0)
                return MRT_E_UNSUPPORTED;
            have_format = 1;
        } else if (strncmp(buf, "EXPOSURE=", 9) == 0) {
            ex = strtof(buf + 9, NULL);
        } else if (strncmp(buf, "GAMMA=", 6) == 0) {
            ga = strtof(buf + 6, NULL);
        }
    }
    if (!have_format || !read_line(&r, buf, sizeof(buf)))
        return MRT_E_INVALID;
    char sy, sx, ay, ax;
    long ny, nx;
    if (sscanf(buf, "%c%c %ld %c%c %ld", &sy, &ay, &ny, &sx, &ax, &nx) != 6
        || ay != 'Y' || ax != 'X' || (sy != '-' && sy != '+')
        || (sx != '-' && sx != '+') || ny < 2 || nx < 2 || ny > 65536
        || nx > 65536)
        return MRT_E_INVALID;
    const uint32_t W = (uint32_t)nx, H = (uint32_t)ny;
    float *out = malloc((size_t)W * H * 3 * sizeof(float));
    uint8_t *line = malloc(4 * (size_t)W);
    if (!out || !line) {
        free(out);
        free(line);
        return MRT_E_NOMEM;
    }
    int rc = MRT_OK;
    for (uint32_t y = 0; y < H && rc == MRT_OK; y++) {
        rc = read_scanline(&r, W, line);
        if (rc != MRT_OK)
            break;
        /* "-Y": the first scanline is the top row; rows are stored bottom
         * first (hdr_set_pixel.c: y = h - 1 - y when dir_y = -1) */
        uint32_t row = sy == '-' ? H - 1 - y : y;
        for (uint32_t x = 0; x < W; x++) {
            uint32_t col = sx == '-' ? W - 1 - x : x;
            const uint8_t *q = line + 4 * (size_t)x;
            double s = ldexp(1.0, (int)q[3] - 128);
            float *o = out + ((size_t)row * W + col) * 3;
            o[0] = (float)(q[0] * s / 255.0);
            o[1] = (float)(q[1] * s / 255.0);
            o[2] = (float)(q[2] * s / 255.0);
        }
    }
    free(line);
    if (rc != MRT_OK) {
        free(out);
        return rc;
    }
    *rgb = out;
    *w = W;
    *h = H;
    if (exposure)
        *exposure = ex;
    if (gamma)
        *gamma = ga;
    return MRT_OK;
}

int mrt_load_hdr(const char *path, float **rgb, uint32_t *w, uint32_t *h,
                 float *exposure, float *gamma)
{
    if (!path || !rgb)
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
    int rc = mrt_decode_hdr(data, got, rgb, w, h, exposure, gamma);
    free(data);
    return rc;
}

void mrt_free(void *p) { free(p); }
