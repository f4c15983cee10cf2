/*
 * ORACLE (test infrastructure only): RNG, viewport, fp32 ray generation and
 * the HDRI importance tables.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.14159265358979323846

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: ten rounds of two 32x32->64 multiplies with the
 * published multipliers and Weyl key increments. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += W0;
            k1 += W1;
        }
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

void orc_uniform4(uint64_t seed, uint32_t pixel, uint32_t sample,
                  uint32_t bounce, uint32_t block, uint32_t stream,
                  float out[4])
{
    uint32_t ctr[4] = {pixel, sample, (bounce << 20) | block, stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    for (int k = 0; k < 4; k++)
        out[k] = (float)(r[k] >> 8) * 0x1p-24f;
}

static void unit3(double v[3])
{
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0) {
        v[0] /= l;
        v[1] /= l;
        v[2] /= l;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
}

static void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* src/ray_tracing/viewport.c:38-88. */
void orc_viewport(const mrt_camera *cam, uint32_t w, uint32_t h,
                  double out[19])
{
    double dir[3] = {cam->dir[0], cam->dir[1], cam->dir[2]};
    unit3(dir);
    double focus = cam->focus_dist > 0.0f ? cam->focus_dist : 1.0;
    double width = 2.0 * tan((double)cam->fov_deg * (PI_D / 180.0) / 2.0)
                   * focus;
    double height = width / ((double)w / (double)h);
    double up[3] = {0, 1, 0};
    if (fabs(dir[1]) > 0.999) {
        up[0] = 1;
        up[1] = 0;
    }
    double nd[3] = {-dir[0], -dir[1], -dir[2]}, u[3], cv[3];
    cross3(up, nd, u);
    unit3(u);
    cross3(dir, u, cv);
    double vu[3], vv[3], du[3], dv[3], ul[3], p00[3];
    for (int k = 0; k < 3; k++) {
        vu[k] = u[k] * width;
        vv[k] = cv[k] * height;
        du[k] = vu[k] / (double)w;
        dv[k] = vv[k] / (double)h;
        ul[k] = ((double)cam->pos[k] + dir[k] * focus) - vu[k] / 2 - vv[k] / 2;
        p00[k] = ul[k] + (du[k] + dv[k]) * 0.5;
    }
    double rad = focus * tan((double)cam->defocus_angle_deg * PI_D / 360.0);
    for (int k = 0; k < 3; k++) {
        out[k] = cam->pos[k];
        out[3 + k] = p00[k];
        out[6 + k] = du[k];
        out[9 + k] = dv[k];
        out[12 + k] = u[k] * rad;
        out[15 + k] = cv[k] * rad;
    }
    out[18] = rad;
}

/* Shirley-Chiu concentric map: uniform on the unit disk, the distribution of
 * random_in_unit_disk (src/ray_tracing/focus.c:16-29) without rejection. */
static void concentric_disk(float a, float b, float *x, float *y)
{
    float sx = 2.0f * a - 1.0f, sy = 2.0f * b - 1.0f;
    if (sx == 0.0f && sy == 0.0f) {
        *x = *y = 0.0f;
        return;
    }
    float r, phi;
    if (fabsf(sx) > fabsf(sy)) {
        r = sx;
        phi = 0.78539816339744831f * (sy / sx);
    } else {
        r = sy;
        phi = 1.57079632679489662f - 0.78539816339744831f * (sx / sy);
    }
    *x = r * cosf(phi);
    *y = r * sinf(phi);
}

/* calc_one_sample_task (src/ray_tracing/calc_sample.c:44-58) with a per-pixel
 * jitter drawn from the counter RNG (the reference draws one jitter per
 * sample for all pixels, render.c:33; per pixel the expectation is the
 * same), then the direction normalised as vec3_unit does. */
void orc_raygen_f32(const mrt_camera *cam, uint32_t w, uint32_t h,
                    uint32_t sample, uint64_t seed, float *orig, float *dir)
{
    double vp[19];
    orc_viewport(cam, w, h, vp);
    float o[3], p00[3], du[3], dv[3], ku[3], kv[3];
    for (int k = 0; k < 3; k++) {
        o[k] = (float)vp[k];
        p00[k] = (float)vp[3 + k];
        du[k] = (float)vp[6 + k];
        dv[k] = (float)vp[9 + k];
        ku[k] = (float)vp[12 + k];
        kv[k] = (float)vp[15 + k];
    }
    int lens = vp[18] > 0.0;
    uint32_t n = w * h;
    for (uint32_t i = 0; i < n; i++) {
        float u4[4];
        orc_uniform4(seed, i, sample, 0, 0, 0, u4);
        float oo[3] = {o[0], o[1], o[2]};
        if (lens) {
            float lx, ly;
            concentric_disk(u4[2], u4[3], &lx, &ly);
            for (int k = 0; k < 3; k++)
                oo[k] = o[k] + (ku[k] * lx + kv[k] * ly);
        }
        float fx = (float)(i % w) + u4[0];
        float fy = (float)(i / w) + u4[1];
        float d[3];
        for (int k = 0; k < 3; k++)
            d[k] = ((p00[k] + du[k] * fx) + dv[k] * fy) - oo[k];
        float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        for (int k = 0; k < 3; k++) {
            orig[(size_t)i * 3 + k] = oo[k];
            dir[(size_t)i * 3 + k] = len > 0.0f ? d[k] / len : 0.0f;
        }
    }
}

/* make_importance_sampling_map and its helpers
 * (importance_sampling_manager.c:41-61, calc_pdf.c:16-114,
 * calc_cdf.c:15-122).  The reference searches each inverse-CDF entry from
 * index 0 (O(W^2 H)); since the targets i/(n-1) increase with i, the first
 * index with cdf >= target never decreases, so one forward scan per row
 * returns the same indices. */
int orc_hdri_tables(const float *rgb, uint32_t w, uint32_t h, float exposure,
                    float *pdf_joint, float *marg_inv, float *cond_inv,
                    double *column_average)
{
    size_t n = (size_t)w * h;
    double *gray = malloc(n * sizeof(double));
    double *raw = malloc(h * sizeof(double));
    double *cdf = malloc((w > h ? w : h) * sizeof(double));
    if (!gray || !raw || !cdf) {
        free(gray);
        free(raw);
        free(cdf);
        return -1;
    }
    double scale = exposure == 0.0f ? 1.0 : (double)powf(2.0f, exposure);
    for (size_t i = 0; i < n; i++)
        gray[i] = 0.2126 * ((double)rgb[i * 3] * scale)
                  + 0.7152 * ((double)rgb[i * 3 + 1] * scale)
                  + 0.0722 * ((double)rgb[i * 3 + 2] * scale);
    for (uint32_t y = 0; y < h; y++) {
        double s = 0;
        for (uint32_t x = 0; x < w; x++)
            s += gray[(size_t)y * w + x];
        raw[y] = s / w;
    }
    double col = 0;
    for (uint32_t y = 0; y < h; y++)
        col += raw[y];
    col = col / h;
    if (column_average)
        *column_average = col;
    for (size_t i = 0; i < n; i++)
        pdf_joint[i] = (float)(gray[i] / col);
    /* marginal: pdf_marginal[y] = raw[y] / (col / 255) */
    double sum = 0;
    for (uint32_t y = 0; y < h; y++) {
        sum += raw[y] / (col / 255);
        cdf[y] = sum;
    }
    if (sum > 0.0)
        for (uint32_t y = 0; y < h; y++)
            cdf[y] /= sum;
    memset(marg_inv, 0, h * sizeof(float));
    for (uint32_t i = 0, y = 0; i < h; i++) {
        double u = (double)i / (h - 1);
        while (y < h && !(cdf[y] >= u))
            y++;
        if (y < h)
            marg_inv[i] = (float)((double)y / (h - 1));
        else
            break;
    }
    /* conditional rows: pdf_conditional = gray / (raw[y] / 255) */
    memset(cond_inv, 0, n * sizeof(float));
    for (uint32_t y = 0; y < h; y++) {
        sum = 0;
        for (uint32_t x = 0; x < w; x++) {
            sum += gray[(size_t)y * w + x] / (raw[y] / 255);
            cdf[x] = sum;
        }
        if (sum > 0.0)
            for (uint32_t x = 0; x < w; x++)
                cdf[x] /= sum;
        for (uint32_t i = 0, x = 0; i < w; i++) {
            double u = (double)i / (double)(w - 1);
            while (x < w && !(cdf[x] >= u))
                x++;
            if (x < w)
                cond_inv[(size_t)y * w + i] = (float)((double)x / (double)(w - 1));
            else
                break;
        }
    }
    free(gray);
    free(raw);
    free(cdf);
    return 0;
}

/* The MIS environment density (a new estimator, not in the reference):
 * cell (y, x) holds the directions whose texel lookup (background.c:34-57)
 * reads texel (y, x); its weight is luminance x (sin e_hi - sin e_lo), the
 * exact integral of cos(elevation) over the cell's elevation band. */
int orc_hdri_mis_tables(const float *rgb, uint32_t w, uint32_t h,
                        float exposure, float *marg, float *cond,
                        float *pcell, double *total_out)
{
    const uint32_t nw = w - 1, nh = h - 1;
    double *wg = malloc((size_t)nw * nh * sizeof(double));
    double *rs = malloc(nh * sizeof(double));
    if (!wg || !rs) {
        free(wg);
        free(rs);
        return -1;
    }
    double scale = exposure == 0.0f ? 1.0 : (double)powf(2.0f, exposure);
    double total = 0.0;
    for (uint32_t y = 0; y < nh; y++) {
        double band = sin(((double)(y + 1) / nh - 0.5) * PI_D)
                      - sin(((double)y / nh - 0.5) * PI_D);
        double s = 0.0;
        for (uint32_t x = 0; x < nw; x++) {
            const float *t = rgb + ((size_t)y * w + x) * 3;
            double lum = 0.2126 * ((double)t[0] * scale)
                         + 0.7152 * ((double)t[1] * scale)
                         + 0.0722 * ((double)t[2] * scale);
            wg[(size_t)y * nw + x] = lum > 0.0 ? lum * band : 0.0;
            s += wg[(size_t)y * nw + x];
        }
        rs[y] = s;
        total += s;
    }
    double a = 0.0;
    marg[0] = 0.0f;
    for (uint32_t y = 0; y < nh; y++) {
        a += rs[y];
        marg[y + 1] = total > 0.0 ? (float)(a / total) : (float)(y + 1) / nh;
        double ra = 0.0;
        cond[(size_t)y * w] = 0.0f;
        for (uint32_t x = 0; x < nw; x++) {
            ra += wg[(size_t)y * nw + x];
            cond[(size_t)y * w + x + 1] =
                rs[y] > 0.0 ? (float)(ra / rs[y]) : (float)(x + 1) / nw;
            pcell[(size_t)y * nw + x] =
                total > 0.0 ? (float)(wg[(size_t)y * nw + x] / total) : 0.0f;
        }
        cond[(size_t)y * w + nw] = 1.0f;
    }
    marg[nh] = 1.0f;
    if (total_out)
        *total_out = total;
    free(wg);
    free(rs);
    return 0;
}
