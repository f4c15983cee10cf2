/*
 * Output formats of the accumulation buffer: the P6 PPM export
 * (src/utils/ppm/export_ppm.c:42-87 with render_float_to_int_gamma,
 * src/upscaling/utils.c:15-35) and the render_state.rts checkpoint
 * (src/utils/export_render_state.c:15-51), and the BVH heat map of render
 * mode 1.
 */
#include "mrt_host_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

/* get_heatmap_fcolor: blue -> cyan -> green -> yellow -> red in four float
 * ramps of ratio = min(current / max, 1); max == 0 gives blue. */
static void heat_color(int cur, int max, float c[3])
{
    c[0] = c[1] = c[2] = 0.0f;
    if (max == 0) {
        c[2] = 1.0f;
        return;
    }
    float r = fminf((float)cur / (float)max, 1.0f);
    if (r < 0.25f) {
        c[2] = 4.0f * r;
    } else if (r < 0.5f) {
        c[1] = 1.0f;
        c[2] = 1.0f - 4.0f * (r - 0.25f);
    } else if (r < 0.75f) {
        c[0] = 4.0f * (r - 0.5f);
        c[1] = 1.0f;
    } else {
        c[0] = 1.0f;
        c[1] = 1.0f - 4.0f * (r - 0.75f);
    }
}

void mrt_heat_map_add(const int32_t *depth, uint32_t w, uint32_t h,
                      double *acc)
{
    if (!depth || !acc)
        return;
    size_t n = (size_t)w * h;
    int max = 0;
    for (size_t i = 0; i < n; i++)
        if (max < depth[i])
            max = depth[i];
    for (size_t i = 0; i < n; i++) {
        if (depth[i] < 0)
            continue;
        float c[3];
        heat_color(depth[i], max, c);
        for (int k = 0; k < 3; k++)
            acc[3 * i + k] += (double)c[k];
    }
}

static unsigned char to_byte(double acc, uint32_t samples, double gamma)
{
    double v = acc / (double)samples;
    if (gamma != 1.0)
        v = pow(v, 1.0 / gamma);
    if (!(v > 0.0))
        v = 0.0;
    if (v > 1.0)
        v = 1.0;
    return (unsigned char)(v * 255.0);
}

int mrt_export_ppm(const char *path, const double *acc, uint32_t w,
                   uint32_t h, uint32_t samples, double gamma)
{
    if (!path || !acc || !w || !h || !samples || !(gamma > 0.0))
        return MRT_E_INVALID;
    FILE *f = fopen(path, "wb");
    if (!f)
        return MRT_E_INVALID;
    size_t n = (size_t)w * h * 3;
    unsigned char *buf = malloc(n);
    if (!buf) {
        fclose(f);
        return MRT_E_NOMEM;
    }
    for (size_t i = 0; i < n; i++)
        buf[i] = to_byte(acc[i], samples, gamma);
    int ok = fprintf(f, "P6\n%u %u\n255\n", w, h) > 0 && fwrite(buf, 1, n, f) == n;
    free(buf);
    ok = (fclose(f) == 0) && ok;
    return ok ? MRT_OK : MRT_E_INVALID;
}

int mrt_export_rts(const char *path, const double *acc, uint32_t w,
                   uint32_t h, uint32_t samples)
{
    if (!path || !acc || !w || !h)
        return MRT_E_INVALID;
    FILE *f = fopen(path, "wb");
    if (!f)
        return MRT_E_INVALID;
    int32_t hdr[3] = {(int32_t)w, (int32_t)h, (int32_t)samples};
    size_t n = (size_t)w * h * 3;
    int ok = fwrite(hdr, sizeof(hdr), 1, f) == 1
             && fwrite(acc, sizeof(double), n, f) == n;
    ok = (fclose(f) == 0) && ok;
    return ok ? MRT_OK : MRT_E_INVALID;
}

/* Unlike init_float_render (:34-51), which checks the window size, this
 * checks the render size the buffer was written at. */
int mrt_load_rts(const char *path, double *acc, uint32_t w, uint32_t h,
                 uint32_t *samples)
{
    if (!path || !acc || !w || !h)
        return MRT_E_INVALID;
    FILE *f = fopen(path, "rb");
    if (!f)
        return MRT_E_INVALID;
    int32_t hdr[3];
    size_t n = (size_t)w * h * 3;
    int ok = fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == (int32_t)w
             && hdr[1] == (int32_t)h && hdr[2] >= 0
             && fread(acc, sizeof(double), n, f) == n;
    fclose(f);
    if (!ok)
        return MRT_E_INVALID;
    if (samples)
        *samples = (uint32_t)hdr[2];
    return MRT_OK;
}
