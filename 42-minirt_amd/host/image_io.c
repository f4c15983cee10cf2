/*
 * Output formats of the accumulation buffer: the P6 PPM export
 * (src/utils/ppm/export_ppm.c:42-87 with render_float_to_int_gamma,
 * src/upscaling/utils.c:15-35) and the render_state.rts checkpoint
 * (src/utils/export_render_state.c:15-51).
 */
#include "mrt_host_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

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
