/*
 * One render_frame tick (src/rendering/loop.c:36-63 -> render_gpu,
 * src/ray_tracing/render_gpu.c:35-61 -> draw_pixels, render.c:28-45) on the
 * device library: sample 0 clears the sums, every tick adds
 * samples_per_tick samples of paths into the host double buffer
 * (mrt_readback_add hands the device sums over and zeroes them).
 */
#include "mrt_host_internal.h"

#include <string.h>

int mrt_frame_tick(mrt_frame *f)
{
    if (!f || !f->ctx || !f->float_render || !f->w || !f->h)
        return MRT_E_INVALID;
    if (f->sample >= f->spp)
        return 0;
    if (f->sample == 0) {
        memset(f->float_render, 0, (size_t)f->w * f->h * 3 * sizeof(double));
        int rc = mrt_clear(f->ctx);
        if (rc != MRT_OK)
            return rc;
    }
    uint32_t n = f->samples_per_tick ? f->samples_per_tick : 1;
    if (n > f->spp - f->sample)
        n = f->spp - f->sample;
    int rc = mrt_render(f->ctx, f->sample, n, f->max_bounces);
    if (rc == MRT_OK)
        rc = mrt_readback_add(f->ctx, f->float_render);
    if (rc != MRT_OK)
        return rc;
    f->sample += n;
    return (int)n;
}
