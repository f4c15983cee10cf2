/*
 * minirt_hip: headless driver of the device library, the counterpart of the
 * reference's main_gpu path (src/main.c:17-30, src/main_shader.c:78-93) with
 * --no-display: parse the scene, then call render_frame's GPU tick until
 * --max-samples, saving render_state.rts every 10 samples with --auto-save
 * and a PPM with --auto-export (src/rendering/loop.c:36-63,
 * src/ray_tracing/render.c:47-84).
 *
 *   minirt_hip scene.rt [--max-samples=N] [--max-bounces=N]
 *       [--samples-per-tick=N] [--seed=S] [--output-dir=DIR/] [--auto-export]
 *       [--auto-save] [--load-render=FILE] [--triangle-one-sided]
 *       [--device=N]
 *   minirt_hip --procedural=plain|mesh ...
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mrt.h"
#include "mrt_host.h"

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

static const char *opt_val(const char *arg, const char *name)
{
    size_t n = strlen(name);
    if (strncmp(arg, name, n) == 0 && arg[n] == '=')
        return arg + n + 1;
    return NULL;
}

int main(int argc, char **argv)
{
    const char *scene_path = NULL, *procedural = NULL, *outdir = "";
    const char *load = NULL, *v;
    uint32_t spp = 64, bounces = 10, per_tick = 1;
    uint64_t seed = 42;
    int auto_export = 0, auto_save = 0, one_sided = 0, device = 0, mis = 0;
    const char *caustic = NULL;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if ((v = opt_val(a, "--max-samples")))
            spp = (uint32_t)strtoul(v, NULL, 10);
        else if ((v = opt_val(a, "--max-bounces")))
            bounces = (uint32_t)strtoul(v, NULL, 10);
        else if ((v = opt_val(a, "--samples-per-tick")))
            per_tick = (uint32_t)strtoul(v, NULL, 10);
        else if ((v = opt_val(a, "--seed")))
            seed = strtoull(v, NULL, 0);
        else if ((v = opt_val(a, "--output-dir")))
            outdir = v;
        else if ((v = opt_val(a, "--load-render")))
            load = v;
        else if ((v = opt_val(a, "--procedural")))
            procedural = v;
        else if ((v = opt_val(a, "--device")))
            device = atoi(v);
        else if (!strcmp(a, "--auto-export"))
            auto_export = 1;
        else if (!strcmp(a, "--auto-save"))
            auto_save = 1;
        else if (!strcmp(a, "--triangle-one-sided"))
            one_sided = 1;
        else if ((v = opt_val(a, "--caustic")))
            caustic = v;             /* N;RADIUS;NORM (option_caustic.c) */
        else if (!strcmp(a, "--mis"))
            mis = 1;
        else if (!strcmp(a, "--no-display") || !strcmp(a, "--compute-gpu"))
            ;
        else if (a[0] != '-' && !scene_path)
            scene_path = a;
        else {
            fprintf(stderr, "Error\nunknown option %s\n", a);
            return 1;
        }
    }
    mrt_host_scene *hs = NULL;
    char err[512];
    int rc;
    if (procedural && !strcmp(procedural, "plain"))
        rc = mrt_gen_plain_scene(seed, 640, 360, &hs);
    else if (procedural && !strcmp(procedural, "mesh"))
        rc = mrt_gen_mesh_scene(seed, 1920, 1080, 8, 256, 0.15f, &hs);
    else if (procedural && !strcmp(procedural, "caustic"))
        rc = mrt_gen_caustic_scene(seed, 1920, 1080, 8, 256, 512, &hs);
    else if (scene_path)
        rc = mrt_parse_rt_file(scene_path, &hs, err, sizeof(err));
    else {
        fprintf(stderr, "Error\nusage: minirt_hip scene.rt [options]\n");
        return 1;
    }
    if (rc != MRT_OK) {
        fprintf(stderr, "Error\n%s\n", scene_path ? err : "scene generation failed");
        return 1;
    }
    if (one_sided)
        mrt_hs_set_flags(hs, mrt_hs_view(hs)->flags | MRT_SCENE_TRIANGLE_ONE_SIDED);
    uint32_t w, h;
    mrt_hs_get_resolution(hs, &w, &h);
    mrt_ctx *ctx = NULL;
    if ((rc = mrt_init(device, w, h, &ctx)) != MRT_OK) {
        fprintf(stderr, "Error\nmrt_init failed (%d)\n", rc);
        mrt_hs_free(hs);
        return 1;
    }
    mrt_set_seed(ctx, (uint64_t)seed);
    if ((rc = mrt_upload_scene(ctx, mrt_hs_view(hs))) != MRT_OK) {
        fprintf(stderr, "Error\n%s\n", mrt_last_error(ctx));
        mrt_destroy(ctx);
        mrt_hs_free(hs);
        return 1;
    }
    const float *sky = NULL;
    uint32_t sw = 0, sh = 0;
    float sex = 0.0f;
    if (mrt_hs_get_hdri(hs, &sky, &sw, &sh, &sex) == MRT_OK
        && (rc = mrt_upload_hdri(ctx, sky, sw, sh, sex)) != MRT_OK) {
        fprintf(stderr, "Error\n%s\n", mrt_last_error(ctx));
        mrt_destroy(ctx);
        mrt_hs_free(hs);
        return 1;
    }
    if (mis)
        mrt_set_option(ctx, MRT_OPT_ESTIMATOR, MRT_EST_MIS);
    if (caustic) {
        unsigned long np = 0;
        double rad = 0.0, norm = 1.0;
        if (sscanf(caustic, "%lu;%lf;%lf", &np, &rad, &norm) < 2
            || mrt_set_option(ctx, MRT_OPT_PHOTON_NORM, norm) != MRT_OK
            || mrt_build_photons(ctx, (uint32_t)np, (float)rad) != MRT_OK) {
            fprintf(stderr, "Warning\ncaustics will be disabled: %s\n",
                    mrt_last_error(ctx));
        }
    }
    double *acc = calloc((size_t)w * h * 3, sizeof(double));
    mrt_frame f = {ctx, w, h, 0, spp, bounces, per_tick, acc};
    if (!acc) {
        rc = MRT_E_NOMEM;
        goto done;
    }
    if (load) {
        uint32_t s = 0;
        if (mrt_load_rts(load, acc, w, h, &s) == MRT_OK && s < spp) {
            /* resume: the counter RNG continues at sample s exactly */
            f.sample = s;
        } else {
            fprintf(stderr, "Warning\ncannot resume from %s\n", load);
        }
    }
    double t0 = now_ms();
    while (f.sample < spp) {
        double ts = now_ms();
        int n = mrt_frame_tick(&f);
        if (n < 0) {
            fprintf(stderr, "Error\n%s\n", mrt_last_error(ctx));
            rc = n;
            goto done;
        }
        printf("Sample %u - %.0fms\n", f.sample, now_ms() - ts);
        if (auto_save && f.sample % 10 == 0) {
            char path[1024];
            snprintf(path, sizeof(path), "%srender_state.rts", outdir);
            mrt_export_rts(path, acc, w, h, f.sample);
        }
    }
    printf("%u samples in %.1f ms\n", spp, now_ms() - t0);
    if (auto_export) {
        char path[1024];
        snprintf(path, sizeof(path), "%sminirt_export.SAMPLES.%u.ppm", outdir,
                 f.sample);
        rc = mrt_export_ppm(path, acc, w, h, f.sample, 1.0);
        printf("exported %s\n", path);
    }
done:
    free(acc);
    mrt_destroy(ctx);
    mrt_hs_free(hs);
    return rc == MRT_OK ? 0 : 1;
}
