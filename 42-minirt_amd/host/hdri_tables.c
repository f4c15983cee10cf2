/*
 * HDRI importance tables on the host, as make_importance_sampling_map
 * builds them (src/ray_tracing/importance_sampling/importance_sampling_manager.c:41-61,
 * calc_pdf.c:16-114, calc_cdf.c:15-122): luminance, row averages, the joint
 * pdf and the quantised inverse marginal / conditional CDFs.
 *
 * The reference searches every inverse-CDF entry from index 0, O(W^2 H).
 * Its targets i/(n-1) increase with i and the CDF never decreases, so the
 * first index reaching target i is never below the one reaching target i-1:
 * a single forward scan per row yields identical tables in O(W H).
 */
#include "mrt_host_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void inverse_row(const double *cdf, uint32_t n, float *inv)
{
    uint32_t x = 0;
    for (uint32_t i = 0; i < n; i++) {
        double u = (double)i / (double)(n - 1);
        while (x < n && !(cdf[x] >= u))
            x++;
        if (x >= n)
            break;                 /* unreached entries stay 0 (ft_calloc) */
        inv[i] = (float)((double)x / (double)(n - 1));
    }
}

int mrt_hdri_build_tables(const float *rgb, uint32_t w, uint32_t h,
                          float exposure, mrt_hdri_tables *out)
{
    if (!rgb || !out || w < 2 || h < 2)
        return MRT_E_INVALID;
    memset(out, 0, sizeof(*out));
    size_t n = (size_t)w * h;
    double *gray = malloc(n * sizeof(double));
    double *rowavg = malloc(h * sizeof(double));
    double *cdf = malloc((w > h ? w : h) * sizeof(double));
    out->pdf_joint = malloc(n * sizeof(float));
    out->cond_inv = calloc(n, sizeof(float));
    out->marg_inv = calloc(h, sizeof(float));
    if (!gray || !rowavg || !cdf || !out->pdf_joint || !out->cond_inv
        || !out->marg_inv) {
        free(gray);
        free(rowavg);
        free(cdf);
        mrt_hdri_free_tables(out);
        return MRT_E_NOMEM;
    }
    out->w = w;
    out->h = h;
    /* get_grey_value (calc_pdf.c:16-35): scale = 2^exposure when nonzero */
    double scale = exposure == 0.0f ? 1.0 : (double)powf(2.0f, exposure);
    for (size_t i = 0; i < n; i++)
        gray[i] = 0.2126 * ((double)rgb[3 * i] * scale)
                  + 0.7152 * ((double)rgb[3 * i + 1] * scale)
                  + 0.0722 * ((double)rgb[3 * i + 2] * scale);
    /* calc_raw_average / calc_column_average */
    double colavg = 0.0;
    for (uint32_t y = 0; y < h; y++) {
        double s = 0.0;
        for (uint32_t x = 0; x < w; x++)
            s += gray[(size_t)y * w + x];
        rowavg[y] = s / w;
    }
    for (uint32_t y = 0; y < h; y++)
        colavg += rowavg[y];
    colavg = colavg / h;
    out->column_average = colavg;
    /* calc_pdfs */
    for (size_t i = 0; i < n; i++)
        out->pdf_joint[i] = (float)(gray[i] / colavg);
    /* build_and_normalize_marginal_cdf + calc_cdf_marginal_inv */
    double sum = 0.0;
    for (uint32_t y = 0; y < h; y++) {
        sum += rowavg[y] / (colavg / 255);
        cdf[y] = sum;
    }
    if (sum > 0.0)
        for (uint32_t y = 0; y < h; y++)
            cdf[y] /= sum;
    inverse_row(cdf, h, out->marg_inv);
    /* process_conditional_row + set_cdf_conditional_row per row */
    for (uint32_t y = 0; y < h; y++) {
        sum = 0.0;
        for (uint32_t x = 0; x < w; x++) {
            sum += gray[(size_t)y * w + x] / (rowavg[y] / 255);
            cdf[x] = sum;
        }
        if (sum > 0.0)
            for (uint32_t x = 0; x < w; x++)
                cdf[x] /= sum;
        inverse_row(cdf, w, out->cond_inv + (size_t)y * w);
    }
    free(gray);
    free(rowavg);
    free(cdf);
    return MRT_OK;
}

void mrt_hdri_free_tables(mrt_hdri_tables *t)
{
    if (!t)
        return;
    free(t->pdf_joint);
    free(t->marg_inv);
    free(t->cond_inv);
    memset(t, 0, sizeof(*t));
}
