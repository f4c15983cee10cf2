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

/*
 * Environment sampling tables for the MIS estimator.  Cells follow the
 * texel lookup of get_background_color_value (background.c:34-57): cell
 * (y, x) covers v in [y, y+1)/(h-1) and u in [x, x+1)/(w-1), i.e. the
 * directions that read texel (y, x).  Cell weight = luminance times the
 * integral of cos(elevation) over the cell's rows,
 * sin(e_hi) - sin(e_lo) with e = (v - 1/2) pi, so rows near the poles get
 * their true small solid angle.  marg has h entries (0 .. 1), cond h-1 rows
 * of w entries (0 .. 1), pcell (h-1)(w-1) cell probabilities.
 */
int mrt_hdri_build_mis(const float *rgb, uint32_t w, uint32_t h,
                       float exposure, mrt_hdri_mis *out)
{
    if (!rgb || !out || w < 2 || h < 2)
        return MRT_E_INVALID;
    memset(out, 0, sizeof(*out));
    const uint32_t nw = w - 1, nh = h - 1;
    double *wgt = malloc((size_t)nw * nh * sizeof(double));
    double *row = malloc(nh * sizeof(double));
    out->marg = malloc(h * sizeof(float));
    out->cond = malloc((size_t)nh * w * sizeof(float));
    out->pcell = malloc((size_t)nh * nw * sizeof(float));
    if (!wgt || !row || !out->marg || !out->cond || !out->pcell) {
        free(wgt);
        free(row);
        mrt_hdri_free_mis(out);
        return MRT_E_NOMEM;
    }
    out->w = w;
    out->h = h;
    const double pi = 3.14159265358979323846;
    double scale = exposure == 0.0f ? 1.0 : (double)powf(2.0f, exposure);
    double total = 0.0;
    for (uint32_t y = 0; y < nh; y++) {
        double elo = ((double)y / nh - 0.5) * pi;
        double ehi = ((double)(y + 1) / nh - 0.5) * pi;
        double band = sin(ehi) - sin(elo);
        double s = 0.0;
        for (uint32_t x = 0; x < nw; x++) {
            const float *t = rgb + ((size_t)y * w + x) * 3;
            double lum = 0.2126 * ((double)t[0] * scale)
                         + 0.7152 * ((double)t[1] * scale)
                         + 0.0722 * ((double)t[2] * scale);
            double v = lum > 0.0 ? lum * band : 0.0;
            wgt[(size_t)y * nw + x] = v;
            s += v;
        }
        row[y] = s;
        total += s;
    }
    out->total = total;
    double acc = 0.0;
    out->marg[0] = 0.0f;
    for (uint32_t y = 0; y < nh; y++) {
        acc += row[y];
        out->marg[y + 1] = total > 0.0 ? (float)(acc / total) : (float)(y + 1) / nh;
        float *c = out->cond + (size_t)y * w;
        double ra = 0.0;
        c[0] = 0.0f;
        for (uint32_t x = 0; x < nw; x++) {
            ra += wgt[(size_t)y * nw + x];
            c[x + 1] = row[y] > 0.0 ? (float)(ra / row[y]) : (float)(x + 1) / nw;
            out->pcell[(size_t)y * nw + x] =
                total > 0.0 ? (float)(wgt[(size_t)y * nw + x] / total) : 0.0f;
        }
        c[nw] = 1.0f;
    }
    out->marg[nh] = 1.0f;
    free(wgt);
    free(row);
    return MRT_OK;
}

void mrt_hdri_free_mis(mrt_hdri_mis *m)
{
    if (!m)
        return;
    free(m->marg);
    free(m->cond);
    free(m->pcell);
    memset(m, 0, sizeof(*m));
}
