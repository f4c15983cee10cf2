/*
 * Wavefront OBJ meshes for the .rt "cu pos dir scale file.obj colour|mat"
 * element (src/parsing/parse_elements4.c:101-128, the reference's OBJ reader in src/utils/obj):
 * "v x y z" positions, "vn x y z" normals, "vt u v" texture coordinates,
 * "f a[/t][/n] b c ..." faces with 1-based or negative indices; other lines
 * are ignored.  A triangle whose three corners name texture coordinates
 * carries them (the (u, v) hit_triangle interpolates, hit_triangle.c:91-92).  A face of n > 3
 * vertices becomes the triangles the reference's ear clipping emits
 * (obj_ear_clipping.c:57-95): (v[n-1], v0, v1), then v0 is dropped and the
 * same is repeated on the remaining polygon.
 *
 * Placement as apply_transformations (obj_finilizer.c:18-46): position =
 * rotate(scale * p) + pos, the rotation taking +Z to `dir` about
 * cross(+Z, dir) (by pi about +X when dir = -Z).  Normals are rotated with
 * the mesh here; the reference leaves them in model space.
 */
#include "mrt_host_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double axis[3], c, s;
    int    rotate;
} rot_t;

static void rot_init(rot_t *R, const float dir[3])
{
    double d[3] = {dir[0], dir[1], dir[2]};
    double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    memset(R, 0, sizeof(*R));
    if (!(l > 0))
        return;
    for (int k = 0; k < 3; k++)
        d[k] /= l;
    double ax[3] = {-d[1], d[0], 0.0};       /* cross((0,0,1), d) */
    double al = sqrt(ax[0] * ax[0] + ax[1] * ax[1]);
    double ang = acos(d[2] > 1 ? 1 : (d[2] < -1 ? -1 : d[2]));
    if (al < 1e-4) {
        if (ang > M_PI - 0.01) {
            R->axis[0] = 1.0;
            R->c = -1.0;
            R->s = 0.0;
            R->rotate = 1;
        }
        return;
    }
    for (int k = 0; k < 3; k++)
        R->axis[k] = ax[k] / al;
    R->c = cos(ang);
    R->s = sin(ang);
    R->rotate = 1;
}

/* Rodrigues' rotation of v about the unit axis. */
static void rot_apply(const rot_t *R, const double v[3], double o[3])
{
    if (!R->rotate) {
        memcpy(o, v, 3 * sizeof(double));
        return;
    }
    const double *k = R->axis;
    double kv = k[0] * v[0] + k[1] * v[1] + k[2] * v[2];
    double kx[3] = {k[1] * v[2] - k[2] * v[1], k[2] * v[0] - k[0] * v[2],
                    k[0] * v[1] - k[1] * v[0]};
    for (int i = 0; i < 3; i++)
        o[i] = v[i] * R->c + kx[i] * R->s + k[i] * kv * (1.0 - R->c);
}

typedef struct {
    float   *v;
    size_t   n, cap;
} fvec;

static int fpush3(fvec *a, double x, double y, double z)
{
    if (a->n + 3 > a->cap) {
        size_t nc = a->cap ? a->cap * 2 : 3072;
        float *p = realloc(a->v, nc * sizeof(float));
        if (!p)
            return MRT_E_NOMEM;
        a->v = p;
        a->cap = nc;
    }
    a->v[a->n++] = (float)x;
    a->v[a->n++] = (float)y;
    a->v[a->n++] = (float)z;
    return MRT_OK;
}

/* index of an OBJ reference (1-based, or negative from the end) */
static long obj_index(long i, size_t count)
{
    if (i > 0)
        return i - 1 < (long)count ? i - 1 : -1;
    if (i < 0)
        return (long)count + i >= 0 ? (long)count + i : -1;
    return -1;
}

int mrt_hs_add_obj_text(mrt_host_scene *hs, const char *text, const float pos[3],
                        const float dir[3], float scale, uint32_t mat)
{
    if (!hs || !text || !pos || !dir)
        return MRT_E_INVALID;
    rot_t R;
    rot_init(&R, dir);
    fvec P = {0}, N = {0}, TC = {0};
    int rc = MRT_OK, added = 0;
    const char *line = text;
    enum { MAXV = 256 };
    long fv[MAXV], fn[MAXV], ft[MAXV];
    while (*line && rc == MRT_OK) {
        const char *end = strchr(line, '\n');
        size_t len = end ? (size_t)(end - line) : strlen(line);
        char buf[4096];
        if (len >= sizeof(buf)) {
            rc = MRT_E_LIMIT;
            break;
        }
        memcpy(buf, line, len);
        buf[len] = 0;
        line = end ? end + 1 : line + len;
        char *s = buf;
        while (*s == ' ' || *s == '\t')
            s++;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            double a, b, c;
            if (sscanf(s + 2, "%lf %lf %lf", &a, &b, &c) != 3) {
                rc = MRT_E_INVALID;
                break;
            }
            double v[3] = {a * scale, b * scale, c * scale}, o[3];
            rot_apply(&R, v, o);
            rc = fpush3(&P, o[0] + pos[0], o[1] + pos[1], o[2] + pos[2]);
        } else if (s[0] == 'v' && s[1] == 'n' && (s[2] == ' ' || s[2] == '\t')) {
            double a, b, c;
            if (sscanf(s + 3, "%lf %lf %lf", &a, &b, &c) != 3) {
                rc = MRT_E_INVALID;
                break;
            }
            double v[3] = {a, b, c}, o[3];
            rot_apply(&R, v, o);
            rc = fpush3(&N, o[0], o[1], o[2]);
        } else if (s[0] == 'v' && s[1] == 't' && (s[2] == ' ' || s[2] == '\t')) {
            double a, b = 0.0;
            int got = sscanf(s + 3, "%lf %lf", &a, &b);
            if (got < 1) {
                rc = MRT_E_INVALID;
                break;
            }
            rc = fpush3(&TC, a, b, 0.0);
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            int nv = 0;
            char *p = s + 2;
            while (*p && rc == MRT_OK) {
                while (*p == ' ' || *p == '\t' || *p == '\r')
                    p++;
                if (!*p)
                    break;
                if (nv == MAXV) {
                    rc = MRT_E_LIMIT;
                    break;
                }
                char *q;
                long vi = strtol(p, &q, 10), ni = 0, ti = 0;
                if (q == p) {
                    rc = MRT_E_INVALID;
                    break;
                }
                p = q;
                if (*p == '/') {
                    p++;
                    ti = strtol(p, &q, 10);           /* texture index */
                    p = q;
                    if (*p == '/') {
                        p++;
                        ni = strtol(p, &q, 10);
                        p = q;
                    }
                }
                while (*p && *p != ' ' && *p != '\t' && *p != '\r')
                    p++;
                fv[nv] = obj_index(vi, P.n / 3);
                fn[nv] = ni ? obj_index(ni, N.n / 3) : -1;
                ft[nv] = ti ? obj_index(ti, TC.n / 3) : -1;
                if (fv[nv] < 0) {
                    rc = MRT_E_INVALID;
                    break;
                }
                nv++;
            }
            if (rc != MRT_OK)
                break;
            if (nv < 3)
                continue;
            /* ear clipping at vertex 0 of the shrinking polygon */
            long pv[MAXV], pn[MAXV], pt[MAXV];
            memcpy(pv, fv, nv * sizeof(long));
            memcpy(pn, fn, nv * sizeof(long));
            memcpy(pt, ft, nv * sizeof(long));
            int m = nv;
            while (m >= 3 && rc == MRT_OK) {
                long tri[3] = {m > 3 ? pv[m - 1] : pv[0], m > 3 ? pv[0] : pv[1],
                               m > 3 ? pv[1] : pv[2]};
                long trn[3] = {m > 3 ? pn[m - 1] : pn[0], m > 3 ? pn[0] : pn[1],
                               m > 3 ? pn[1] : pn[2]};
                long trt[3] = {m > 3 ? pt[m - 1] : pt[0], m > 3 ? pt[0] : pt[1],
                               m > 3 ? pt[1] : pt[2]};
                float v9[9], n9[9];
                int have_n = 1;
                for (int k = 0; k < 3; k++) {
                    memcpy(v9 + 3 * k, P.v + 3 * tri[k], 3 * sizeof(float));
                    if (trn[k] < 0)
                        have_n = 0;
                    else
                        memcpy(n9 + 3 * k, N.v + 3 * trn[k], 3 * sizeof(float));
                }
                int r = mrt_hs_add_triangle(hs, v9, have_n ? n9 : NULL, mat);
                if (r >= 0 && trt[0] >= 0 && trt[1] >= 0 && trt[2] >= 0) {
                    float uv[6];
                    for (int k = 0; k < 3; k++) {
                        uv[2 * k] = TC.v[3 * trt[k]];
                        uv[2 * k + 1] = TC.v[3 * trt[k] + 1];
                    }
                    int e = mrt_hs_set_triangle_uv(hs, (uint32_t)r, uv);
                    r = e < 0 ? e : r;
                }
                if (r < 0)
                    rc = r;
                else
                    added++;
                if (m == 3)
                    break;
                memmove(pv, pv + 1, (m - 1) * sizeof(long));
                memmove(pn, pn + 1, (m - 1) * sizeof(long));
                memmove(pt, pt + 1, (m - 1) * sizeof(long));
                m--;
            }
        }
    }
    free(P.v);
    free(N.v);
    free(TC.v);
    return rc < 0 ? rc : added;
}

int mrt_hs_add_obj(mrt_host_scene *hs, const char *path, const float pos[3],
                   const float dir[3], float scale, uint32_t mat)
{
    if (!path)
        return MRT_E_INVALID;
    FILE *f = fopen(path, "rb");
    if (!f)
        return MRT_E_INVALID;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0 || sz > (1L << 30)) {
        fclose(f);
        return MRT_E_LIMIT;
    }
    char *text = malloc((size_t)sz + 1);
    if (!text) {
        fclose(f);
        return MRT_E_NOMEM;
    }
    size_t got = fread(text, 1, (size_t)sz, f);
    fclose(f);
    text[got] = 0;
    int rc = mrt_hs_add_obj_text(hs, text, pos, dir, scale, mat);
    free(text);
    return rc;
}
