/*
 * .rt scene reader for the elements the path tracer renders (the grammar of
 * src/parsing/parse_scene.c:99-125 and parse_elements*.c):
 *   A ratio r,g,b                       (parse_elements_utils5.c:36-53)
 *   C pos dir fov [defocus [focus]]     (parse_elements2.c:41-67)
 *   L|l pos brightness r,g,b [radius [shadow_samples]]   (:69-122)
 *   dl dir brightness r,g,b             (parse_elements5.c:16-42)
 *   sp pos diameter r,g,b|mat           (parse_elements2.c:124-149)
 *   pl pos normal r,g,b|mat             (parse_elements6.c:16-39)
 *   mat name color metallic roughness ior transmission ao emission_strength
 *       emission_color [scale]          (parse_elements3.c:59-136)
 *   W win_w win_h [render_w render_h]   (parse_elements5.c:44-83)
 *   tex name image file.hdr             (parse_elements5.c:85-142)
 *   A ratio texname                     an HDRI skybox (get_hdr_pixel_skybox)
 *   cu pos dir scale file.obj r,g,b|mat (parse_elements4.c:101-128)
 *   cy pos dir diameter height r,g,b|mat [mat_top [mat_bot]]
 *   co apex dir diameter height r,g,b|mat [mat_top]
 *   hb pos dir height a b c shape r,g,b|mat   (parse_elements3.c:15-57,
 *                                              parse_elements4.c:15-99)
 *   tex name image file.png|.bmp        image textures of materials
 *   tex name checker_local|checker_global scale c1 c2
 * Every material slot but ior and scale may name a texture.  Relative
 * file paths are tried from the working directory, then from the .rt file's
 * directory.  Numbers go through strtod; the reference's ft_atod is not
 * available, so the rounding of parsed decimals is not pinned to it.
 */
#include "mrt_host_internal.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXTOK 16

#define MAXTEX 64

typedef struct {
    char  name[64];
    char  path[1024];
    int   is_hdr;             /* an image texture in Radiance format */
    int   kind;               /* MRT_TEX_IMAGE or a checker            */
    float scale, c1[3], c2[3];
    int   scene_index;        /* texture in the scene once a material uses it */
} texdef;

typedef struct {
    mrt_host_scene *hs;
    char           *err;
    size_t          errlen;
    int             line;
    int             n_amb, n_cam, n_L, n_W;
    char            dir[1024];        /* directory of the .rt file, or "" */
    texdef          tex[MAXTEX];
    int             n_tex;
} parser;

/* path as given if it opens, else relative to the .rt file's directory */
static void resolve(const parser *P, const char *path, char *out, size_t cap)
{
    snprintf(out, cap, "%s", path);
    FILE *f = fopen(out, "rb");
    if (f) {
        fclose(f);
        return;
    }
    size_t a = strlen(P->dir), b = strlen(path);
    if (path[0] != '/' && a && a + 1 + b < cap) {
        memcpy(out, P->dir, a);
        out[a] = '/';
        memcpy(out + a + 1, path, b + 1);
    }
}

static int ends_with(const char *s, const char *suf)
{
    size_t a = strlen(s), b = strlen(suf);
    if (a < b)
        return 0;
    for (size_t i = 0; i < b; i++)
        if (tolower((unsigned char)s[a - b + i]) != suf[i])
            return 0;
    return 1;
}

static const texdef *find_tex(const parser *P, const char *name)
{
    for (int i = 0; i < P->n_tex; i++)
        if (strcmp(P->tex[i].name, name) == 0)
            return &P->tex[i];
    return NULL;
}

static int perr(parser *P, int code, const char *msg, const char *tok)
{
    if (P->err && P->errlen)
        snprintf(P->err, P->errlen, "line %d: %s%s%s", P->line, msg,
                 tok ? ": " : "", tok ? tok : "");
    return code;
}

static int num(const char *s, double *out)
{
    char *end;
    errno = 0;
    double v = strtod(s, &end);
    if (end == s || *end || errno || !isfinite(v))
        return 0;
    *out = v;
    return 1;
}

static int vec3(const char *s, double v[3])
{
    char buf[256];
    if (strlen(s) >= sizeof(buf))
        return 0;
    strcpy(buf, s);
    char *p = buf;
    for (int k = 0; k < 3; k++) {
        char *c = strchr(p, ',');
        if (k < 2) {
            if (!c)
                return 0;
            *c = 0;
        } else if (c) {
            return 0;
        }
        if (!num(p, &v[k]))
            return 0;
        p = c ? c + 1 : p;
    }
    return 1;
}

static int color(const char *s, float c[3])
{
    double v[3];
    if (!vec3(s, v))
        return 0;
    for (int k = 0; k < 3; k++) {
        if (v[k] < 0 || v[k] > 255 || v[k] != floor(v[k]))
            return 0;
        c[k] = (float)v[k];
    }
    return 1;
}

static int unit_vec(const char *s, float out[3])
{
    double v[3];
    if (!vec3(s, v))
        return 0;
    for (int k = 0; k < 3; k++)
        if (v[k] < -1.0 || v[k] > 1.0)
            return 0;
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(l > 0.0))
        return 0;
    for (int k = 0; k < 3; k++)
        out[k] = (float)(v[k] / l);
    return 1;
}

static int pos3(const char *s, float out[3])
{
    double v[3];
    if (!vec3(s, v))
        return 0;
    for (int k = 0; k < 3; k++)
        out[k] = (float)v[k];
    return 1;
}

/* A colour creates the reference's "no material" object: diffuse with
 * albedo colour / 255 (get_hit_register_color, get_hit_color.c:17-36). */
static int color_or_mat(parser *P, const char *s, uint32_t *mat)
{
    if (strchr(s, ',')) {
        float c[3];
        if (!color(s, c))
            return perr(P, MRT_E_INVALID, "bad colour", s);
        mrt_material m;
        memset(&m, 0, sizeof(m));
        for (int k = 0; k < 3; k++)
            m.color[k] = (float)(c[k] / 255.0);
        m.ao = 1.0f;
        int r = mrt_hs_add_material(P->hs, &m, NULL);
        if (r < 0)
            return r;
        *mat = (uint32_t)r;
        return MRT_OK;
    }
    int r = mrt_hs_find_material(P->hs, s);
    if (r < 0)
        return perr(P, MRT_E_INVALID, "unknown material", s);
    *mat = (uint32_t)r;
    return MRT_OK;
}

static int is_name(const char *s)
{
    for (; *s; s++)
        if (isalpha((unsigned char)*s) || *s == '_')
            return 1;
    return 0;
}

/* The scene texture of a "tex" definition, added on first use.  An image
 * that does not decode becomes the reference's fallback pattern,
 * get_solid_texture_default (get_solid_texture.c:15-28): a global checker
 * of scale 2 between black and (243, 0, 243); a material's .hdr texture
 * reads as black (get_tex_color drops get_hdr_value's result,
 * get_tex_color.c:56-57). */
static int scene_tex(parser *P, texdef *d, uint32_t *out)
{
    if (d->scene_index >= 0) {
        *out = (uint32_t)d->scene_index;
        return MRT_OK;
    }
    mrt_texture t;
    memset(&t, 0, sizeof(t));
    uint8_t *rgba = NULL;
    if (d->kind != MRT_TEX_IMAGE) {
        t.type = (uint32_t)d->kind;
        t.scale = d->scale;
        memcpy(t.c1, d->c1, sizeof(t.c1));
        memcpy(t.c2, d->c2, sizeof(t.c2));
    } else if (d->is_hdr) {
        t.type = MRT_TEX_BLACK;
    } else if (mrt_load_image(d->path, &rgba, &t.width, &t.height) == MRT_OK) {
        t.type = MRT_TEX_IMAGE;
    } else {
        t.type = MRT_TEX_CHECKER_GLOBAL;
        t.scale = 2.0f;
        t.c2[0] = 243.0f;
        t.c2[2] = 243.0f;
    }
    int r = mrt_hs_add_texture(P->hs, &t, rgba, d->name);
    mrt_free(rgba);
    if (r < 0)
        return perr(P, r, "texture rejected", d->name);
    d->scene_index = r;
    *out = (uint32_t)r;
    return MRT_OK;
}

/* A material slot: a number in [lo, hi] or a texture name
 * (parse_double_b_or_tex / parse_double_or_tex / parse_color_or_tex,
 * parse_elements_utils*.c).  is_color: an r,g,b colour instead of a number. */
static int slot(parser *P, const char *s, int is_color, double lo, double hi,
                float *val, uint32_t *tex)
{
    *tex = MRT_NO_TEX;
    double v;
    int named = is_color ? !strchr(s, ',') : (!num(s, &v) && is_name(s));
    if (named) {
        texdef *d = (texdef *)find_tex(P, s);
        if (!d)
            return perr(P, MRT_E_INVALID, "unknown texture", s);
        return scene_tex(P, d, tex);
    }
    if (is_color) {
        float c[3];
        if (!color(s, c))
            return perr(P, MRT_E_INVALID, "bad colour", s);
        for (int k = 0; k < 3; k++)
            val[k] = (float)(c[k] / 255.0);
        return MRT_OK;
    }
    if (!num(s, &v) || v < lo || v > hi)
        return perr(P, MRT_E_INVALID, "material value out of range", s);
    *val = (float)v;
    return MRT_OK;
}

/* mat name colour metallic roughness ior transmission ao emission_strength
 * emission_colour [scale [normal_map [normal_intensity]]] (parse_material,
 * parse_elements3.c:59-136); every slot but ior and scale may name a
 * texture. */
static int p_mat(parser *P, char **t, int n)
{
    if (n < 10 || n > 13)
        return perr(P, MRT_E_INVALID, "mat takes 9 to 12 fields", NULL);
    if (mrt_hs_find_material(P->hs, t[1]) >= 0)
        return perr(P, MRT_E_INVALID, "material defined twice", t[1]);
    mrt_material m;
    memset(&m, 0, sizeof(m));
    m.ao = 1.0f;
    mrt_material_tex mt;
    mt.normal = MRT_NO_TEX;
    mt.normal_intensity = 1.0f;
    mt.scale = 1.0f;
    double ior;
    int r;
    if ((r = slot(P, t[2], 1, 0, 0, m.color, &mt.color))
        || (r = slot(P, t[3], 0, 0.0, 1.0, &m.metallic, &mt.metallic))
        || (r = slot(P, t[4], 0, 0.0, 1.0, &m.roughness, &mt.roughness)))
        return r;
    if (!num(t[5], &ior) || ior < 0)
        return perr(P, MRT_E_INVALID, "bad ior", t[5]);
    m.ior = (float)ior;
    if ((r = slot(P, t[6], 0, 0.0, 1.0, &m.transmission, &mt.transmission))
        || (r = slot(P, t[7], 0, 0.0, 1.0, &m.ao, &mt.ao))
        || (r = slot(P, t[8], 0, 0.0, INFINITY, &m.emission_strength,
                     &mt.emission_strength))
        || (r = slot(P, t[9], 1, 0, 0, m.emission_color, &mt.emission_color)))
        return r;
    if (n >= 11) {
        double sc;
        if (!num(t[10], &sc))
            return perr(P, MRT_E_INVALID, "bad texture scale", t[10]);
        mt.scale = (float)sc;
    }
    if (n >= 12) {
        texdef *d = (texdef *)find_tex(P, t[11]);
        if (!d)
            return perr(P, MRT_E_INVALID, "unknown normal map", t[11]);
        if ((r = scene_tex(P, d, &mt.normal)))
            return r;
    }
    if (n >= 13) {
        double ni;
        if (!num(t[12], &ni))
            return perr(P, MRT_E_INVALID, "bad normal intensity", t[12]);
        mt.normal_intensity = (float)ni;
    }
    r = mrt_hs_add_material(P->hs, &m, t[1]);
    if (r < 0)
        return r;
    const uint32_t *sl = &mt.color;
    int textured = 0;
    for (int k = 0; k < 8; k++)
        textured |= sl[k] != MRT_NO_TEX;
    if (textured && (r = mrt_hs_set_material_tex(P->hs, (uint32_t)r, &mt)) < 0)
        return perr(P, r, "material textures", t[1]);
    return MRT_OK;
}

/* tex name image path | tex name checker_local|checker_global ... */
static int p_tex(parser *P, char **t, int n)
{
    if (n < 3)
        return perr(P, MRT_E_INVALID, "bad texture", NULL);
    if (P->n_tex == MAXTEX)
        return perr(P, MRT_E_LIMIT, "too many textures", t[1]);
    if (find_tex(P, t[1]))
        return perr(P, MRT_E_INVALID, "texture defined twice", t[1]);
    texdef *d = &P->tex[P->n_tex++];
    memset(d, 0, sizeof(*d));
    d->scene_index = -1;
    snprintf(d->name, sizeof(d->name), "%s", t[1]);
    if (strcmp(t[2], "image") == 0) {
        if (n != 4)
            return perr(P, MRT_E_INVALID, "bad image texture", t[1]);
        resolve(P, t[3], d->path, sizeof(d->path));
        d->is_hdr = ends_with(t[3], ".hdr");
        d->kind = MRT_TEX_IMAGE;
    } else if (strcmp(t[2], "checker_local") == 0
               || strcmp(t[2], "checker_global") == 0) {
        double sc;
        /* tex name checker_* scale c1 c2 (parse_texture_type,
         * parse_elements5.c:85-112) */
        if (n != 6 || !num(t[3], &sc) || sc < 0 || !color(t[4], d->c1)
            || !color(t[5], d->c2))
            return perr(P, MRT_E_INVALID, "bad checker texture", t[1]);
        d->scale = (float)sc;
        d->kind = strcmp(t[2], "checker_local") == 0 ? MRT_TEX_CHECKER_LOCAL
                                                      : MRT_TEX_CHECKER_GLOBAL;
    } else {
        return perr(P, MRT_E_INVALID, "unknown texture kind", t[2]);
    }
    return MRT_OK;
}

/* A ratio texname: the HDRI skybox.  Texels are c 2^(e-128)/255 scaled at
 * lookup by 2^(exposure + ratio - 1); a GAMMA other than 1 is applied after
 * that scale (get_hdr_pixel_skybox), so it is baked in here with an
 * exposure that makes the lookup scale 1. */
static int set_skybox(parser *P, const texdef *d, float ratio)
{
    if (!d->is_hdr)
        return perr(P, MRT_E_UNSUPPORTED, "skybox textures must be .hdr images",
                    d->name);
    float *rgb = NULL, ex = 0.0f, ga = 1.0f;
    uint32_t w = 0, h = 0;
    int r = mrt_load_hdr(d->path, &rgb, &w, &h, &ex, &ga);
    if (r < 0)
        return perr(P, r, "cannot read the HDR image", d->path);
    if (ga != 1.0f && ga > 0.0f) {
        double sc = pow(2.0, (double)ex + ratio - 1.0);
        for (size_t i = 0; i < (size_t)w * h * 3; i++)
            rgb[i] = (float)pow(rgb[i] * sc, 1.0 / ga);
        ex = 1.0f - ratio;
    }
    r = mrt_hs_set_hdri(P->hs, rgb, w, h, ex);
    mrt_free(rgb);
    return r < 0 ? perr(P, r, "HDRI", d->name) : MRT_OK;
}

/* A material given by name only (parse_only_mat,
 * parse_elements_utils5.c:55-71). */
static int only_mat(parser *P, const char *s, uint32_t *mat)
{
    int r = mrt_hs_find_material(P->hs, s);
    if (r < 0)
        return perr(P, MRT_E_INVALID, "unknown material", s);
    *mat = (uint32_t)r;
    return MRT_OK;
}

/* "cy pos dir diameter height colour|mat [mat_top] [mat_bot]",
 * "co apex dir diameter height colour|mat [mat_top]" and
 * "hb pos dir height a b c shape colour|mat" (parse_cylinder / parse_cone /
 * parse_hyperboloid, parse_elements3.c:15-57, parse_elements4.c:15-99).
 * Sizes must be > 0 here: the reference also takes 0, where its
 * intersections divide by zero. */
static int p_quadric(parser *P, char **t, int n)
{
    const char *id = t[0];
    mrt_quadric q;
    memset(&q, 0, sizeof(q));
    double v[5];
    int nsz, imat;
    if (!strcmp(id, "cy")) {
        q.kind = MRT_QUADRIC_CYLINDER;
        if (n < 6 || n > 8)
            return perr(P, MRT_E_INVALID, "cy takes 5 to 7 fields", NULL);
        nsz = 2;
        imat = 5;
    } else if (!strcmp(id, "co")) {
        q.kind = MRT_QUADRIC_CONE;
        if (n < 6 || n > 7)
            return perr(P, MRT_E_INVALID, "co takes 5 or 6 fields", NULL);
        nsz = 2;
        imat = 5;
    } else {
        q.kind = MRT_QUADRIC_HYPERBOLOID;
        if (n != 9)
            return perr(P, MRT_E_INVALID, "hb takes 8 fields", NULL);
        nsz = 5;
        imat = 8;
    }
    if (!pos3(t[1], q.pos))
        return perr(P, MRT_E_INVALID, "bad position", t[1]);
    if (!unit_vec(t[2], q.axis))
        return perr(P, MRT_E_INVALID, "bad orientation", t[2]);
    for (int k = 0; k < nsz; k++) {
        if (!num(t[3 + k], &v[k]))
            return perr(P, MRT_E_INVALID, "bad size", t[3 + k]);
        /* the hyperboloid's shape may take any sign */
        if (!(q.kind == MRT_QUADRIC_HYPERBOLOID && k == 4) && !(v[k] > 0.0))
            return perr(P, MRT_E_INVALID, "size must be > 0", t[3 + k]);
    }
    if (q.kind == MRT_QUADRIC_HYPERBOLOID) {
        q.height = (float)v[0];
        q.a = (float)v[1];
        q.b = (float)v[2];
        q.c = (float)v[3];
        q.shape = (float)v[4];
    } else {
        q.radius = (float)(v[0] * 0.5);
        q.height = (float)v[1];
    }
    int r = color_or_mat(P, t[imat], &q.mat);
    if (r)
        return r;
    q.mat_top = q.mat_bot = q.mat;
    if (q.kind != MRT_QUADRIC_HYPERBOLOID && n > 6 && (r = only_mat(P, t[6], &q.mat_top)))
        return r;
    if (q.kind == MRT_QUADRIC_CYLINDER && n > 7 && (r = only_mat(P, t[7], &q.mat_bot)))
        return r;
    r = mrt_hs_add_quadric(P->hs, &q);
    return r < 0 ? perr(P, r, "quadric rejected", NULL) : MRT_OK;
}

static int p_line(parser *P, char **t, int n, int pass)
{
    const char *id = t[0];
    if (pass == 0)
        return strcmp(id, "tex") == 0 ? p_tex(P, t, n) : MRT_OK;
    if (pass == 1)
        return strcmp(id, "mat") == 0 ? p_mat(P, t, n) : MRT_OK;
    if (strcmp(id, "mat") == 0 || strcmp(id, "tex") == 0)
        return MRT_OK;
    if (strcmp(id, "A") == 0) {
        double r;
        float c[3];
        if (n != 3 || !num(t[1], &r) || r < 0 || r > 1)
            return perr(P, MRT_E_INVALID, "bad ambient light", NULL);
        if (!strchr(t[2], ',')) {
            const texdef *d = find_tex(P, t[2]);
            if (!d)
                return perr(P, MRT_E_INVALID, "unknown texture", t[2]);
            int rc = set_skybox(P, d, (float)r);
            if (rc != MRT_OK)
                return rc;
            c[0] = c[1] = c[2] = 255.0f;
        } else if (!color(t[2], c)) {
            return perr(P, MRT_E_INVALID, "bad ambient light", NULL);
        }
        mrt_hs_set_ambient(P->hs, (float)r, c);
        P->n_amb++;
        return MRT_OK;
    }
    if (strcmp(id, "cu") == 0) {
        float pos[3], dir[3];
        double sc;
        uint32_t mat;
        if (n != 6 || !pos3(t[1], pos) || !unit_vec(t[2], dir) || !num(t[3], &sc)
            || !(sc > 0))
            return perr(P, MRT_E_INVALID, "bad custom object", NULL);
        int rc = color_or_mat(P, t[5], &mat);
        if (rc != MRT_OK)
            return rc;
        char path[1024];
        resolve(P, t[4], path, sizeof(path));
        rc = mrt_hs_add_obj(P->hs, path, pos, dir, (float)sc, mat);
        if (rc < 0)
            return perr(P, rc, "cannot read the OBJ mesh", t[4]);
        return MRT_OK;
    }
    if (strcmp(id, "C") == 0) {
        mrt_camera cam;
        memset(&cam, 0, sizeof(cam));
        double fov, a = 0, f = 0;
        if (n < 4 || n > 6 || !pos3(t[1], cam.pos) || !unit_vec(t[2], cam.dir)
            || !num(t[3], &fov) || fov < 0 || fov > 180 || fov != floor(fov)
            || (n > 4 && (!num(t[4], &a) || a < 0))
            || (n > 5 && (!num(t[5], &f) || f < 0)))
            return perr(P, MRT_E_INVALID, "bad camera", NULL);
        cam.fov_deg = (float)fov;
        cam.defocus_angle_deg = (float)a;
        cam.focus_dist = f > 0 ? (float)f : 1.0f;
        mrt_hs_set_camera(P->hs, &cam);
        P->n_cam++;
        return MRT_OK;
    }
    if (strcmp(id, "L") == 0 || strcmp(id, "l") == 0) {
        mrt_light L;
        memset(&L, 0, sizeof(L));
        double b, rad = 0, ss = 0;
        L.type = MRT_LIGHT_POINT;
        if (n < 4 || n > 6 || !pos3(t[1], L.vec) || !num(t[2], &b) || b < 0
            || b > 1 || !color(t[3], L.color)
            || (n > 4 && (!num(t[4], &rad) || rad < 0))
            || (n > 5 && (!num(t[5], &ss) || ss < 0 || ss != floor(ss))))
            return perr(P, MRT_E_INVALID, "bad light", NULL);
        L.brightness = (float)b;
        L.radius = (float)rad;
        L.shadow_samples = n > 4 ? (n > 5 ? (int32_t)ss : 5) : 0;
        if (id[0] == 'L')
            P->n_L++;
        int r = mrt_hs_add_light(P->hs, &L);
        return r < 0 ? perr(P, r, "too many lights", NULL) : MRT_OK;
    }
    if (strcmp(id, "dl") == 0) {
        mrt_light L;
        memset(&L, 0, sizeof(L));
        double b;
        L.type = MRT_LIGHT_DIRECTIONAL;
        if (n != 4 || !unit_vec(t[1], L.vec) || !num(t[2], &b) || b < 0 || b > 1
            || !color(t[3], L.color))
            return perr(P, MRT_E_INVALID, "bad directional light", NULL);
        L.brightness = (float)b;
        int r = mrt_hs_add_light(P->hs, &L);
        return r < 0 ? perr(P, r, "too many lights", NULL) : MRT_OK;
    }
    if (strcmp(id, "sp") == 0) {
        float c[3];
        double d;
        uint32_t mat;
        if (n != 4 || !pos3(t[1], c) || !num(t[2], &d) || !(d > 0))
            return perr(P, MRT_E_INVALID, "bad sphere", NULL);
        int r = color_or_mat(P, t[3], &mat);
        if (r)
            return r;
        r = mrt_hs_add_sphere(P->hs, c, (float)(d * 0.5), mat);
        return r < 0 ? perr(P, r, "sphere rejected", NULL) : MRT_OK;
    }
    if (strcmp(id, "pl") == 0) {
        float p[3], nn[3];
        uint32_t mat;
        double v[3];
        if (n != 4 || !pos3(t[1], p) || !vec3(t[2], v))
            return perr(P, MRT_E_INVALID, "bad plane", NULL);
        /* the normal is kept as written (parse_elements6.c:16-39) */
        for (int k = 0; k < 3; k++) {
            if (v[k] < -1.0 || v[k] > 1.0)
                return perr(P, MRT_E_INVALID, "bad plane normal", t[2]);
            nn[k] = (float)v[k];
        }
        int r = color_or_mat(P, t[3], &mat);
        if (r)
            return r;
        r = mrt_hs_add_plane(P->hs, p, nn, mat);
        return r < 0 ? perr(P, r, "plane rejected", NULL) : MRT_OK;
    }
    if (strcmp(id, "W") == 0) {
        double v[4];
        if ((n != 3 && n != 5))
            return perr(P, MRT_E_INVALID, "bad window", NULL);
        for (int k = 1; k < n; k++)
            if (!num(t[k], &v[k - 1]) || v[k - 1] < 1 || v[k - 1] > 16384
                || v[k - 1] != floor(v[k - 1]))
                return perr(P, MRT_E_INVALID, "bad window size", t[k]);
        if (n == 5)
            mrt_hs_set_resolution(P->hs, (uint32_t)v[2], (uint32_t)v[3]);
        else
            mrt_hs_set_resolution(P->hs, (uint32_t)v[0], (uint32_t)v[1]);
        P->n_W++;
        return MRT_OK;
    }
    if (!strcmp(id, "cy") || !strcmp(id, "co") || !strcmp(id, "hb"))
        return p_quadric(P, t, n);
    return perr(P, MRT_E_INVALID, "unknown element", id);
}

static int parse_text(const char *text, const char *dir, mrt_host_scene **out,
                      char *err, size_t errlen)
{
    if (!text || !out)
        return MRT_E_INVALID;
    *out = NULL;
    parser P;
    memset(&P, 0, sizeof(P));
    if (dir)
        snprintf(P.dir, sizeof(P.dir), "%s", dir);
    P.err = err;
    P.errlen = errlen;
    if (err && errlen)
        err[0] = 0;
    P.hs = mrt_hs_new();
    if (!P.hs)
        return MRT_E_NOMEM;
    size_t len = strlen(text);
    char *buf = malloc(len + 1);
    if (!buf) {
        mrt_hs_free(P.hs);
        return MRT_E_NOMEM;
    }
    int rc = MRT_OK;
    for (int pass = 0; pass < 3 && rc == MRT_OK; pass++) {
        memcpy(buf, text, len + 1);
        P.line = 0;
        char *save = NULL;
        for (char *ln = buf; ln && rc == MRT_OK; ln = save) {
            save = strchr(ln, '\n');
            if (save)
                *save++ = 0;
            P.line++;
            char *hash = strchr(ln, '#');
            if (hash)
                *hash = 0;
            char *t[MAXTOK];
            int n = 0;
            for (char *p = strtok(ln, " \t\r"); p; p = strtok(NULL, " \t\r")) {
                if (n == MAXTOK) {
                    rc = perr(&P, MRT_E_INVALID, "too many fields", NULL);
                    break;
                }
                t[n++] = p;
            }
            if (rc == MRT_OK && n > 0)
                rc = p_line(&P, t, n, pass);
        }
    }
    free(buf);
    if (rc == MRT_OK) {
        if (P.n_amb != 1)
            rc = perr(&P, MRT_E_INVALID, "exactly one A is required", NULL);
        else if (P.n_cam != 1)
            rc = perr(&P, MRT_E_INVALID, "exactly one C is required", NULL);
        else if (P.n_L > 1 || P.n_W > 1)
            rc = perr(&P, MRT_E_INVALID, "at most one L and one W", NULL);
    }
    if (rc != MRT_OK) {
        mrt_hs_free(P.hs);
        return rc;
    }
    *out = P.hs;
    return MRT_OK;
}

int mrt_parse_rt_string(const char *text, mrt_host_scene **out, char *err,
                        size_t errlen)
{
    return parse_text(text, NULL, out, err, errlen);
}

int mrt_parse_rt_file(const char *path, mrt_host_scene **out, char *err,
                      size_t errlen)
{
    if (!path || !out)
        return MRT_E_INVALID;
    FILE *f = fopen(path, "rb");
    if (!f) {
        if (err && errlen)
            snprintf(err, errlen, "cannot open %s", path);
        return MRT_E_INVALID;
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0 || sz > (64L << 20)) {
        fclose(f);
        return MRT_E_LIMIT;
    }
    char *text = malloc((size_t)sz + 1);
    if (!text) {
        fclose(f);
        return MRT_E_NOMEM;
    }
    size_t rd = fread(text, 1, (size_t)sz, f);
    fclose(f);
    text[rd] = 0;
    char dir[1024];
    snprintf(dir, sizeof(dir), "%s", path);
    char *slash = strrchr(dir, '/');
    if (slash)
        *slash = 0;
    else
        dir[0] = 0;
    int rc = parse_text(text, dir, out, err, errlen);
    free(text);
    return rc;
}
