/*
 * Host scene container: growable arrays behind an mrt_scene view.
 * Plays the role of t_scene's element list (includes/scene_structs.h:362-383)
 * for the procedural generators and the .rt reader.
 */
#include "mrt_host_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int grow(void **p, size_t *cap, size_t need, size_t elem)
{
    if (need <= *cap)
        return 0;
    size_t ncap = *cap ? *cap : 16;
    while (ncap < need)
        ncap *= 2;
    void *np = realloc(*p, ncap * elem);
    if (!np)
        return MRT_E_NOMEM;
    *p = np;
    *cap = ncap;
    return 0;
}

mrt_host_scene *mrt_hs_new(void)
{
    mrt_host_scene *hs = calloc(1, sizeof(*hs));
    if (!hs)
        return NULL;
    hs->camera.dir[2] = 1.0f;
    hs->camera.fov_deg = 60.0f;
    hs->camera.focus_dist = 1.0f;
    hs->ior_all = 1.0f;
    hs->render_w = 640;
    hs->render_h = 360;
    return hs;
}

void mrt_hs_free(mrt_host_scene *hs)
{
    if (!hs)
        return;
    free(hs->tri_pos);
    free(hs->tri_nrm);
    free(hs->tri_mat);
    free(hs->sph);
    free(hs->sph_mat);
    free(hs->quad);
    free(hs->tex);
    free(hs->tex_names);
    free(hs->texels);
    free(hs->mtex);
    free(hs->tri_uv);
    free(hs->pln);
    free(hs->pln_mat);
    free(hs->materials);
    free(hs->mat_names);
    free(hs->lights);
    free(hs->hdri_rgb);
    free(hs);
}

static void mtex_default(mrt_material_tex *t)
{
    t->color = t->metallic = t->roughness = t->transmission = t->ao = MRT_NO_TEX;
    t->emission_strength = t->emission_color = t->normal = MRT_NO_TEX;
    t->normal_intensity = 1.0f;
    t->scale = 1.0f;
}

int mrt_hs_add_texture(mrt_host_scene *hs, const mrt_texture *t,
                       const uint8_t *rgba, const char *name)
{
    if (!hs || !t || t->type > MRT_TEX_BLACK)
        return MRT_E_INVALID;
    mrt_texture c = *t;
    c.texel0 = 0;
    c.pad = 0;
    if (c.type == MRT_TEX_IMAGE) {
        uint64_t n = (uint64_t)c.width * c.height;
        if (!rgba || c.width == 0 || c.height == 0)
            return MRT_E_INVALID;
        if ((uint64_t)hs->n_texels + n >= (1ull << 31))
            return MRT_E_LIMIT;
        if (grow((void **)&hs->texels, &hs->cap_texels,
                 (size_t)(hs->n_texels + n) * 4, 1))
            return MRT_E_NOMEM;
        memcpy(hs->texels + (size_t)hs->n_texels * 4, rgba, (size_t)n * 4);
        c.texel0 = hs->n_texels;
        hs->n_texels += (uint32_t)n;
    } else {
        c.width = c.height = 0;
    }
    if (grow((void **)&hs->tex, &hs->cap_tex, hs->n_tex + 1, sizeof(mrt_texture))
        || grow((void **)&hs->tex_names, &hs->cap_tex_names, hs->n_tex + 1,
                sizeof(*hs->tex_names)))
        return MRT_E_NOMEM;
    hs->tex[hs->n_tex] = c;
    memset(hs->tex_names[hs->n_tex], 0, sizeof(*hs->tex_names));
    if (name)
        strncpy(hs->tex_names[hs->n_tex], name, sizeof(*hs->tex_names) - 1);
    return (int)hs->n_tex++;
}

int mrt_hs_find_texture(const mrt_host_scene *hs, const char *name)
{
    if (!hs || !name)
        return MRT_E_INVALID;
    for (uint32_t i = 0; i < hs->n_tex; i++)
        if (strcmp(hs->tex_names[i], name) == 0)
            return (int)i;
    return MRT_E_INVALID;
}

int mrt_hs_set_material_tex(mrt_host_scene *hs, uint32_t mat,
                            const mrt_material_tex *mt)
{
    if (!hs || !mt || mat >= hs->n_mat)
        return MRT_E_INVALID;
    const uint32_t *slot = &mt->color;
    for (int k = 0; k < 8; k++)
        if (slot[k] != MRT_NO_TEX && slot[k] >= hs->n_tex)
            return MRT_E_INVALID;
    if (!hs->mtex) {
        if (grow((void **)&hs->mtex, &hs->cap_mtex, hs->n_mat,
                 sizeof(mrt_material_tex)))
            return MRT_E_NOMEM;
        for (uint32_t i = 0; i < hs->n_mat; i++)
            mtex_default(&hs->mtex[i]);
    }
    hs->mtex[mat] = *mt;
    return MRT_OK;
}

int mrt_hs_set_triangle_uv(mrt_host_scene *hs, uint32_t tri, const float uv[6])
{
    if (!hs || !uv || tri >= hs->n_tri)
        return MRT_E_INVALID;
    if (!hs->tri_uv) {
        if (grow((void **)&hs->tri_uv, &hs->cap_tri_uv, (size_t)hs->n_tri * 6,
                 sizeof(float)))
            return MRT_E_NOMEM;
        memset(hs->tri_uv, 0, (size_t)hs->n_tri * 6 * sizeof(float));
    }
    memcpy(hs->tri_uv + (size_t)tri * 6, uv, 6 * sizeof(float));
    return MRT_OK;
}

int mrt_hs_add_material(mrt_host_scene *hs, const mrt_material *m,
                        const char *name)
{
    if (!hs || !m)
        return MRT_E_INVALID;
    if (grow((void **)&hs->materials, &hs->cap_mat, hs->n_mat + 1,
             sizeof(mrt_material))
        || grow((void **)&hs->mat_names, &hs->cap_mat_names, hs->n_mat + 1,
                sizeof(*hs->mat_names)))
        return MRT_E_NOMEM;
    if (hs->mtex) {
        if (grow((void **)&hs->mtex, &hs->cap_mtex, hs->n_mat + 1,
                 sizeof(mrt_material_tex)))
            return MRT_E_NOMEM;
        mtex_default(&hs->mtex[hs->n_mat]);
    }
    hs->materials[hs->n_mat] = *m;
    memset(hs->mat_names[hs->n_mat], 0, sizeof(*hs->mat_names));
    if (name)
        strncpy(hs->mat_names[hs->n_mat], name, sizeof(*hs->mat_names) - 1);
    return (int)hs->n_mat++;
}

int mrt_hs_find_material(const mrt_host_scene *hs, const char *name)
{
    if (!hs || !name)
        return MRT_E_INVALID;
    for (uint32_t i = 0; i < hs->n_mat; i++)
        if (strcmp(hs->mat_names[i], name) == 0)
            return (int)i;
    return MRT_E_INVALID;
}

int mrt_hs_add_sphere(mrt_host_scene *hs, const float c[3], float r,
                      uint32_t mat)
{
    if (!hs || !c || !(r > 0.0f) || mat >= hs->n_mat)
        return MRT_E_INVALID;
    if (hs->n_sph >= (1u << 30))
        return MRT_E_LIMIT;
    if (grow((void **)&hs->sph, &hs->cap_sph, (size_t)(hs->n_sph + 1) * 4,
             sizeof(float))
        || grow((void **)&hs->sph_mat, &hs->cap_sph_mat, hs->n_sph + 1,
                sizeof(uint32_t)))
        return MRT_E_NOMEM;
    float *s = hs->sph + (size_t)hs->n_sph * 4;
    s[0] = c[0];
    s[1] = c[1];
    s[2] = c[2];
    s[3] = r;
    hs->sph_mat[hs->n_sph] = mat;
    return (int)hs->n_sph++;
}

int mrt_hs_add_quadric(mrt_host_scene *hs, const mrt_quadric *q)
{
    if (!hs || !q || q->kind > MRT_QUADRIC_HYPERBOLOID || q->mat >= hs->n_mat
        || q->mat_top >= hs->n_mat || q->mat_bot >= hs->n_mat)
        return MRT_E_INVALID;
    mrt_quadric c = *q;
    double l = sqrt((double)c.axis[0] * c.axis[0] + (double)c.axis[1] * c.axis[1]
                    + (double)c.axis[2] * c.axis[2]);
    int ok = l > 0.0 && isfinite(l) && c.height > 0.0f && isfinite(c.height);
    for (int k = 0; k < 3; k++)
        ok = ok && isfinite(c.pos[k]);
    if (c.kind == MRT_QUADRIC_HYPERBOLOID)
        ok = ok && c.a > 0.0f && c.b > 0.0f && c.c > 0.0f && isfinite(c.a)
             && isfinite(c.b) && isfinite(c.c) && isfinite(c.shape);
    else
        ok = ok && c.radius > 0.0f && isfinite(c.radius);
    if (!ok)
        return MRT_E_INVALID;
    for (int k = 0; k < 3; k++)
        c.axis[k] = (float)(c.axis[k] / l);
    if (hs->n_quad >= (1u << 30))
        return MRT_E_LIMIT;
    if (grow((void **)&hs->quad, &hs->cap_quad, hs->n_quad + 1,
             sizeof(mrt_quadric)))
        return MRT_E_NOMEM;
    hs->quad[hs->n_quad] = c;
    return (int)hs->n_quad++;
}

int mrt_hs_add_plane(mrt_host_scene *hs, const float p[3], const float n[3],
                     uint32_t mat)
{
    if (!hs || !p || !n || mat >= hs->n_mat)
        return MRT_E_INVALID;
    if (hs->n_pln >= 4096)
        return MRT_E_LIMIT;
    if (grow((void **)&hs->pln, &hs->cap_pln, (size_t)(hs->n_pln + 1) * 6,
             sizeof(float))
        || grow((void **)&hs->pln_mat, &hs->cap_pln_mat, hs->n_pln + 1,
                sizeof(uint32_t)))
        return MRT_E_NOMEM;
    float *q = hs->pln + (size_t)hs->n_pln * 6;
    memcpy(q, p, 3 * sizeof(float));
    memcpy(q + 3, n, 3 * sizeof(float));
    hs->pln_mat[hs->n_pln] = mat;
    return (int)hs->n_pln++;
}

int mrt_hs_add_triangle(mrt_host_scene *hs, const float v[9], const float n[9],
                        uint32_t mat)
{
    if (!hs || !v || mat >= hs->n_mat)
        return MRT_E_INVALID;
    if (hs->n_tri >= (1u << 28))
        return MRT_E_LIMIT;
    size_t need = (size_t)(hs->n_tri + 1) * 9;
    if (grow((void **)&hs->tri_pos, &hs->cap_tri_pos, need, sizeof(float))
        || grow((void **)&hs->tri_nrm, &hs->cap_tri_nrm, need, sizeof(float))
        || grow((void **)&hs->tri_mat, &hs->cap_tri_mat, hs->n_tri + 1,
                sizeof(uint32_t)))
        return MRT_E_NOMEM;
    float *p = hs->tri_pos + (size_t)hs->n_tri * 9;
    float *q = hs->tri_nrm + (size_t)hs->n_tri * 9;
    memcpy(p, v, 9 * sizeof(float));
    if (n) {
        memcpy(q, n, 9 * sizeof(float));
    } else {
        float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        double c[3] = {(double)e1[1] * e2[2] - (double)e1[2] * e2[1],
                       (double)e1[2] * e2[0] - (double)e1[0] * e2[2],
                       (double)e1[0] * e2[1] - (double)e1[1] * e2[0]};
        double l = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        l = l > 0.0 ? 1.0 / __builtin_sqrt(l) : 0.0;
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++)
                q[k * 3 + j] = (float)(c[j] * l);
    }
    if (hs->tri_uv) {
        if (grow((void **)&hs->tri_uv, &hs->cap_tri_uv, need / 9 * 6,
                 sizeof(float)))
            return MRT_E_NOMEM;
        memset(hs->tri_uv + (size_t)hs->n_tri * 6, 0, 6 * sizeof(float));
    }
    hs->tri_mat[hs->n_tri] = mat;
    return (int)hs->n_tri++;
}

int mrt_hs_add_light(mrt_host_scene *hs, const mrt_light *l)
{
    if (!hs || !l)
        return MRT_E_INVALID;
    if (hs->n_light >= MRT_MAX_LIGHTS)
        return MRT_E_LIMIT;
    if (grow((void **)&hs->lights, &hs->cap_light, hs->n_light + 1,
             sizeof(mrt_light)))
        return MRT_E_NOMEM;
    hs->lights[hs->n_light] = *l;
    return (int)hs->n_light++;
}

void mrt_hs_set_camera(mrt_host_scene *hs, const mrt_camera *c)
{
    if (hs && c)
        hs->camera = *c;
}

void mrt_hs_set_ambient(mrt_host_scene *hs, float ratio, const float rgb[3])
{
    if (!hs)
        return;
    hs->ambient_ratio = ratio;
    if (rgb)
        memcpy(hs->ambient_color, rgb, 3 * sizeof(float));
}

void mrt_hs_set_resolution(mrt_host_scene *hs, uint32_t w, uint32_t h)
{
    if (hs) {
        hs->render_w = w;
        hs->render_h = h;
    }
}

void mrt_hs_get_resolution(const mrt_host_scene *hs, uint32_t *w, uint32_t *h)
{
    if (!hs)
        return;
    if (w)
        *w = hs->render_w;
    if (h)
        *h = hs->render_h;
}

void mrt_hs_set_flags(mrt_host_scene *hs, uint32_t flags)
{
    if (hs)
        hs->flags = flags;
}

int mrt_hs_set_hdri(mrt_host_scene *hs, const float *rgb, uint32_t w,
                    uint32_t h, float exposure)
{
    if (!hs || !rgb || w < 2 || h < 2)
        return MRT_E_INVALID;
    size_t n = (size_t)w * h * 3;
    float *p = malloc(n * sizeof(float));
    if (!p)
        return MRT_E_NOMEM;
    memcpy(p, rgb, n * sizeof(float));
    free(hs->hdri_rgb);
    hs->hdri_rgb = p;
    hs->hdri_w = w;
    hs->hdri_h = h;
    hs->hdri_exposure = exposure;
    return MRT_OK;
}

int mrt_hs_get_hdri(const mrt_host_scene *hs, const float **rgb, uint32_t *w,
                    uint32_t *h, float *exposure)
{
    if (!hs || !hs->hdri_rgb)
        return MRT_E_STATE;
    if (rgb)
        *rgb = hs->hdri_rgb;
    if (w)
        *w = hs->hdri_w;
    if (h)
        *h = hs->hdri_h;
    if (exposure)
        *exposure = hs->hdri_exposure;
    return MRT_OK;
}

const mrt_scene *mrt_hs_view(mrt_host_scene *hs)
{
    if (!hs)
        return NULL;
    mrt_scene *v = &hs->view;
    v->n_triangles = hs->n_tri;
    v->tri_pos = hs->tri_pos;
    v->tri_nrm = hs->tri_nrm;
    v->tri_mat = hs->tri_mat;
    v->n_spheres = hs->n_sph;
    v->sph = hs->sph;
    v->sph_mat = hs->sph_mat;
    v->n_quadrics = hs->n_quad;
    v->quadrics = hs->quad;
    v->n_textures = hs->n_tex;
    v->textures = hs->tex;
    v->n_texels = hs->n_texels;
    v->texels = hs->texels;
    v->material_tex = hs->mtex;
    v->tri_uv = hs->tri_uv;
    v->n_planes = hs->n_pln;
    v->pln = hs->pln;
    v->pln_mat = hs->pln_mat;
    v->n_materials = hs->n_mat;
    v->materials = hs->materials;
    v->n_lights = hs->n_light;
    v->lights = hs->lights;
    v->camera = hs->camera;
    v->ambient_ratio = hs->ambient_ratio;
    memcpy(v->ambient_color, hs->ambient_color, sizeof(v->ambient_color));
    v->ior_all = hs->ior_all;
    v->flags = hs->flags;
    return v;
}
