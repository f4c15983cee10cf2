/* Private host-side declarations shared by the host C files. */
#ifndef MRT_HOST_INTERNAL_H
#define MRT_HOST_INTERNAL_H

#include "mrt_host.h"

#define MRT_MAX_LIGHTS 64u

struct mrt_host_scene {
    mrt_scene     view;
    float        *tri_pos, *tri_nrm;
    uint32_t     *tri_mat;
    size_t        cap_tri_pos, cap_tri_nrm, cap_tri_mat;
    uint32_t      n_tri;
    float        *sph;
    uint32_t     *sph_mat;
    size_t        cap_sph, cap_sph_mat;
    uint32_t      n_sph;
    mrt_quadric  *quad;
    size_t        cap_quad;
    uint32_t      n_quad;
    mrt_texture  *tex;
    char        (*tex_names)[32];
    size_t        cap_tex, cap_tex_names;
    uint32_t      n_tex;
    uint8_t      *texels;
    size_t        cap_texels;
    uint32_t      n_texels;
    mrt_material_tex *mtex;           /* NULL until a material has a slot */
    size_t        cap_mtex;
    float        *tri_uv;             /* NULL until a triangle has (u, v) */
    size_t        cap_tri_uv;
    float        *pln;
    uint32_t     *pln_mat;
    size_t        cap_pln, cap_pln_mat;
    uint32_t      n_pln;
    mrt_material *materials;
    char        (*mat_names)[32];
    size_t        cap_mat, cap_mat_names;
    uint32_t      n_mat;
    mrt_light    *lights;
    size_t        cap_light;
    uint32_t      n_light;
    mrt_camera    camera;
    float         ambient_ratio;
    float         ambient_color[3];
    float         ior_all;
    uint32_t      flags;
    uint32_t      render_w, render_h;
    float        *hdri_rgb;
    uint32_t      hdri_w, hdri_h;
    float         hdri_exposure;
};

/* splitmix64: the seed expander used by every generator. */
static inline uint64_t mrt_splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform double in [0,1) with 53 random bits. */
static inline double mrt_u01(uint64_t *s)
{
    return (double)(mrt_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

#endif
