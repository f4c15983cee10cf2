/*
 * mrt_host.h - host-side C helpers exported by libmrt_hip.so: scene
 * containers, the SAH BVH builder, procedural scene generators, the .rt
 * reader, the HDRI inverse-CDF tables, PPM / .rts output and the
 * render-frame loop that drives the device entry points of mrt.h.
 *
 * None of these touch the GPU; they run (and are tested) on any host.
 */
#ifndef MRT_HOST_H
#define MRT_HOST_H

#include "mrt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- BVH (replaces src/ray_tracing/bvh/bvh_build.c:17-124) ------------ */

/* A binary node holding the boxes of BOTH children, so one 64-byte fetch
 * decides which children a ray enters and in which order.
 * child[k] >= 0 : index of an inner node.
 * child[k] <  0 : leaf; v = ~child[k], first primitive = v >> 3,
 *                 count = (v & 7) + 1 (at most MRT_BVH_MAX_LEAF).
 * An absent child has a zero-size box at 1e30 that no ray reaches. */
typedef struct mrt_bvh_node {
    float   lo0[3], hi0[3];
    float   lo1[3], hi1[3];
    int32_t child[2];
    uint32_t pad[2];
} mrt_bvh_node;                       /* 64 bytes */

/* One primitive in leaf order (48 bytes, three 16-byte words).
 * triangle: a = v0.xyz | src id,  b = (v1-v0).xyz | kind 0,
 *           c = (v2-v0).xyz | material
 * sphere:   a = centre.xyz | src id,  b = (radius, radius^2, 0) | kind 1,
 *           c = (0, 0, 0) | material
 * quadric:  a = pos.xyz | src id,  b = axis.xyz | kind 2, 3 or 4
 *           (cylinder, cone, hyperboloid), c = (radius or a, b, c, height);
 *           its shading record (mrt_bvh.shade_nrm) holds shape, mat,
 *           mat_top, mat_bot (the last three as uint32 bits).
 * "src id" is the MRT_PRIM_* tagged index into the scene arrays. */
typedef struct mrt_prim {
    float a[4], b[4], c[4];
} mrt_prim;

#define MRT_PRIMKIND_TRIANGLE 0u
#define MRT_PRIMKIND_SPHERE   1u
#define MRT_PRIMKIND_QUADRIC  2u      /* + MRT_QUADRIC_* */
#define MRT_BVH_MAX_LEAF      4u
#define MRT_BVH_MAX_DEPTH     60u     /* the device traversal stack holds 64 */

typedef struct mrt_bvh {
    uint32_t      n_nodes;
    mrt_bvh_node *nodes;              /* node 0 is the root                  */
    uint32_t      n_prims;
    mrt_prim     *prims;              /* leaf order                          */
    float        *shade_nrm;          /* leaf order, 9 floats: vertex normals */
    uint32_t      depth;              /* deepest inner-node level (root = 1)  */
    uint32_t      n_leaves;
    double        sah_cost;           /* C_trav = 1, C_isect = 1, root area   */
} mrt_bvh;

/* Binned SAH over the scene's triangles and spheres, leaves of at most
 * max_leaf primitives (1..MRT_BVH_MAX_LEAF).  Planes stay out of the tree
 * (the reference tests them linearly, hit_register.c:79-107). */
int  mrt_bvh_build(const mrt_scene *scene, uint32_t max_leaf, mrt_bvh *out);
void mrt_bvh_free(mrt_bvh *bvh);

/* A 4-wide node, the layout the device traverses: the boxes of four
 * children as structure-of-arrays (one 16-byte word per bound and axis) and
 * their codes, in two 64-byte lines.  Child codes follow mrt_bvh_node (>= 0:
 * inner node, < 0: leaf); an unused slot has a box at 1e30 that no ray
 * reaches. */
typedef struct mrt_bvh4_node {
    float   lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    int32_t child[4];
    uint32_t pad[4];
} mrt_bvh4_node;                      /* 128 bytes */

#define MRT_BVH4_MAX_DEPTH 40u        /* stack: 3 entries per level + 1   */
#define MRT_BVH4_TOP       64u        /* nodes laid out breadth-first at the
                                         front of the array (the top levels
                                         the trace kernels may keep in LDS) */

typedef struct mrt_bvh4 {
    uint32_t       n_nodes;
    mrt_bvh4_node *nodes;             /* node 0 is the root; the first
                                         MRT_BVH4_TOP nodes in breadth-first
                                         order, the rest in pre-order       */
    uint32_t       depth;             /* deepest node level (root = 1)      */
} mrt_bvh4;

/* Collapses a binary BVH into a 4-wide one over the same primitive order:
 * every node takes the two children of its binary node and keeps opening
 * the inner child of largest surface area until it holds four. */
int  mrt_bvh4_build(const mrt_bvh *bvh2, mrt_bvh4 *out);
void mrt_bvh4_free(mrt_bvh4 *bvh4);

/* ---- scene container ---------------------------------------------------- */

typedef struct mrt_host_scene mrt_host_scene;

mrt_host_scene *mrt_hs_new(void);
void            mrt_hs_free(mrt_host_scene *hs);
/* Returns the index of the new element, or a negative MRT_E_*. */
int  mrt_hs_add_material(mrt_host_scene *hs, const mrt_material *m,
                         const char *name);
int  mrt_hs_find_material(const mrt_host_scene *hs, const char *name);
int  mrt_hs_add_sphere(mrt_host_scene *hs, const float centre[3], float radius,
                       uint32_t mat);
int  mrt_hs_add_plane(mrt_host_scene *hs, const float point[3],
                      const float normal[3], uint32_t mat);
/* Copies q with its axis normalised (vec3_unit); sizes must be > 0 and
 * finite, the materials defined. */
int  mrt_hs_add_quadric(mrt_host_scene *hs, const mrt_quadric *q);
/* A texture; an IMAGE copies width * height RGBA8 texels from rgba (its
 * texel0 is assigned here).  name may be NULL. */
int  mrt_hs_add_texture(mrt_host_scene *hs, const mrt_texture *t,
                        const uint8_t *rgba, const char *name);
int  mrt_hs_find_texture(const mrt_host_scene *hs, const char *name);
/* Texture slots of material mat (materials without slots get all
 * MRT_NO_TEX, normal_intensity 1, scale 1). */
int  mrt_hs_set_material_tex(mrt_host_scene *hs, uint32_t mat,
                             const mrt_material_tex *mt);
/* (u, v) at the three vertices of triangle tri (others default to 0). */
int  mrt_hs_set_triangle_uv(mrt_host_scene *hs, uint32_t tri,
                            const float uv[6]);
int  mrt_hs_add_triangle(mrt_host_scene *hs, const float v[9],
                         const float n[9] /* NULL: flat */, uint32_t mat);
int  mrt_hs_add_light(mrt_host_scene *hs, const mrt_light *l);
void mrt_hs_set_camera(mrt_host_scene *hs, const mrt_camera *c);
void mrt_hs_set_ambient(mrt_host_scene *hs, float ratio, const float rgb255[3]);
void mrt_hs_set_resolution(mrt_host_scene *hs, uint32_t w, uint32_t h);
void mrt_hs_get_resolution(const mrt_host_scene *hs, uint32_t *w, uint32_t *h);
void mrt_hs_set_flags(mrt_host_scene *hs, uint32_t flags);
/* The HDRI (width*height*3 linear floats, row 0 at the nadir); copied. */
int  mrt_hs_set_hdri(mrt_host_scene *hs, const float *rgb, uint32_t w,
                     uint32_t h, float exposure);
int  mrt_hs_get_hdri(const mrt_host_scene *hs, const float **rgb, uint32_t *w,
                     uint32_t *h, float *exposure);
/* A view of the arrays; valid until the next mrt_hs_add_* call. */
const mrt_scene *mrt_hs_view(mrt_host_scene *hs);

/* ---- procedural scenes --------------------------------------------------- */

/* Ground plane y=0, three radius-1 spheres (lambert 0.8, GGX metal
 * roughness 0.1, dielectric IOR 1.5) centred at x in {-2,0,2} plus a seeded
 * jitter in [-0.5,0.5) on x and z, one point light, ambient 0.1, pinhole
 * camera with a 60 degree field of view; width x height. */
int  mrt_gen_plain_scene(uint64_t seed, uint32_t width, uint32_t height,
                         mrt_host_scene **out);

/* Displaced icosphere (subdiv levels, seeded fBm radial displacement of
 * amplitude amp) over a cells x cells heightfield floor; per-triangle
 * materials drawn 60% lambert / 25% GGX (roughness U[0.05,0.6]) / 15%
 * dielectric; lit by the plain scene's point light and ambient. */
int  mrt_gen_mesh_scene(uint64_t seed, uint32_t width, uint32_t height,
                        uint32_t subdiv, uint32_t floor_cells, float amp,
                        mrt_host_scene **out);

/* The HDRI + caustics scene: the mesh scene without its point light
 * (ambient ratio 1, so the HDRI is unscaled), a rippled water heightfield of
 * water_cells^2 cells (IOR 1.33) just above the floor and a glass sphere
 * (IOR 1.5) beside the icosphere. */
int  mrt_gen_caustic_scene(uint64_t seed, uint32_t width, uint32_t height,
                           uint32_t subdiv, uint32_t floor_cells,
                           uint32_t water_cells, mrt_host_scene **out);

/* Adds a displaced icosphere to hs; returns the triangle count or < 0. */
int  mrt_gen_icosphere(mrt_host_scene *hs, const float centre[3], float radius,
                       uint32_t subdiv, float amp, uint64_t seed,
                       int material_mix /* 0: every triangle uses `mat`;
                                           1: per-triangle 60/25/15 mix;
                                           2: uniform over materials
                                              0 .. mat-1 of hs */,
                       uint32_t mat);

/* Adds a displaced geodesic icosphere of 20 * freq^2 triangles (each
 * icosahedron face split into a freq x freq triangular grid projected onto
 * the sphere; freq 100 gives 200,000 triangles), displaced and coloured as
 * mrt_gen_icosphere's; returns the triangle count or < 0. */
int  mrt_gen_icosphere_freq(mrt_host_scene *hs, const float centre[3],
                            float radius, uint32_t freq, float amp,
                            uint64_t seed, int material_mix, uint32_t mat);

/* Equirectangular sky: analytic gradient plus a sun disc of the given
 * angular radius (degrees) and radiance at a seeded elevation in
 * [15, 60] degrees.  rgb receives w*h*3 floats, row 0 at the nadir. */
int  mrt_gen_hdri_sky(uint64_t seed, uint32_t w, uint32_t h,
                      float sun_radius_deg, float sun_radiance, float *rgb);

/* ---- scene files: Radiance .hdr and Wavefront .obj ---------------------- */

/* Decodes an RGBE .hdr image (src/utils/hdr/parse_hdr*.c): *rgb receives
 * w*h*3 floats c * 2^(e-128) / 255, row 0 at the bottom of the image; the
 * header's EXPOSURE and GAMMA (defaults 0 and 1) are returned.  Free *rgb
 * with mrt_free. */
int  mrt_decode_hdr(const uint8_t *data, size_t size, float **rgb, uint32_t *w,
                    uint32_t *h, float *exposure, float *gamma);
int  mrt_load_hdr(const char *path, float **rgb, uint32_t *w, uint32_t *h,
                  float *exposure, float *gamma);
void mrt_free(void *p);

/* BMP / PNG texture images as RGBA8 (w * h * 4 bytes, free with mrt_free),
 * rows as the reference's readers order them (src/utils/bmp, src/utils/png):
 * the file's rows for BMP, the last scanline first for PNG. */
int  mrt_decode_image(const uint8_t *data, size_t size, uint8_t **rgba,
                      uint32_t *w, uint32_t *h);
int  mrt_load_image(const char *path, uint8_t **rgba, uint32_t *w,
                    uint32_t *h);

/* Adds an OBJ mesh placed as the .rt "cu" element places it (scale, then
 * the rotation taking +Z to dir, then pos); returns the triangle count. */
int  mrt_hs_add_obj(mrt_host_scene *hs, const char *path, const float pos[3],
                    const float dir[3], float scale, uint32_t mat);
int  mrt_hs_add_obj_text(mrt_host_scene *hs, const char *text,
                         const float pos[3], const float dir[3], float scale,
                         uint32_t mat);

/* ---- HDRI tables (src/ray_tracing/importance_sampling/calc_*.c) --------- */

typedef struct mrt_hdri_tables {
    uint32_t w, h;
    float   *pdf_joint;          /* w*h                                    */
    float   *marg_inv;           /* h   quantised inverse marginal CDF     */
    float   *cond_inv;           /* w*h quantised inverse conditional CDF  */
    double   column_average;
} mrt_hdri_tables;

int  mrt_hdri_build_tables(const float *rgb, uint32_t w, uint32_t h,
                           float exposure, mrt_hdri_tables *out);
void mrt_hdri_free_tables(mrt_hdri_tables *t);

/* Tables of the MIS estimator: a piecewise-constant density over the cells
 * that read each texel, weighted by luminance x solid angle. */
typedef struct mrt_hdri_mis {
    uint32_t w, h;
    float   *marg;               /* h:         row CDF, marg[0] = 0, [h-1] = 1 */
    float   *cond;               /* (h-1) * w: per-row column CDFs          */
    float   *pcell;              /* (h-1) * (w-1): cell probabilities       */
    double   total;              /* sum of the cell weights                 */
} mrt_hdri_mis;

int  mrt_hdri_build_mis(const float *rgb, uint32_t w, uint32_t h,
                        float exposure, mrt_hdri_mis *out);
void mrt_hdri_free_mis(mrt_hdri_mis *m);

/* ---- .rt reader (src/parsing/parse_scene.c:99-125) ---------------------- */

/* Reads a .rt scene: A C L l dl sp pl cy co hb cu mat tex W (image textures
 * other than .hdr skyboxes, checker textures and textured materials are
 * rejected with MRT_E_UNSUPPORTED).  err receives a message on failure. */
int  mrt_parse_rt_file(const char *path, mrt_host_scene **out, char *err,
                       size_t errlen);
int  mrt_parse_rt_string(const char *text, mrt_host_scene **out, char *err,
                         size_t errlen);

/* ---- output (src/utils/ppm/export_ppm.c, src/utils/export_render_state.c) */

/* P6 PPM of clamp01(pow(acc / samples, 1 / gamma)) * 255. */
int  mrt_export_ppm(const char *path, const double *float_render,
                    uint32_t w, uint32_t h, uint32_t samples, double gamma);
/* render_state.rts: [i32 w][i32 h][i32 sample][w*h*3 f64]. */
int  mrt_export_rts(const char *path, const double *float_render,
                    uint32_t w, uint32_t h, uint32_t samples);
int  mrt_load_rts(const char *path, double *float_render, uint32_t w,
                  uint32_t h, uint32_t *samples);

/* BVH heat map, as set_heat_map_color (src/ray_tracing/calc_sample.c:16-42)
 * with get_heatmap_fcolor (src/math/color/get_heat_map_color.c:16-34): adds
 * the ramp colour of depth[i] / max(depth) to float_render[i] (3 doubles)
 * for every pixel with depth[i] >= 0 (a miss is -1).  depth is the BVH work
 * per pixel, e.g. MRT_AOV_NODES of the pixels MRT_AOV_PRIM_ID reports hit. */
void mrt_heat_map_add(const int32_t *depth, uint32_t w, uint32_t h,
                      double *float_render);

/* ---- render-frame loop (src/rendering/loop.c:36-63) --------------------- */

typedef struct mrt_frame {
    mrt_ctx  *ctx;
    uint32_t  w, h;
    uint32_t  sample;            /* screen.sample                          */
    uint32_t  spp;               /* screen.spp (--max-samples)             */
    uint32_t  max_bounces;       /* controls.max_bounces                   */
    uint32_t  samples_per_tick;  /* samples rendered per render_frame call */
    double   *float_render;      /* host-owned W*H*3 sums                  */
} mrt_frame;

/* One tick: on sample 0 clears the sums, then renders samples_per_tick
 * samples (never past spp) and adds them into float_render.  Returns the
 * number of samples added, 0 once spp is reached, < 0 on failure. */
int  mrt_frame_tick(mrt_frame *f);

#ifdef __cplusplus
}
#endif

#endif /* MRT_HOST_H */
