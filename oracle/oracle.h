/*
 * ORACLE - test infrastructure only.
 *
 * A CPU restatement, in plain C, of the reference's CPU render path
 * (src/main_cpu.c:45-61 -> render_cpu -> calc_one_sample -> path_trace) on
 * the counter-based RNG stream that the HIP path also consumes.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the timed CPU baseline; the product never calls it.
 *
 * Every function names the reference file:line it follows.  Two builds of the
 * same source: *_f64 keeps the reference's precision (fp64 vectors, fp32
 * Moller-Trumbore scalars, includes/structs.h:99-110) and *_f32 runs the
 * same steps in fp32 in the order the HIP kernels use.
 *
 * Pinning (what checks each part): see tests/test_oracle_*.py.  Philox is
 * pinned by the Random123 known-answer vectors and an independent numpy
 * restatement; traversal by brute force; intersections by closed forms;
 * shading by the closed-form scenes; HDRI tables by direct re-derivation
 * on small maps.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stdint.h>
#include "mrt.h"
#include "mrt_host.h"   /* node / primitive layout of the BVH being traversed */

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TRAV_BRUTE      0   /* every primitive, no BVH                 */
#define ORC_TRAV_REFERENCE  1   /* hit_bvh: both children, no culling      */
#define ORC_TRAV_ORDERED    2   /* near child first, t-culled (HIP order):
                                   4-wide when the BVH carries nodes4        */

typedef struct orc_bvh {
    const mrt_bvh_node *nodes;
    uint32_t            n_nodes;
    const mrt_prim     *prims;     /* only a[3] (source id) is read          */
    uint32_t            n_prims;
    /* the 4-wide collapse of `nodes` (mrt_bvh4_build); when present the
     * ordered traversal walks it, in the device's order */
    const mrt_bvh4_node *nodes4;
    uint32_t             n_nodes4;
} orc_bvh;

typedef struct orc_hdri {
    uint32_t     w, h;
    const float *rgb;              /* w*h*3 linear texels, row 0 at nadir    */
    float        exposure;
    const float *pdf_joint;        /* w*h  (reference tables)                */
    const float *marg_inv;         /* h                                      */
    const float *cond_inv;         /* w*h                                    */
    const float *mis_marg;         /* h        (orc_hdri_mis_tables)         */
    const float *mis_cond;         /* (h-1)*w                                */
    const float *mis_pcell;        /* (h-1)*(w-1)                            */
} orc_hdri;

/* A photon map: positions and powers (3 floats each) and the gather rule:
 * k > 0: the k nearest within squared distance `radius` (the reference's
 * k-NN, get_caustic.c:37-88); k == 0: every photon within distance
 * `radius`.  Estimate = sum of powers / (pi r^2) / norm. */
typedef struct orc_photon_map {
    const float *pos;
    const float *pow;
    uint32_t     n;
    uint32_t     k;
    float        radius;
    float        norm;
} orc_photon_map;

typedef struct orc_params {
    uint32_t width, height;
    uint32_t first_sample, n_samples, max_depth;
    uint64_t seed;
    int      traversal;            /* ORC_TRAV_*                             */
    int      estimator;            /* MRT_EST_*                              */
    int      threads;              /* 0 = OpenMP default                     */
    uint32_t row_begin, row_end;   /* rows rendered (end 0 = height)         */
} orc_params;

typedef struct orc_counters {
    uint64_t rays_closest, rays_shadow;
    uint64_t nodes_closest, prims_closest, nodes_shadow, prims_shadow;
} orc_counters;

/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
/* The 4 uniforms of block `block` of bounce `bounce`: ctr = {pixel, sample,
 * bounce << 20 | block, stream}, key = seed; u = (x >> 8) * 2^-24. */
void orc_uniform4(uint64_t seed, uint32_t pixel, uint32_t sample,
                  uint32_t bounce, uint32_t block, uint32_t stream,
                  float out[4]);

/* init_viewport_values / init_viewport_vectors
 * (src/ray_tracing/viewport.c:38-88) in double: out = origin[3],
 * pixel00[3], du[3], dv[3], disk_u[3], disk_v[3], defocus_radius. */
void orc_viewport(const mrt_camera *cam, uint32_t w, uint32_t h,
                  double out[19]);

/* Primary rays of sample `sample` in fp32, as the HIP ray generator makes
 * them: orig/dir are w*h*3 floats. */
void orc_raygen_f32(const mrt_camera *cam, uint32_t w, uint32_t h,
                    uint32_t sample, uint64_t seed, float *orig, float *dir);

/* Closest hits of n given rays (fp32 inputs, widened for f64).  prim gets
 * MRT_PRIM_* tagged ids, t the distance, nodes / prims the BVH nodes visited
 * and primitives tested (ORC_TRAV_ORDERED counts what the HIP kernel
 * counts).  Any output pointer may be NULL. */
void orc_trace_rays_f32(const mrt_scene *s, const orc_bvh *b, int traversal,
                        uint32_t n, const float *orig, const float *dir,
                        uint32_t *prim, float *t, uint32_t *nodes,
                        uint32_t *prims, int threads);
void orc_trace_rays_f64(const mrt_scene *s, const orc_bvh *b, int traversal,
                        uint32_t n, const float *orig, const float *dir,
                        uint32_t *prim, double *t, uint32_t *nodes,
                        uint32_t *prims, int threads);

/* Closest hits with the shading frame: normal facing the ray (3 per ray,
 * normal-mapped), front side flag, material of the part hit and its
 * (textured) colour (3 per ray; albedo may be NULL). */
void orc_hit_info_f32(const mrt_scene *s, const orc_bvh *b, int traversal,
                      uint32_t n, const float *orig, const float *dir,
                      uint32_t *prim, float *t, float *nrm, uint8_t *front,
                      uint32_t *mat, float *albedo);
void orc_hit_info_f64(const mrt_scene *s, const orc_bvh *b, int traversal,
                      uint32_t n, const float *orig, const float *dir,
                      uint32_t *prim, double *t, double *nrm, uint8_t *front,
                      uint32_t *mat, float *albedo);

/* Any-hit: occluded[i] = 1 when a primitive lies at t in [0.001, tmax[i])
 * (tmax capped at 1000 inclusive), as check_plight_hit
 * (src/ray_tracing/light/p_light.c:48-64). */
void orc_occluded_f32(const mrt_scene *s, const orc_bvh *b, uint32_t n,
                      const float *orig, const float *dir, const float *tmax,
                      uint8_t *occluded, uint32_t *nodes, uint32_t *prims);

/* Path tracing of rows [row_begin, row_end) x all columns, samples
 * first_sample .. first_sample + n_samples - 1, added into acc
 * (width*height*3 doubles, like screen.float_render). */
int  orc_render_f32(const mrt_scene *s, const orc_bvh *b, const orc_hdri *hd,
                    const orc_photon_map *pm, const orc_params *p,
                    double *acc, orc_counters *cnt);
int  orc_render_f64(const mrt_scene *s, const orc_bvh *b, const orc_hdri *hd,
                    const orc_photon_map *pm, const orc_params *p,
                    double *acc, orc_counters *cnt);

/* Photon emission (caustic_manager.c:20-102, trace_photon_path.c:18-75):
 * n_per photons from every point light and, with an HDRI, n_per photons
 * from the environment through the disk of the scene's bounding sphere.
 * Stored photons go to pos / pow (3 floats each) and gid (global photon
 * index) in increasing gid order; returns how many were stored.  The
 * arrays hold (point lights + (hd != NULL)) * n_per entries. */
uint32_t orc_emit_photons_f32(const mrt_scene *s, const orc_bvh *b,
                              const orc_hdri *hd, uint32_t n_per,
                              uint64_t seed, float *pos, float *pow,
                              uint32_t *gid);
uint32_t orc_emit_photons_f64(const mrt_scene *s, const orc_bvh *b,
                              const orc_hdri *hd, uint32_t n_per,
                              uint64_t seed, float *pos, float *pow,
                              uint32_t *gid);

/* Forward-mode derivative of the rendered image with respect to one
 * parameter, on the MIS estimator (or no HDRI), exact light sums, no photon
 * map: the independent reference for the HIP adjoint.  Albedo, emission
 * (E = strength * colour, taken as the parameter) and texel derivatives are
 * exact derivatives of the estimator; roughness contributes the
 * score-function term d log p(m)/dr of every GGX micro-normal sample. */
#define ORC_P_ALBEDO    0
#define ORC_P_ROUGHNESS 1
#define ORC_P_EMISSION  2
#define ORC_P_TEXEL     3
typedef struct orc_param {
    int      kind;
    uint32_t mat;      /* material (albedo, roughness, emission)            */
    uint32_t chan;     /* colour channel (albedo, emission, texel)          */
    uint32_t texel;    /* y * w + x (texel)                                 */
} orc_param;
/* img and dimg: width*height*3 doubles, sums over the samples (added to). */
int  orc_render_grad_f32(const mrt_scene *s, const orc_bvh *b,
                         const orc_hdri *hd, const orc_params *p,
                         const orc_param *q, double *img, double *dimg);
int  orc_render_grad_f64(const mrt_scene *s, const orc_bvh *b,
                         const orc_hdri *hd, const orc_params *p,
                         const orc_param *q, double *img, double *dimg);

/* The caustic estimate at n points (3 floats each) into out (3 floats). */
void orc_caustic_f32(const orc_photon_map *pm, uint32_t n, const float *pts,
                     float *out);

/* HDRI importance tables (src/ray_tracing/importance_sampling/calc_pdf.c,
 * calc_cdf.c, importance_sampling_manager.c): pdf_joint w*h, marg_inv h,
 * cond_inv w*h. */
int  orc_hdri_tables(const float *rgb, uint32_t w, uint32_t h, float exposure,
                     float *pdf_joint, float *marg_inv, float *cond_inv,
                     double *column_average);

/* Tables of the MIS environment estimator (not in the reference). */
int  orc_hdri_mis_tables(const float *rgb, uint32_t w, uint32_t h,
                         float exposure, float *marg, float *cond,
                         float *pcell, double *total);

/* The samplers of ggx_sample_hemisphere (src/material/utils.c:39-63),
 * cos_weighted_sample_hemishphere (material_default.c:20-39) and
 * random_in_unit_sphere (src/math/random.c:42-50), fp64, for tests. */
void orc_sample_ggx_f64(const double n[3], double rough, float u1, float u2,
                        double out[3]);
void orc_sample_cos_f64(const double n[3], float u1, float u2, double out[3]);
void orc_unit_sphere_f64(const float u[4], double out[3]);

/* The transmission branch of a smooth dielectric (refracted_ray,
 * src/material/refraction.c:16-76, with its TIR fallback to
 * metallic_color, src/material/metallic.c:25-48) for one direction d at a
 * surface whose normal n faces it; power starts at 1.  Returns 1 when the
 * ray was totally reflected. */
int  orc_refract_f32(const double d[3], const double n[3], int front,
                     float ior, float ior_all, const float color[3],
                     double out[3], double power[3]);
int  orc_refract_f64(const double d[3], const double n[3], int front,
                     float ior, float ior_all, const float color[3],
                     double out[3], double power[3]);

#ifdef __cplusplus
}
#endif

#endif
