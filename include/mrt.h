/*
 * mrt.h - C ABI of libmrt_hip.so, the MI355X (gfx950) wavefront path tracer
 * that replaces miniRT's render back-ends (src/ray_tracing, src/gpu,
 * src/caustic and the HDRI sampler) behind the reference's render-frame /
 * accumulation-buffer contract.
 *
 * Plain C types only: pointers, sizes and the POD structs below.  Every entry
 * point returns an int (MRT_OK = 0, negative MRT_E_* on failure) and never
 * calls exit(); mrt_last_error() gives the message of the last failure on a
 * context.  One context per host thread; calls on one context are not
 * thread-safe (the caller serialises them, as the reference's main thread
 * does with its GL context).
 */
#ifndef MRT_H
#define MRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MRT_OK              0
#define MRT_E_INVALID      -1   /* bad argument or scene data               */
#define MRT_E_NOMEM        -2   /* host or device allocation failed         */
#define MRT_E_HIP          -3   /* a HIP runtime call failed                */
#define MRT_E_STATE        -4   /* call out of order (e.g. render w/o scene) */
#define MRT_E_UNSUPPORTED  -5   /* feature not built in this library        */
#define MRT_E_LIMIT        -6   /* a size exceeds a documented maximum      */
#define MRT_E_NODEVICE     -7   /* no usable gfx950 device                  */

/* ---- scene contract (host-owned; the library copies on upload) --------- */

/* One material.  Mirrors t_mat (includes/scene_structs.h:84-106) with the
 * texture slots dropped: color is color_value / 255 (color_to_fcolor,
 * src/math/color/conversions.c:15-18), emission_color likewise.  An object
 * given a bare colour in the .rt file is a material with metallic = 0,
 * ior = 0, ao = 1 and no emission, which takes exactly the reference's
 * "no material" branch (default_mat, src/ray_tracing/path_trace.c:18-19). */
typedef struct mrt_material {
    float color[3];
    float metallic;           /* 1 metal, 0 dielectric/diffuse, else random */
    float roughness;          /* GGX alpha = roughness^2                     */
    float ior;                /* 0 => diffuse                                */
    float transmission;
    float ao;
    float emission_strength;
    float emission_color[3];
} mrt_material;               /* 48 bytes */

#define MRT_LIGHT_POINT        0
#define MRT_LIGHT_DIRECTIONAL  1

/* A light: t_light / t_dlight (includes/scene_structs.h:137-152). */
typedef struct mrt_light {
    int32_t type;             /* MRT_LIGHT_*                                 */
    float   vec[3];           /* position (point) or orientation (dir.)      */
    float   brightness;       /* 0..1                                        */
    float   color[3];         /* 0..255, as t_color                          */
    float   radius;           /* soft shadows when > 0 and samples > 1       */
    int32_t shadow_samples;
} mrt_light;                  /* 40 bytes */

/* Camera: t_camera (includes/scene_structs.h:128-135). */
typedef struct mrt_camera {
    float pos[3];
    float dir[3];             /* normalised by the library                   */
    float fov_deg;            /* horizontal field of view                    */
    float defocus_angle_deg;  /* 0 = pinhole                                 */
    float focus_dist;         /* 0 => 1, as the reference's UI default       */
} mrt_camera;

#define MRT_SCENE_TRIANGLE_ONE_SIDED  0x1u  /* --triangle-one-sided         */
#define MRT_SCENE_LIGHT_EXACT         0x2u  /* do not truncate light sums to
                                               integers (reference t_lcolor
                                               is uint64_t and truncates)   */

/* Cylinders, cones and hyperboloids: t_cylinder / t_cone / t_hyperboloid
 * (includes/scene_structs.h) as the .rt "cy", "co" and "hb" elements give
 * them (parse_elements3.c:15-57, parse_elements4.c:15-99).
 *   cylinder:    centre pos, |proj on axis| <= height / 2 of the centre,
 *                radius = diameter / 2, closed by two caps
 *                (hit_cylinder.c:107-134);
 *   cone:        apex pos, body from the apex to apex + axis * height, base
 *                radius = diameter / 2, closed by the base cap
 *                (hit_cone.c:111-129);
 *   hyperboloid: centre pos, x^2/a^2 + y^2/b^2 - z^2/c^2 = shape in a frame
 *                whose z is the axis, |z| <= height / 2, no caps
 *                (hit_hyperboloid.c:79-104).
 * axis is a unit vector (the parser normalises it).  mat_top / mat_bot are
 * the cap materials (material_top / material_bot, or mat when the file
 * names none); the cone has no bottom cap. */
#define MRT_QUADRIC_CYLINDER     0u
#define MRT_QUADRIC_CONE         1u
#define MRT_QUADRIC_HYPERBOLOID  2u

typedef struct mrt_quadric {
    uint32_t kind;            /* MRT_QUADRIC_*                               */
    float    pos[3];
    float    axis[3];
    float    height;
    float    radius;          /* cylinder, cone                              */
    float    a, b, c;         /* hyperboloid semi-axes                       */
    float    shape;           /* hyperboloid right-hand side                 */
    uint32_t mat, mat_top, mat_bot;
} mrt_quadric;                /* 64 bytes */

/* Textures: the .rt "tex" element (parse_texture, parse_elements5.c:
 * 85-142) and get_tex_color / get_tex_transparency
 * (hit_register/get_tex_color.c:27-67).
 *   IMAGE           nearest texel ((int)(u (w-1)), (int)(v (h-1))) of an
 *                   RGBA8 image, rows in file order; colour rgb / 255,
 *                   transparency 1 - a / 255;
 *   CHECKER_LOCAL   c1 where floor(u scale) + floor(v scale) is even, else c2;
 *   CHECKER_GLOBAL  c1 where floor(s x) + floor(s y + 1/2) + floor(s z) of
 *                   the hit point is even, else c2 (get_solid_texture,
 *                   textures/get_solid_texture.c:30-43).
 * Checker colours are 0..255 as in the file. */
#define MRT_TEX_IMAGE           0u
#define MRT_TEX_CHECKER_LOCAL   1u
#define MRT_TEX_CHECKER_GLOBAL  2u
#define MRT_TEX_BLACK           3u   /* a material's .hdr texture: the
                                        reference's get_tex_color returns 0 */
#define MRT_NO_TEX              0xFFFFFFFFu

typedef struct mrt_texture {
    uint32_t type;            /* MRT_TEX_*                                   */
    uint32_t width, height;   /* IMAGE                                       */
    uint32_t texel0;          /* IMAGE: first texel in mrt_scene.texels      */
    float    scale;           /* checkers                                    */
    float    c1[3], c2[3];
    uint32_t pad;
} mrt_texture;                /* 48 bytes */

/* The texture slots of one material (t_mat, includes/scene_structs.h:
 * 84-106): a texture index or MRT_NO_TEX per slot.  metallic, roughness,
 * transmission, ao and emission_strength take the texture's red channel,
 * color and emission_color its colour (apply_all_map, apply_map.c:60-86);
 * normal is a tangent-space normal map mixed in by normal_intensity
 * (apply_normal_map, apply_map.c:37-58).  scale multiplies the (u, v) of
 * spheres and quadrics (then modf) and 0.1 * scale those of planes (then
 * v - floor(v)). */
typedef struct mrt_material_tex {
    uint32_t color, metallic, roughness, transmission, ao;
    uint32_t emission_strength, emission_color, normal;
    float    normal_intensity;
    float    scale;
} mrt_material_tex;           /* 40 bytes */

typedef struct mrt_scene {
    uint32_t            n_triangles;
    const float        *tri_pos;      /* 9 floats / tri: v0 v1 v2             */
    const float        *tri_nrm;      /* 9 floats / tri: vertex normals, or
                                         NULL for the geometric normal        */
    const uint32_t     *tri_mat;      /* material index / tri                 */
    uint32_t            n_spheres;
    const float        *sph;          /* 4 floats / sphere: centre, radius    */
    const uint32_t     *sph_mat;
    uint32_t            n_planes;
    const float        *pln;          /* 6 floats / plane: point, normal      */
    const uint32_t     *pln_mat;
    uint32_t            n_materials;
    const mrt_material *materials;
    uint32_t            n_lights;
    const mrt_light    *lights;
    mrt_camera          camera;
    float               ambient_ratio;     /* A ratio                          */
    float               ambient_color[3];  /* A colour, 0..255                 */
    float               ior_all;           /* scene->ior_all (1.0)             */
    uint32_t            flags;             /* MRT_SCENE_*                      */
    uint32_t            n_quadrics;
    const mrt_quadric  *quadrics;
    /* textures (all may be 0 / NULL) */
    uint32_t            n_textures;
    const mrt_texture  *textures;
    uint32_t            n_texels;
    const uint8_t      *texels;        /* RGBA8, 4 bytes / texel           */
    const mrt_material_tex *material_tex;  /* NULL, or one per material    */
    const float        *tri_uv;        /* NULL, or 6 floats / tri: (u, v)
                                          at v0, v1, v2 (OBJ "vt")          */
} mrt_scene;

/* Primitive ids reported by MRT_AOV_PRIM_ID: the index into the array the
 * primitive came from, tagged with its kind.  Miss = MRT_PRIM_NONE. */
#define MRT_PRIM_NONE      0xFFFFFFFFu
#define MRT_PRIM_TRIANGLE  0x00000000u
#define MRT_PRIM_SPHERE    0x40000000u
#define MRT_PRIM_PLANE     0x80000000u
#define MRT_PRIM_QUADRIC   0xC0000000u
#define MRT_PRIM_KIND_MASK 0xC0000000u

/* Estimator for the HDRI term (see mrt_upload_hdri). */
#define MRT_EST_REFERENCE  0   /* default_mat's importance_sampling term      */
#define MRT_EST_MIS        1   /* proper solid-angle pdf + balance heuristic  */

/* Options for mrt_set_option. */
#define MRT_OPT_SEED        1  /* RNG key, an integer in [0, 2^53] (a larger
                                  key needs mrt_set_seed)                   */
#define MRT_OPT_ESTIMATOR   2  /* MRT_EST_*                                  */
#define MRT_OPT_STATS       3  /* MRT_STATS_* flags                          */
#define MRT_OPT_PATH_BUDGET 4  /* paths in flight per wavefront pass
                                  (default 2^27: 64 spp at 1080p, ~33 GB of
                                  queues per render lane)                   */
#define MRT_OPT_BLOCK       5  /* threads per block for the trace kernels    */
#define MRT_OPT_PHOTON_K    6  /* caustic gather: k nearest photons (default
                                  15, get_caustic.c); 0 = every photon within
                                  the radius                                */
#define MRT_OPT_PHOTON_NORM 7  /* caustic norm factor (--caustic=..;..;NORM) */

#define MRT_OPT_BUILDER     8  /* MRT_BUILDER_*                              */
#define MRT_OPT_LANES      10  /* 1..4 render lanes: the passes of a render
                                  of more samples than the path budget holds
                                  run on up to this many HIP streams, each
                                  with a pass's queues, a pass starting once
                                  the one before it has traced its first
                                  bounce, so that one pass's launches fill
                                  the GPU while the other's late bounces wait
                                  for their longest rays (same image).
                                  Default 2; the passes of a batch are split
                                  into this many parts of >= 65,536 paths  */

#define MRT_OPT_TRACE_TOP  11  /* 1 (default): the trace kernel of scenes of
                                  triangles and spheres keeps the tree's top
                                  levels (the first nodes, laid out
                                  breadth-first) in LDS; 0: every node from
                                  global memory (same results)              */

#define MRT_OPT_TRACE_REFILL 12 /* 1 (default): with MRT_OPT_TRACE_TOP, a
                                  lane of the trace kernel whose ray is done
                                  takes the next ray of its wave's chunk
                                  instead of waiting for the wave's slowest
                                  ray; 0: a wave traces 64 rays to the end
                                  (same results)                            */

/* BVH builders (MRT_OPT_BUILDER). */
#define MRT_BUILDER_HOST    0  /* binned SAH on the host (default)            */
#define MRT_BUILDER_DEVICE  1  /* PLOC clustering on the device; falls back to
                                  the host below 64 primitives or when the
                                  tree would be too deep for the traversal  */

/* Flags of MRT_OPT_STATS. */
#define MRT_STATS_COUNT     1  /* count BVH nodes / primitives per ray       */
#define MRT_STATS_TIME      2  /* time every kernel launch with HIP events   */

/* Arbitrary output variables of primary rays (mrt_render_aov). */
#define MRT_AOV_PRIM_ID     0  /* uint32 per pixel                           */
#define MRT_AOV_T           1  /* float per pixel, distance along unit dir   */
#define MRT_AOV_NORMAL      2  /* 3 floats per pixel, facing the ray         */
#define MRT_AOV_NODES       3  /* uint32 per pixel: BVH nodes visited        */
#define MRT_AOV_PRIMS       4  /* uint32 per pixel: primitives tested        */

/* Counters and kernel times since the last mrt_reset_stats. */
typedef struct mrt_stats {
    uint64_t paths;            /* camera paths started                        */
    uint64_t rays_closest;     /* closest-hit segments traced                 */
    uint64_t rays_shadow;      /* shadow / visibility rays traced             */
    uint64_t nodes_closest;    /* BVH nodes visited by closest-hit rays       */
    uint64_t prims_closest;    /* primitives tested by closest-hit rays       */
    uint64_t nodes_shadow;
    uint64_t prims_shadow;
    uint64_t launches_trace;   /* trace launches: the closest-hit rays of a
                                  bounce with the shadow rays of the one
                                  before                                      */
    uint64_t launches_shadow;  /* shadow-only launches (0: shadow rays ride
                                  in the trace launches)                      */
    uint64_t launches_shade;
    uint64_t nonfinite;        /* NaN/Inf samples zeroed                      */
    double   ms_raygen;        /* summed HIP-event times per kernel class     */
    double   ms_trace;         /* trace launches (closest hit + shadow rays) */
    double   ms_shade;         /* shade, light sums, queue advance           */
    double   ms_shadow;        /* shadow-only launches                       */
    double   ms_accum;
    double   ms_total;         /* mrt_render wall time on the stream          */
    double   ms_caustic;       /* photon-map gather                           */
    uint64_t gather_queries;   /* caustic gathers (MRT_STATS_COUNT)           */
    uint64_t gather_photons;   /* photon records those gathers examined       */
    uint64_t gather_cells;     /* nonempty buckets whose bounds they read     */
    uint64_t gather_lookups;   /* grid cells they looked up                   */
} mrt_stats;

typedef struct mrt_ctx mrt_ctx;

/* Replaces init_shader (src/gpu/init_shader.c:133-159, includes/gpu.h:32):
 * picks the device, creates the stream and the W x H accumulation buffer. */
int  mrt_init(int device, uint32_t width, uint32_t height, mrt_ctx **out);

/* Replaces convert_scene_build (src/gpu/gpu_scene.c:104-124, includes/gpu.h:86)
 * together with init_bvh / build_bvh (src/ray_tracing/bvh/bvh_manager.c:17-37):
 * builds the BVH on the host and uploads every array.  The library copies;
 * the caller keeps ownership of the scene. */
int  mrt_upload_scene(mrt_ctx *ctx, const mrt_scene *scene);

/* Replaces convert_scene (camera/viewport only, src/gpu/gpu_scene.c:62-75). */
int  mrt_set_camera(mrt_ctx *ctx, const mrt_camera *camera);

/* Replaces send_importance_sampling (src/gpu/importance_sampling.c:69-96) and
 * make_importance_sampling_map (src/ray_tracing/importance_sampling/
 * importance_sampling_manager.c:41-61): rgb is width*height linear texels,
 * row 0 at the nadir (the reference's hdr.pixels order), exposure as
 * t_hdr.exposure.  Uploads the texels and builds the inverse-CDF tables of
 * the reference estimator and the MIS tables on the device (bit for bit
 * the tables mrt_hdri_build_tables / mrt_hdri_build_mis give on the host,
 * where the reference builds them in O(W^2 H)). */
int  mrt_upload_hdri(mrt_ctx *ctx, const float *rgb, uint32_t width,
                     uint32_t height, float exposure);

/* No counterpart in the reference: the uploaded scene's BVH as the kernels
 * read it: counts, 4-wide depth, whether the device built it, and (when the
 * pointers are not NULL) copies of the 4-wide nodes (n_nodes x 128 bytes,
 * mrt_bvh4_node), the primitive records (n_prims x 48, mrt_prim) and the
 * shading records (n_prims x 9 floats). */
int  mrt_get_scene_bvh(mrt_ctx *ctx, uint32_t *n_nodes, uint32_t *n_prims,
                       uint32_t *depth, int *on_device, void *nodes,
                       void *prims, float *shade_nrm);

/* No counterpart in the reference: copies the device's HDRI tables back
 * (for inspection and tests): pdf_joint, cond_inv (w*h), marg_inv (h) as
 * mrt_hdri_tables; mis_marg (h), mis_cond ((h-1)*w), mis_pcell
 * ((h-1)*(w-1)) as mrt_hdri_mis.  Any pointer may be NULL. */
int  mrt_get_hdri_tables(mrt_ctx *ctx, float *pdf_joint, float *marg_inv,
                         float *cond_inv, float *mis_marg, float *mis_cond,
                         float *mis_pcell);

/* Replaces caustic_manager (src/caustic/caustic_manager.c:83-102) and
 * build_caustic_grid (src/gpu/convert_caustic.c:85-102), the handling of
 * --caustic=N;RADIUS;NORM (src/options/option_caustic.c:63-88): traces
 * n_photons photons from every point light (and from the HDRI when one is
 * uploaded), keeps those that reach a diffuse surface through a specular
 * event, and builds the hashed grid the caustic gather reads.  With k
 * nearest photons (MRT_OPT_PHOTON_K > 0) `radius` bounds the squared
 * distance of the k-th, as the reference compares it; with k = 0 every
 * photon within distance `radius` is gathered.  n_photons = 0 turns
 * caustics off. */
int  mrt_build_photons(mrt_ctx *ctx, uint32_t n_photons, float radius);

/* The stored photons in sorted (grid) order: positions and powers are 3
 * floats each, ids the global photon index; *n_out gets the count. */
int  mrt_get_photons(mrt_ctx *ctx, uint32_t max, float *pos, float *power,
                     uint32_t *ids, uint32_t *n_out);

/* The caustic term (estimate_caustics_physically_based,
 * src/caustic/get_caustic.c:67-88) at n host points (3 floats each). */
int  mrt_caustic_at(mrt_ctx *ctx, uint32_t n, const float *pts, float *rgb);

/* Replaces compute_frame_gpu (src/gpu/use_shader.c:81-100): traces
 * n_samples paths per pixel, sample indices first_sample .. +n_samples-1,
 * each of at most max_depth segments, and adds them to the device
 * accumulation buffer.  Asynchronous on the context's stream. */
int  mrt_render(mrt_ctx *ctx, uint32_t first_sample, uint32_t n_samples,
                uint32_t max_depth);

/* No counterpart in the reference: allocates the wavefront queues that
 * renders of n_samples samples per call need (at most the path budget), so
 * that the first such mrt_render does not allocate (an interactive host's
 * first frame, or a timed render).  mrt_render allocates them itself when
 * this was not called. */
int  mrt_reserve(mrt_ctx *ctx, uint32_t n_samples);

/* Replaces get_result (src/gpu/use_shader.c:53-79): copies the device sums
 * after every render before it, adds them into the host-owned W*H*3 double
 * buffer, as add_fcolor_float3 does into screen.float_render, and zeroes
 * the device sums in stream order (the next render starts from zero).  A
 * frame of at least 2^20 floats is copied in chunks, each added on the
 * context's host threads as it arrives.  Like get_result, which adds the
 * one sample each dispatch wrote, every sample rendered since the previous
 * call reaches float_render exactly once: the compute_frame_gpu sequence
 * mrt_render(sample, 1) -> mrt_readback_add -> sample++ leaves k samples
 * after k ticks.  Returns once float_render holds them. */
int  mrt_readback_add(mrt_ctx *ctx, double *float_render);

/* Copies the device sums (W*H*3 floats) to out and leaves them in place. */
int  mrt_readback_f32(mrt_ctx *ctx, float *out);

/* Zeroes the device accumulation buffer. */
int  mrt_clear(mrt_ctx *ctx);

/* Render modes 1/2 of the reference (src/gpu/use_shader.c:69-71,
 * src/shaders/shader.comp:52-57) generalised: traces the primary rays of
 * sample `sample` and writes one MRT_AOV_* value per pixel to host `out`. */
int  mrt_render_aov(mrt_ctx *ctx, int aov, uint32_t sample, void *out);

/* Traces caller-supplied rays (host arrays of n origins and n unit
 * directions, 3 floats each) and returns closest-hit primitive ids and t;
 * used by parity tests and by hosts that pick objects with a ray
 * (src/edition_mod/select_obj.c:40-66 calls hit_register_all the same way). */
int  mrt_trace_rays(mrt_ctx *ctx, uint32_t n, const float *orig,
                    const float *dir, uint32_t *prim_id, float *t);

/* No counterpart in the reference.  mrt_trace_rays without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight interactive loops such as an object picker called per
 * click, harnesses that forbid hipMalloc): the rays and results are staged
 * in a fixed 8 MiB device array that the library holds on each GPU until
 * the process ends.  At most MRT_TRACE_NOALLOC_MAX rays per call (larger
 * n returns MRT_E_LIMIT).  All contexts on one GPU share that array: calls
 * that overlap on one GPU (from several host threads) race, so the caller
 * serialises them.  Same results as mrt_trace_rays. */
#define MRT_TRACE_NOALLOC_MAX 262144u
int  mrt_trace_rays_noalloc(mrt_ctx *ctx, uint32_t n, const float *orig,
                            const float *dir, uint32_t *prim_id, float *t);

int  mrt_set_option(mrt_ctx *ctx, int option, double value);
/* The 64-bit RNG key of the context's camera paths and photons; the
 * reference seeds its generator from RANDOM_SEED (src/math/random.c:17-30).
 * Samples are keyed by (seed, pixel, sample, bounce), so equal seeds give
 * equal images on any device. */
int  mrt_set_seed(mrt_ctx *ctx, uint64_t seed);
int  mrt_get_stats(mrt_ctx *ctx, mrt_stats *out);
int  mrt_reset_stats(mrt_ctx *ctx);
int  mrt_synchronize(mrt_ctx *ctx);

/* Replaces clean_shaders (src/gpu/clean_shaders.c:69-84). */
void mrt_destroy(mrt_ctx *ctx);

const char *mrt_last_error(const mrt_ctx *ctx);
const char *mrt_version(void);

/* ---- differentiable rendering -------------------------------------------
 *
 * No counterpart in the reference.  Parameters: per material the albedo
 * (rgb), the GGX roughness and the emission E (rgb, E = strength * colour;
 * setting it stores strength 1 and colour E), and the HDRI texels.  The
 * forward image is mrt_render's mean over n_samples; mrt_adjoint replays the
 * same paths (same RNG keys) and returns dL/dparam for a given dL/dI.  The
 * HDRI sampling tables stay those built by mrt_upload_hdri.  Requirements:
 * the MIS estimator when an HDRI is present, MRT_SCENE_LIGHT_EXACT when
 * lights are present, no photon map, depth <= 8, <= 64 materials. */

/* Replace material parameters; any array may be NULL (kept).  n must equal
 * the scene's material count. */
int  mrt_diff_set_materials(mrt_ctx *ctx, uint32_t n, const float *albedo,
                            const float *roughness, const float *emission);
/* Replace the HDRI texels (same size as uploaded; tables unchanged). */
int  mrt_diff_set_hdri(mrt_ctx *ctx, const float *rgb);
/* d_image: W*H*3 dL/dI of the mean image.  g_mat: n_materials * 7 floats
 * (albedo rgb, roughness, emission rgb); g_hdri: env_w*env_h*3 or NULL.
 * Albedo, emission and texel gradients are exact derivatives of the
 * replayed estimator; the roughness gradient is its score-function term
 * (an unbiased estimate of the derivative of the expected image). */
int  mrt_adjoint(mrt_ctx *ctx, uint32_t first_sample, uint32_t n_samples,
                 uint32_t max_depth, const float *d_image, float *g_mat,
                 float *g_hdri);

/* ---- batches: many independent scenes in one wavefront -----------------
 *
 * No counterpart in the reference, which renders one scene per process.
 * A batch holds scenes of any size and resolution, each with its own
 * sample count; mrt_batch_render traces all of their paths through one
 * shared set of queues, so small scenes fill the GPU together.  Paths are
 * numbered scene after scene (sample-major inside a scene) and every pass
 * takes the next `path budget` of them; the pass is captured once in a HIP
 * graph and replayed.  Samples use the same RNG keys as mrt_render with the
 * scene's seed, so a batched scene equals its plain render up to the order
 * of the fp32 sums.  Scenes must not use an HDRI or caustics. */
typedef struct mrt_batch mrt_batch;

/* Queue and launch shape of a batch (mrt_batch_autotune picks them). */
typedef struct mrt_batch_tuning {
    uint32_t path_budget;        /* paths per pass                           */
    uint32_t trace_blocks_per_cu;/* persistent blocks of the trace kernels   */
    uint32_t shade_blocks_per_cu;/* blocks of the shading kernels            */
    uint32_t use_graph;          /* replay a captured pass (1), or launch the
                                    passes directly, pipelined on the render
                                    lanes (0)                                */
    uint32_t trace_refill;       /* 1: the per-lane refill trace kernel on
                                    its resident grid, where the batch allows
                                    it (triangle / sphere scenes of one
                                    sidedness, at most 512 scenes); 0: the
                                    persistent trace kernel on
                                    trace_blocks_per_cu blocks per CU        */
    double   ms_per_mpath;       /* measured by the autotuner, else 0        */
} mrt_batch_tuning;

int  mrt_batch_create(int device, mrt_batch **out);
/* Copies the scene; renders it at width x height with `spp` samples per
 * pixel and RNG key `seed`.  *id receives the scene's index. */
int  mrt_batch_add_scene(mrt_batch *b, const mrt_scene *scene, uint32_t width,
                         uint32_t height, uint32_t spp, uint64_t seed,
                         uint32_t *id);
/* Builds every BVH (in parallel on the host) and uploads the batch. */
int  mrt_batch_commit(mrt_batch *b);
int  mrt_batch_autotune(mrt_batch *b, uint32_t max_depth,
                        mrt_batch_tuning *chosen);
int  mrt_batch_set_tuning(mrt_batch *b, const mrt_batch_tuning *t);
int  mrt_batch_get_tuning(const mrt_batch *b, mrt_batch_tuning *t);
/* Renders every scene's samples (accumulating onto earlier renders). */
int  mrt_batch_render(mrt_batch *b, uint32_t max_depth);
/* W*H*3 float sums of scene `id`. */
int  mrt_batch_readback(mrt_batch *b, uint32_t id, float *out);
int  mrt_batch_clear(mrt_batch *b);
int  mrt_batch_set_option(mrt_batch *b, int option, double value);
int  mrt_batch_get_stats(mrt_batch *b, mrt_stats *out);
void mrt_batch_destroy(mrt_batch *b);
const char *mrt_batch_last_error(const mrt_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* MRT_H */
