// Kernel-side structures shared by mrt_kernels.hip and mrt_api.hip.
#pragma once

#include "mrt_device.h"
#include "mrt_host.h"

namespace mrt {

#ifndef MRT_TRACE_BLOCK
#define MRT_TRACE_BLOCK 128
#endif
#ifndef MRT_TRACE_SS
#define MRT_TRACE_SS 16
#endif
constexpr int TRACE_BLOCK = MRT_TRACE_BLOCK;  // threads per trace block
static_assert(TRACE_BLOCK % 64 == 0, "trace blocks are whole waves");
constexpr int TRACE_SS = MRT_TRACE_SS;        // stack entries kept in LDS
constexpr int TRACE_OVF = STACK4_MAX - TRACE_SS;  // the rest, per lane
// The trace kernel that keeps the tree's top nodes in LDS: TOP_NODES nodes
// (the first levels, breadth-first at the front of the node array, 112 B
// each) and a TRACE_SS_TOP-entry stack per lane, 9,520 B per 128-thread
// block, so that 16 blocks (8 waves per SIMD) still fit in a CU's 160 KB.
#ifndef MRT_TOP_NODES
#define MRT_TOP_NODES 21
#endif
constexpr int TOP_NODES = MRT_TOP_NODES;
constexpr int TRACE_SS_TOP = 14;
static_assert(TOP_NODES <= (int)MRT_BVH4_TOP, "cached nodes are laid out first");
// Scenes whose roots the batched refill trace kernel keeps in LDS (2 KB
// beside its 7 KB stack and 2.3 KB tree top: 12 blocks, 6 waves per SIMD,
// still fit a CU); larger batches take k_trace.
constexpr uint32_t BATCH_LDS_SCENES = 512;
// per-lane columns of the overflow buffer: enough for either stack split
constexpr int TRACE_OVF_BUF = STACK4_MAX - (TRACE_SS_TOP < TRACE_SS ? TRACE_SS_TOP
                                                                    : TRACE_SS);

// Device buffer for the traversal stack entries beyond the LDS part, one
// column per lane of a persistent trace grid of up to `blocks` blocks
// (TRACE_OVF * blocks * TRACE_BLOCK uint32).  With it the trace kernels use
// no scratch memory, whose set-up the runtime repeats per launch when the
// whole grid's scratch is large.
struct StackOvf {
    uint32_t *buf;
    uint32_t  blocks;
};

// Work distribution of the persistent trace kernels.  The items form
// chunks of CHUNK; wave w of the grid takes chunk w first, without an
// atomic, and the chunks beyond the grid's waves form NSET interleaved sets
// (the k-th chunk of set s is waves + s + NSET k), each claimed through its
// own cursor.  All waves thus stay on one narrow, coherent window of the
// queue (neighbouring pixels share BVH nodes in every L2), while atomics on
// one address, which the memory system serialises (~12 ns each on MI355X,
// 79 us for one per wave of a 6656-wave grid), are split 8 ways and skipped
// where the static chunks cover the queue.
constexpr uint32_t CHUNK = 64;     // items per chunk: one per lane
#ifndef MRT_NSET
#define MRT_NSET 8
#endif
constexpr int NSET = MRT_NSET;

// device counters (uint32), each on its own 128-byte line so that atomics
// on one never queue behind atomics on another.  The shadow-ray and
// pending-sum counts come in two copies indexed by the bounce's parity: the
// shadow rays of bounce b are traced and resolved while bounce b + 1 runs.
constexpr int CTR_LINE = 32;
enum : int {
    CTR_CUR = 0,                   // paths in the current queue
    CTR_NEXT = 1 * CTR_LINE,       // paths appended to the next queue
    CTR_SHADOW = 2 * CTR_LINE,     // shadow rays of an even / odd bounce
    CTR_PEND = 4 * CTR_LINE,       // pending light sums of an even / odd bounce
    CTR_PASS = 6 * CTR_LINE,       // paths of the current pass (batches)
    CTR_SETS = 7 * CTR_LINE,       // NSET work cursors of k_trace
    CTR_COUNT = (7 + NSET) * CTR_LINE
};
__host__ __device__ constexpr int ctr_shadow(uint32_t bounce)
{
    return CTR_SHADOW + (int)(bounce & 1u) * CTR_LINE;
}
__host__ __device__ constexpr int ctr_pend(uint32_t bounce)
{
    return CTR_PEND + (int)(bounce & 1u) * CTR_LINE;
}

#ifndef MRT_SHADE_BLOCK
#define MRT_SHADE_BLOCK 256
#endif
// threads per block of k_shade (256: 843 vs 837 Msamples/s with 512 on the
// 1080p mesh workload)
constexpr int SHADE_BLOCK = MRT_SHADE_BLOCK;

// device statistics (uint64)
enum : int {
    STAT_RAYS_CLOSEST = 0,
    STAT_RAYS_SHADOW = 1,
    STAT_NODES_CLOSEST = 2,
    STAT_PRIMS_CLOSEST = 3,
    STAT_NODES_SHADOW = 4,
    STAT_PRIMS_SHADOW = 5,
    STAT_NONFINITE = 6,
    STAT_GATHER_QUERIES = 7,   // caustic gathers (pending records)
    STAT_GATHER_PHOTONS = 8,   // photon records they examined
    STAT_GATHER_CELLS = 9,     // nonempty buckets whose bounds they read
    STAT_GATHER_LOOKUPS = 10,  // grid cells they looked up (occupancy bits)
    STAT_COUNT = 16
};

// Photon map in HBM, sorted by hash bucket (hash_cell of
// src/gpu/convert_caustic_utils.c:15-24 modulo `table`), photon id second.
// The map's records are interleaved, photon j at rec[2 j] = (position,
// photon id) and rec[2 j + 1] = (power, 0): a bucket's few photons then
// share cache lines instead of touching one line in each of two arrays.
struct PhotonMapDev {
    const float4   *rec;      // 2 per photon, in bucket order
    const uint2    *range;    // per bucket: first photon, one past the last
    const uint32_t *occ;      // one bit per bucket: holds photons
    uint32_t n, table;
    float    cell;            // grid cell edge >= the search distance
    uint32_t k;               // k nearest (0: every photon within radius)
    float    radius, norm;    // --caustic=N;RADIUS;NORM
};

// Photon emission: n_per photons per point light, then n_per from the
// environment through the disk of the scene's bounding sphere (sc, sr).
struct PhotonEmit {
    uint32_t n_per, n_point, has_env, n_total;
    float    sc[3], sr;
    uint32_t point_idx[64];
};

struct FrameParams {
    float    cam_o[3], p00[3], du[3], dv[3], disk_u[3], disk_v[3];
    float    cam_pos[3];
    float    bg_clamped[3];   // clamp(ambient colour / 255 * ratio, 0, 16)
    float    ior_all;
    uint32_t lens;
    uint32_t W, H, WH;
    uint32_t sample0;         // sample index of batch slot 0
    uint32_t slot0;           // path index (sample-major over the frame) of
                              // slot 0 of a render lane
    uint32_t bounce;
    uint32_t last_bounce;
    uint32_t n_lights;
    uint32_t one_sided;
    uint32_t flags;           // MRT_SCENE_*
    uint64_t seed;
    // environment (HDRI), present when has_env
    const float4 *env_tex;    // env_w * env_h texels, row 0 at the nadir
    const float  *env_pdf;    // reference tables: pdf_joint, inverse CDFs
    const float  *env_marg_inv;
    const float  *env_cond_inv;
    const float  *mis_marg;   // MIS tables (mrt_hdri_mis)
    const float  *mis_cond;
    const float  *mis_pcell;
    // guide tables of the MIS CDFs (cdf_search_guided): guide_m + 1 entries
    // for the marginal, guide_c + 1 per conditional row; powers of two
    const uint32_t *mis_marg_guide;
    const uint32_t *mis_cond_guide;
    uint32_t guide_m, guide_c;
    uint32_t env_w, env_h;
    float    env_scale;       // 2^(exposure + ratio - 1)
    uint32_t has_env;
    uint32_t estimator;       // MRT_EST_*
    PhotonMapDev pm;          // caustic photon map, present when has_photons
    uint32_t has_photons;
    uint32_t light0, plane0;  // first light / plane of the scene in the
                              // shared arrays (0 for a single scene)
    uint32_t textured;        // some material has a texture slot
};

// One scene of a batch: its frame, where its BVH and planes sit in the
// shared arrays, and its share of the global path index space
// (W*H*spp paths, sample-major).
struct BatchScene {
    uint32_t    root, n_nodes, n_planes;
    uint32_t    accum0;       // first float of its W*H*3 sums
    unsigned long long path0; // first global path index
    unsigned long long paths; // W*H*spp
    uint32_t    light0, flags;   // fp.light0, fp.flags: in the first cache
                                 // line, for the light sums of k_resolve
    FrameParams fp;
};

// Device-side cursor of a batch: the first global path of the current
// wavefront pass and the total.  Kernels read it, so one captured graph
// of a pass replays without any host-side argument change.
struct BatchCtl {
    unsigned long long base, total;
    uint32_t           n_scenes, pad;
};

constexpr uint32_t TAG_ENV = 0x40000000u;   // shadow-ray tag of the HDRI ray

// The adjoint's tape: float4 rows of tape_P entries (one per path slot).
// Rows 0 and 1 end each path: (escape radiance before thr, its texel
// weight) and (texel index, vertex count); vertex b has rows
// tape_row(b, 0..3): (thr in, material), (scatter factor, GGX score),
// (albedo derivative, environment-sample weight), (next-event radiance,
// environment-sample texel).
__host__ __device__ constexpr size_t tape_row(uint32_t bounce, uint32_t k)
{
    return 2 + 4 * (size_t)bounce + k;
}

struct ShadeBufs {
    const float4 *q_o, *q_d, *q_t, *hit;
    float4       *qn_o, *qn_d, *qn_t;
    float4       *acc;
    const float4 *planes, *prims, *mats, *lights;
    const float  *nrm;
    float4       *s_o, *s_d;
    float2       *s_c;
    uint8_t      *vis;
    float4       *p_a;
    uint4        *p_b;
    float4       *p_c;
    float4       *p_x;         // hit point of each pending record (photons)
    float4       *p_k;         // caustic term of each pending record
    uint32_t      cap_paths;   // capacity of the path / pending arrays
    uint32_t      cap_shadow;  // capacity of the shadow queue
    // batches only: the scene table and (scene, local path index) per slot
    const BatchScene *bs;
    uint2            *slot_info;
    TexDev            tx;      // textures (when FrameParams::textured)
    // the adjoint's recording replay only (else null)
    float4           *tape;
    uint32_t          tape_P;
};

// Moves one scene's nodes / records, already copied into a batch's shared
// arrays, to their place there: links by node0, leaf ranges by prim0,
// materials by mat0.
hipError_t launch_rebase(hipStream_t st, float4 *nodes, uint32_t n_nodes,
                         float4 *prims, float *nrm, uint32_t n_prims,
                         uint32_t node0, uint32_t prim0, uint32_t mat0);
// tex[i] = (rgb[3i], rgb[3i + 1], rgb[3i + 2], 0) for i < n.
hipError_t launch_rgb_to_float4(hipStream_t st, const float *rgb, float4 *tex,
                                size_t n);
// Builds the HDRI's reference and MIS tables from its texels (tex, w * h)
// on the device, bit for bit as mrt_hdri_build_tables / _mis do on the
// host.  sc = 2^exposure (or 1); band[y] = sin(e_hi) - sin(e_lo) of the h-1
// MIS rows; scratch holds 3 h + 2 + w h doubles.
// Guide tables of the MIS CDFs built by launch_env_tables: gm (Gm + 1
// entries) for the marginal, gc ((h - 1) rows of Gc + 1) for the
// conditional rows; Gm, Gc powers of two.
hipError_t launch_env_guides(hipStream_t st, const float *mis_marg,
                             const float *mis_cond, uint32_t w, uint32_t h,
                             uint32_t Gm, uint32_t Gc, uint32_t *gm,
                             uint32_t *gc);
hipError_t launch_env_tables(hipStream_t st, const float4 *tex, uint32_t w,
                             uint32_t h, double sc, const double *band,
                             double *scratch, float *pdf, float *marg_inv,
                             float *cond_inv, float *mis_marg, float *mis_cond,
                             float *mis_pcell);
hipError_t launch_raygen(hipStream_t st, const FrameParams &fp, float4 *q_o,
                         float4 *q_d, float4 *q_t, float4 *acc, uint32_t P,
                         uint32_t *ctr);
// One launch of k_trace: the closest-hit rays of the current queue (when
// closest) and the shadow rays counted at ctr[shadow_ctr] (when >= 0).
struct TraceWork {
    const float4 *q_o, *q_d;      // closest-hit rays: origin | ao, dir | slot
    const float4 *q_t;            // a batch's: .w holds the ray's scene
    float4       *hit;            // (t, code, u, v) per closest-hit ray
    const float4 *s_o, *s_d;      // shadow rays: origin | tmax, dir | tag
    uint8_t      *vis;            // 1: the shadow ray reaches its light
    uint32_t      cap_shadow;     // shadow queue capacity
    int           shadow_ctr;     // counter of the shadow rays, or -1
    bool          closest;        // trace ctr[CTR_CUR] closest-hit rays
};
hipError_t launch_trace(hipStream_t st, uint32_t blocks, bool count,
                        const SceneDev &S, const StackOvf &ovf,
                        const TraceWork &w, uint32_t *ctr,
                        unsigned long long *stats, bool top, bool refill,
                        uint32_t refill_blocks);
hipError_t launch_shade(hipStream_t st, uint32_t blocks, const FrameParams &fp,
                        const ShadeBufs &B, uint32_t *ctr);
hipError_t launch_resolve(hipStream_t st, uint32_t blocks,
                          const FrameParams &fp, const ShadeBufs &B,
                          uint32_t *ctr);
// After the shade of bounce b: the next queue becomes current, and the
// shadow / pending counts of bounce b - 1 (resolved by now) are added to
// the statistics and cleared for bounce b + 1.
hipError_t launch_advance(hipStream_t st, uint32_t *ctr,
                          unsigned long long *stats, uint32_t cap,
                          uint32_t bounce);
hipError_t launch_accumulate(hipStream_t st, const float4 *acc, float *accum,
                             uint32_t WH, uint32_t nb,
                             unsigned long long *stats);
hipError_t launch_aov(hipStream_t st, const FrameParams &fp, const SceneDev &S,
                      const float4 *prims, const float *nrm, uint32_t sample,
                      int aov, void *out);
hipError_t launch_trace_rays(hipStream_t st, const SceneDev &S,
                             const float4 *prims, uint32_t n,
                             const float *orig, const float *dir,
                             uint32_t *prim_id, float *t);
int trace_blocks_per_cu();
// resident blocks per CU of the per-lane refill trace kernel (single
// scenes, or batch = true)
int refill_blocks_per_cu(bool batch = false);

// Path-replay adjoint: dL/dI per pixel in, gradients out.
constexpr int ADJ_MAXD = 8;      // deepest path the adjoint replays
constexpr int ADJ_MAXMAT = 64;   // materials with gradients (LDS sums)
constexpr int ADJ_NG = 7;        // per material: albedo rgb, roughness, emission rgb
struct AdjBufs {
    const float *dimg;           // W*H*3, dL/dI of the mean image
    float       *g_mat;          // n_mat * ADJ_NG
    float       *g_tex;          // env_w*env_h*3, or null
    uint32_t     n_mat;
    float        inv_spp;
};
// Rows 0-1 of the tape for P paths: nothing escaped, max_depth vertices.
hipError_t launch_tape_init(hipStream_t st, float4 *tape, uint32_t P,
                            uint32_t max_depth);
// The backward pass over the tape of a recorded replay of P paths.
hipError_t launch_adjoint_back(hipStream_t st, uint32_t blocks,
                               const FrameParams &fp, const ShadeBufs &B,
                               const AdjBufs &A, uint32_t P);

// batched scenes (the same stages over the scene table of a batch)
// A pass starts at global path `base`, or at the device cursor ctl->base
// when base is BASE_FROM_CTL (a captured pass replayed as a graph).
constexpr unsigned long long BASE_FROM_CTL = ~0ull;
// Camera rays of a render lane's part [o0, o0 + Pl) of a pass of P paths
// (B, q_* and ctr already that lane's), whose paths lie in scenes
// [s_lo, s_hi]; lane 0 (o0 = 0) also records the pass's path count for
// k_baccumulate.
hipError_t launch_braygen(hipStream_t st, const ShadeBufs &B,
                          const BatchCtl *ctl, unsigned long long base,
                          uint32_t s_lo, uint32_t s_hi, float4 *q_o, float4 *q_d,
                          float4 *q_t, uint32_t o0, uint32_t Pl, uint32_t P,
                          uint32_t *ctr);
// refill_blocks > 0: the batched refill kernel on that grid (triangle /
// sphere scenes of one sidedness, at most BATCH_LDS_SCENES of them), with
// nodes [top_base, top_base + ntop) (ntop <= TOP_NODES, the top of one
// scene's tree) staged in LDS; else k_trace on `blocks`.
hipError_t launch_btrace(hipStream_t st, uint32_t blocks, bool count,
                         const SceneDev &S, const StackOvf &ovf,
                         const ShadeBufs &B, const TraceWork &w,
                         uint32_t *ctr, unsigned long long *stats,
                         uint32_t refill_blocks, uint32_t n_scenes,
                         uint32_t top_base, uint32_t ntop);
hipError_t launch_bshade(hipStream_t st, uint32_t blocks,
                         const FrameParams &fp, const ShadeBufs &B,
                         uint32_t *ctr);
hipError_t launch_bresolve(hipStream_t st, uint32_t blocks,
                           const FrameParams &fp, const ShadeBufs &B,
                           uint32_t *ctr);
hipError_t launch_baccumulate(hipStream_t st, const ShadeBufs &B,
                              const BatchCtl *ctl, unsigned long long base,
                              float *accum, uint32_t P, uint32_t *ctr,
                              unsigned long long *stats);
hipError_t launch_bnext(hipStream_t st, BatchCtl *ctl, uint32_t P);

hipError_t launch_photon_emit(hipStream_t st, const FrameParams &fp,
                              const SceneDev &S, const PhotonEmit &pe,
                              const ShadeBufs &B, float4 *pos, float4 *pow,
                              uint32_t *counter, uint32_t cap);
hipError_t launch_photon_keys(hipStream_t st, const float4 *pos, uint32_t n,
                              float cell, uint32_t table,
                              unsigned long long *keys, uint32_t *vals);
hipError_t launch_photon_reorder(hipStream_t st, const float4 *pos,
                                 const float4 *pow,
                                 const unsigned long long *keys,
                                 const uint32_t *vals, uint32_t n,
                                 float4 *rec, uint2 *range, uint32_t *occ);
hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, size_t n);
hipError_t launch_caustic(hipStream_t st, uint32_t blocks,
                          const FrameParams &fp, const ShadeBufs &B,
                          uint32_t *ctr, bool count, unsigned long long *stats);
hipError_t launch_caustic_at(hipStream_t st, const PhotonMapDev &pm,
                             uint32_t n, const float *pts, float *out);

}  // namespace mrt
