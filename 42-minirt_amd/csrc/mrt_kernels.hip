// Wavefront path-tracing kernels for gfx950 (MI355X).
//
// One pass of a batch of paths (nb samples x W*H pixels) runs as
//   raygen -> { trace -> shade -> shadow -> resolve -> advance } x depth
//   -> accumulate
// with every queue in HBM as structure-of-arrays of 16-byte words:
//   path queue   q_o = (origin, ao of the last material)
//                q_d = (direction, path slot)
//                q_t = (throughput, MIS pdf of the last BSDF sample; in a
//                       batch, which has no HDRI, the path's scene)
//   hit record   hit = (t, hit code, barycentric u, v)
//   shadow queue s_o = (origin, t limit), s_d = (direction, light tag),
//                s_c = (n.l, (r.v)^64), vis (1 = unoccluded)
//   pending      p_a = (albedo * throughput, slot), p_b = (next-queue index,
//                first shadow ray, shadow-ray count, flags),
//                p_c = (HDRI radiance term, 1 - cos theta)
//   per-slot sums acc (float4), summed per pixel in sample order at the end.
// A path owns its slot, so every sum is race-free and deterministic; queue
// positions come from wave-aggregated atomics.
#include "mrt_kernels.h"

namespace mrt {

// ---- wave-aggregated allocation (wave64) ----------------------------------
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Every lane of the wave calls this (n = 0 for lanes that want nothing);
// returns the first of the n entries reserved for this lane.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t *counter, uint32_t n)
{
    uint32_t lane = lane_id();
    uint32_t incl = n;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t o = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off)
            incl += o;
    }
    uint32_t total = __shfl(incl, 63, 64);
    uint32_t base = 0;
    if (lane == 63 && total)
        base = atomicAdd(counter, total);
    base = __shfl(base, 63, 64);
    return base + incl - n;
}

// Block-wide version for three counters at once: one atomic per counter per
// block instead of one per wave.  Every thread of the block calls it.
__device__ __forceinline__ void block_alloc3(uint32_t *ca, uint32_t *cb,
                                             uint32_t *cc, uint32_t na,
                                             uint32_t nb, uint32_t nc,
                                             uint32_t &oa, uint32_t &ob,
                                             uint32_t &oc)
{
    constexpr int NW = SHADE_BLOCK / 64;
    __shared__ uint32_t s_w[NW][3];
    __shared__ uint32_t s_base[3];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t ia = na, ib = nb, ic = nc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t xa = __shfl_up(ia, off, 64), xb = __shfl_up(ib, off, 64),
                 xc = __shfl_up(ic, off, 64);
        if (lane >= (uint32_t)off) {
            ia += xa;
            ib += xb;
            ic += xc;
        }
    }
    if (lane == 63) {
        s_w[wave][0] = ia;
        s_w[wave][1] = ib;
        s_w[wave][2] = ic;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const uint32_t k = threadIdx.x;
        uint32_t tot = 0;
        for (int w = 0; w < NW; w++) {
            uint32_t v = s_w[w][k];
            s_w[w][k] = tot;
            tot += v;
        }
        uint32_t *c = k == 0 ? ca : (k == 1 ? cb : cc);
        s_base[k] = tot ? atomicAdd(c, tot) : 0u;
    }
    __syncthreads();
    oa = s_base[0] + s_w[wave][0] + ia - na;
    ob = s_base[1] + s_w[wave][1] + ib - nb;
    oc = s_base[2] + s_w[wave][2] + ic - nc;
    __syncthreads();
}

// Next chunk of a persistent trace kernel (the scheme is described with
// CHUNK): nch chunks in all, W = waves in the grid; set / tries keep the
// wave's place among the NSET cursors at `sets` (one per CTR_LINE).  The
// wave's own set is claimed from directly; the others are read first, so a
// used-up set costs a load, not an atomic.  Returns false when all chunks
// are taken.  Wave-uniform: every lane gets lane 0's answer.
__device__ __forceinline__ bool next_chunk(uint32_t *sets, uint32_t nch,
                                           uint32_t W, uint32_t &set,
                                           uint32_t &tries, uint32_t &chunk)
{
    if (nch <= W)
        return false;
    const uint32_t nd = nch - W;
    while (tries < (uint32_t)NSET) {
        const uint32_t cnt = set < nd ? (nd - set + NSET - 1) / NSET : 0u;
        uint32_t *cur = sets + set * CTR_LINE;
        uint32_t k = cnt;
        if (lane_id() == 0 && cnt) {
            uint32_t seen = tries == 0 ? 0u
                : __hip_atomic_load(cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            k = seen < cnt ? atomicAdd(cur, 1u) : seen;
        }
        k = __shfl(k, 0, 64);
        if (k < cnt) {
            chunk = W + set + NSET * k;
            return true;
        }
        set = set + 1 == (uint32_t)NSET ? 0u : set + 1;
        tries++;
    }
    return false;
}

// Resets the work cursors of k_trace (one thread).
__device__ __forceinline__ void reset_sets(uint32_t *ctr)
{
    for (int k = 0; k < NSET; k++)
        ctr[CTR_SETS + k * CTR_LINE] = 0;
}

// Clears every queue counter of a new pass (one thread).
__device__ __forceinline__ void reset_queues(uint32_t *ctr, uint32_t paths)
{
    ctr[CTR_CUR] = paths;
    ctr[CTR_NEXT] = 0;
    for (uint32_t p = 0; p < 2; p++) {
        ctr[ctr_shadow(p)] = 0;
        ctr[ctr_pend(p)] = 0;
    }
    reset_sets(ctr);
}

__device__ __forceinline__ void wave_add_u64(unsigned long long *dst, uint32_t v)
{
    uint64_t s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor((unsigned long long)s, off, 64);
    if (lane_id() == 0 && s)
        atomicAdd(dst, (unsigned long long)s);
}

// ---- camera rays (calc_one_sample_task, src/ray_tracing/calc_sample.c:44-58)
__device__ __forceinline__ void primary_ray(const FrameParams &fp,
                                            uint32_t pixel, uint32_t sample,
                                            V3 &o, V3 &d)
{
    U4 r = rng4(fp.seed, pixel, sample, 0u, 0u);
    V3 op = v3(fp.cam_o[0], fp.cam_o[1], fp.cam_o[2]);
    if (fp.lens) {
        // concentric map of block 0 dims 2,3 (random_in_unit_disk's law)
        float sx = 2.0f * r.z - 1.0f, sy = 2.0f * r.w - 1.0f;
        float lx = 0.0f, ly = 0.0f;
        if (sx != 0.0f || sy != 0.0f) {
            float rr, phi;
            if (fabsf(sx) > fabsf(sy)) {
                rr = sx;
                phi = 0.78539816339744831f * (sy / sx);
            } else {
                rr = sy;
                phi = 1.57079632679489662f - 0.78539816339744831f * (sx / sy);
            }
            lx = rr * cosf(phi);
            ly = rr * sinf(phi);
        }
        op = v3(op.x + (fp.disk_u[0] * lx + fp.disk_v[0] * ly),
                op.y + (fp.disk_u[1] * lx + fp.disk_v[1] * ly),
                op.z + (fp.disk_u[2] * lx + fp.disk_v[2] * ly));
    }
    float fx = (float)(pixel % fp.W) + r.x;
    float fy = (float)(pixel / fp.W) + r.y;
    V3 dd = v3(((fp.p00[0] + fp.du[0] * fx) + fp.dv[0] * fy) - op.x,
               ((fp.p00[1] + fp.du[1] * fx) + fp.dv[1] * fy) - op.y,
               ((fp.p00[2] + fp.du[2] * fx) + fp.dv[2] * fy) - op.z);
    float len = sqrtf((dd.x * dd.x + dd.y * dd.y) + dd.z * dd.z);
    o = op;
    d = len > 0.0f ? v3(dd.x / len, dd.y / len, dd.z / len) : v3(0.f, 0.f, 0.f);
}

__global__ void __launch_bounds__(256)
k_raygen(FrameParams fp, float4 *__restrict__ q_o, float4 *__restrict__ q_d,
         float4 *__restrict__ q_t, float4 *__restrict__ acc, uint32_t P,
         uint32_t *__restrict__ ctr)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        reset_queues(ctr, P);
    if (i >= P)
        return;
    const uint32_t g = i + fp.slot0;
    uint32_t pixel = g % fp.WH, sample = fp.sample0 + g / fp.WH;
    V3 o, d;
    primary_ray(fp, pixel, sample, o, d);
    q_o[i] = make_float4(o.x, o.y, o.z, 1.0f);
    q_d[i] = make_float4(d.x, d.y, d.z, __uint_as_float(i));
    q_t[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    acc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The scene of a path of a batch: shared node / primitive arrays, its own
// root, planes and flags.
__device__ __forceinline__ SceneDev scene_of(const SceneDev &S,
                                             const BatchScene *bs, uint32_t sid)
{
    const BatchScene &b = bs[sid];
    SceneDev r = S;
    r.root = b.root;
    r.n_nodes = b.n_nodes;
    r.planes = S.planes + 2 * (size_t)b.fp.plane0;
    r.n_planes = b.n_planes;
    r.one_sided = b.fp.one_sided;
    return r;
}

// Shadow-ray tag: scene of a batch (bits 18..29), light (12..17), ray of a
// soft light (0..11); TAG_ENV marks the environment ray.
__device__ __forceinline__ uint32_t tag_scene(uint32_t tag) { return (tag >> 18) & 0xFFFu; }
__device__ __forceinline__ uint32_t tag_light(uint32_t tag) { return (tag >> 12) & 63u; }

// ---- closest-hit and shadow rays (persistent, wave-granular) ----------------
// One launch traces the closest-hit rays of the current bounce and the
// shadow rays of the previous one: the two sets are independent, and in one
// grid the longest rays of one overlap the work of the other (a launch of
// its own waits for its longest ray: ~300 traversal steps late in a path,
// whatever the number of rays).  Chunks [0, ncc) hold closest-hit rays,
// [ncc, ncc + ncs) shadow rays; a chunk is one kind, so the branch is
// wave-uniform.
template <bool COUNT, bool BATCH, bool EXT, int SS = TRACE_SS, bool TOP = false>
__device__ __forceinline__ void
trace_work(const SceneDev &S, const TraceWork &w, uint32_t *__restrict__ ctr,
           unsigned long long *__restrict__ stats,
           const BatchScene *__restrict__ bs, const uint2 *__restrict__ slot_info,
           uint32_t *__restrict__ ovf_buf, uint32_t *s_code,
           const float4 *top = nullptr, uint32_t ntop = 0)
{
    const uint32_t nc = w.closest ? ctr[CTR_CUR] : 0u;
    const uint32_t ns = w.shadow_ctr >= 0 ? min(ctr[w.shadow_ctr], w.cap_shadow) : 0u;
    uint32_t *lc = s_code + threadIdx.x;
    uint32_t *ovf = ovf_buf + blockIdx.x * TRACE_BLOCK + threadIdx.x;
    const uint32_t os = gridDim.x * TRACE_BLOCK;
    uint32_t nodes_c = 0, prims_c = 0, nodes_s = 0, prims_s = 0;
    const uint32_t W = gridDim.x * (TRACE_BLOCK / 64);
    const uint32_t ncc = nc / CHUNK + (nc % CHUNK != 0);
    const uint32_t nch = ncc + ns / CHUNK + (ns % CHUNK != 0);
    uint32_t chunk = blockIdx.x * (TRACE_BLOCK / 64) + (threadIdx.x >> 6);
    uint32_t set = blockIdx.x % NSET, tries = 0;
    for (bool more = chunk < nch; more;
         more = next_chunk(&ctr[CTR_SETS], nch, W, set, tries, chunk)) {
        if (chunk < ncc) {
            const uint32_t i = chunk * CHUNK + lane_id();
            if (i < nc) {
                float4 o4 = w.q_o[i], d4 = w.q_d[i];
                float t, bu, bv;
                uint32_t code;
                if (BATCH) {
                    const SceneDev Si =
                        scene_of(S, bs, slot_info[__float_as_uint(d4.w)].x);
                    closest_hit<COUNT, TRACE_SS, EXT>(Si, xyz(o4), xyz(d4), lc,
                                                      TRACE_BLOCK, ovf, os, t,
                                                      code, bu, bv, nodes_c,
                                                      prims_c);
                } else {
                    closest_hit<COUNT, SS, EXT, TOP>(S, xyz(o4), xyz(d4), lc,
                                                     TRACE_BLOCK, ovf, os, t,
                                                     code, bu, bv, nodes_c,
                                                     prims_c, top, ntop);
                }
                w.hit[i] = make_float4(t, __uint_as_float(code), bu, bv);
            }
        } else {
            const uint32_t i = (chunk - ncc) * CHUNK + lane_id();
            if (i < ns) {
                float4 o4 = w.s_o[i], d4 = w.s_d[i];
                bool occ;
                if (BATCH) {
                    const SceneDev Si =
                        scene_of(S, bs, tag_scene(__float_as_uint(d4.w)));
                    occ = occluded<COUNT, TRACE_SS, EXT>(Si, xyz(o4), xyz(d4), o4.w,
                                                         lc, TRACE_BLOCK, ovf, os,
                                                         nodes_s, prims_s);
                } else {
                    occ = occluded<COUNT, SS, EXT, TOP>(S, xyz(o4), xyz(d4), o4.w,
                                                        lc, TRACE_BLOCK, ovf, os,
                                                        nodes_s, prims_s, top,
                                                        ntop);
                }
                w.vis[i] = occ ? 0 : 1;
            }
        }
    }
    if (COUNT) {
        wave_add_u64(&stats[STAT_NODES_CLOSEST], nodes_c);
        wave_add_u64(&stats[STAT_PRIMS_CLOSEST], prims_c);
        wave_add_u64(&stats[STAT_NODES_SHADOW], nodes_s);
        wave_add_u64(&stats[STAT_PRIMS_SHADOW], prims_s);
    }
}

// Scenes of triangles and spheres: 8 waves per SIMD (63 VGPRs without
// spills, with the 8 KB LDS stack of a 128-thread block); 1 % more
// Msamples/s on the 1080p mesh workload than the compiler's own choice of
// 75 VGPRs and 6 waves.
template <bool COUNT, bool BATCH>
__global__ void __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(8, 8)))
k_trace(SceneDev S, TraceWork w, uint32_t *__restrict__ ctr,
        unsigned long long *__restrict__ stats,
        const BatchScene *__restrict__ bs, const uint2 *__restrict__ slot_info,
        uint32_t *__restrict__ ovf_buf)
{
    __shared__ uint32_t s_code[TRACE_SS * TRACE_BLOCK];
    trace_work<COUNT, BATCH, false>(S, w, ctr, stats, bs, slot_info, ovf_buf, s_code);
}

// k_trace with the tree's top nodes staged in LDS once per block (the
// persistent grid's blocks live for the whole launch), for single scenes.
template <bool COUNT>
__global__ void __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(8, 8)))
k_trace_top(SceneDev S, TraceWork w, uint32_t *__restrict__ ctr,
            unsigned long long *__restrict__ stats, uint32_t *__restrict__ ovf_buf)
{
    __shared__ uint32_t s_code[TRACE_SS_TOP * TRACE_BLOCK];
    __shared__ float4 s_top[TOP_NODES * 7];
    const uint32_t ntop = min((uint32_t)TOP_NODES, S.n_nodes);
    for (uint32_t k = threadIdx.x; k < ntop * 7; k += TRACE_BLOCK)
        s_top[k] = S.nodes[(size_t)(k / 7) * 8 + k % 7];
    __syncthreads();
    trace_work<COUNT, false, false, TRACE_SS_TOP, true>(
        S, w, ctr, stats, nullptr, nullptr, ovf_buf, s_code, s_top, ntop);
}

// k_trace_top with per-lane refill (scenes of triangles and spheres: one,
// or with BATCH a batch's).  A wave takes 64-item chunks as k_trace does (closest-hit rays
// and then shadow rays numbered as one item space), but a lane whose ray is
// done takes the next item of the wave's chunk instead of idling until the
// wave's slowest ray ends: secondary rays of one wave differ several times
// over in traversal length.  Refills happen when REFILL_MIN lanes are idle
// (or all of them), so their loads and set-up run for many lanes at once.
// The traversal is traverse()'s, with the ray's kind per lane: closest-hit
// rays sort the hit children, shadow rays keep slot order and stop at the
// first hit, as there.  An iteration is a node step for the lanes at inner
// nodes, then a leaf step for the wave: the next one or two records of
// every lane at a leaf are tested on the wave's lanes, one record per lane
// (the rest of a leaf stays the lane's code for the next iteration).
constexpr uint32_t REFILL_MIN = 16;

// 6 waves per SIMD (80 VGPRs; one register spilled outside the step loop):
// 1,379 against 1,225 Msamples/s for the 8-wave k_trace_top on the 1080p
// mesh workload (one lane) before the wave-wide leaf step; at 8 waves it
// spilled 68 B per lane and halved the rate, at 5 it hid less latency.
// REFILL_MIN 16 against 8 / 24: 1,367 / 1,364.  The wave-wide leaf step
// against each lane testing its own pair: 1,580 -> 1,698 Msamples/s (the
// leaf code ran once per record slot for the few lanes at leaves, about as
// many instructions as the node step).
//
// BATCH: the rays of many scenes (triangles and spheres, one sidedness for
// all) whose trees share the node array.  A ray starts at its scene's root,
// read from a table of the scenes' roots in LDS: a closest-hit ray's scene
// comes with the ray (q_t.w, written by the batch's raygen and shade), a
// shadow ray's in its tag.  The ray state is the single scene's, so the
// kernel keeps its registers; the LDS copy of a tree's top holds the top of
// one scene's tree (the host names the scene with most of the pass's
// paths), and a ray that ends reads its scene's planes only when the batch
// has planes.
template <bool COUNT, bool BATCH>
__global__ void __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(6, 6)))
k_trace_refill(SceneDev S, TraceWork w, uint32_t *__restrict__ ctr,
               unsigned long long *__restrict__ stats, uint32_t *__restrict__ ovf_buf,
               const BatchScene *__restrict__ bs, uint32_t n_scenes,
               uint32_t top_base, uint32_t ntop)
{
    constexpr int SS = TRACE_SS_TOP;
    __shared__ uint32_t s_code[SS * TRACE_BLOCK];
    __shared__ float4 s_top[TOP_NODES * 7];
    __shared__ uint32_t s_root[BATCH ? BATCH_LDS_SCENES : 1];
    // per wave: the owner (lane | record << 6) of each of its lanes' leaf
    // tests in the current step
    __shared__ uint32_t s_job[TRACE_BLOCK];
    // the top ntop nodes of one tree, nodes [top_base, top_base + ntop):
    // a single scene's (top_base 0), or those of the scene that holds most
    // of a batch pass's paths (each scene's tree starts at its root, top
    // levels first)
    ntop = min(ntop, (uint32_t)TOP_NODES);
    for (uint32_t k = threadIdx.x; k < ntop * 7; k += TRACE_BLOCK)
        s_top[k] = S.nodes[(size_t)(top_base + k / 7) * 8 + k % 7];
    if (BATCH) {
        for (uint32_t k = threadIdx.x; k < n_scenes; k += TRACE_BLOCK)
            s_root[k] = bs[k].n_nodes ? bs[k].root : (uint32_t)INT32_MIN;
    }
    __syncthreads();
    const uint32_t nc = w.closest ? ctr[CTR_CUR] : 0u;
    const uint32_t ns = w.shadow_ctr >= 0 ? min(ctr[w.shadow_ctr], w.cap_shadow) : 0u;
    const uint32_t nall = nc + ns;
    const uint32_t nch = nall / CHUNK + (nall % CHUNK != 0);
    const uint32_t W = gridDim.x * (TRACE_BLOCK / 64);
    uint32_t *lds = s_code + threadIdx.x;
    uint32_t *ovf = ovf_buf + blockIdx.x * TRACE_BLOCK + threadIdx.x;
    const uint32_t os = gridDim.x * TRACE_BLOCK;
    const int32_t DONE = INT32_MIN;
    const float INF = __int_as_float(0x7f800000);
    const uint32_t lane = lane_id();
    // wave-uniform: the chunk being handed out, [pos, end)
    uint32_t chunk = blockIdx.x * (TRACE_BLOCK / 64) + (threadIdx.x >> 6);
    uint32_t set = blockIdx.x % NSET, tries = 0;
    uint32_t pos = 0, end = 0;
    bool more = chunk < nch;
    if (more) {
        pos = chunk * CHUNK;
        end = min(pos + CHUNK, nall);
    }
    // per lane: the item and its ray's traversal state
    uint32_t item = HIT_NONE;
    bool any = false, hit = false;
    V3 o = v3(0.f, 0.f, 0.f), d = o, inv = o;
    float tclose = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t hslot = 0, nodes_c = 0, prims_c = 0, nodes_s = 0, prims_s = 0;
    int sp = 0;
    int32_t code = DONE;
    // the top entry, DONE for an empty stack: an LDS read at a clamped slot
    // (in bounds for every lane, so it needs no branch), replaced by the
    // overflow column's entry when the top lies there
    auto pop = [&]() -> int32_t {
        const int t = sp - 1;
        int32_t v = (int32_t)lds[(uint32_t)min(max(t, 0), SS - 1) * TRACE_BLOCK];
        if (t >= SS)
            v = (int32_t)ovf[(uint32_t)(t - SS) * os];
        sp = max(t, 0);
        return t < 0 ? DONE : v;
    };
    auto push = [&](int32_t cc) {
        if (sp < SS)
            lds[sp * TRACE_BLOCK] = (uint32_t)cc;
        else if (sp < STACK4_MAX)
            ovf[(uint32_t)(sp - SS) * os] = (uint32_t)cc;
        sp = sp < STACK4_MAX ? sp + 1 : sp;
    };
    for (;;) {
        uint64_t idle = __ballot(item == HIT_NONE);
        if (idle && (more || pos < end)
            && (__popcll(idle) >= REFILL_MIN || idle == __ballot(true))) {
            // hand the idle lanes the next items of the wave's chunk,
            // claiming further chunks as it runs out
            while (idle) {
                if (pos >= end) {
                    if (!more || !next_chunk(&ctr[CTR_SETS], nch, W, set, tries, chunk)) {
                        more = false;
                        break;
                    }
                    pos = chunk * CHUNK;
                    end = min(pos + CHUNK, nall);
                }
                // keep the wave's cursor in scalar registers
                pos = __builtin_amdgcn_readfirstlane(pos);
                end = __builtin_amdgcn_readfirstlane(end);
                const uint32_t avail = end - pos;
                const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
                if (((idle >> lane) & 1ull) && rank < avail) {
                    item = pos + rank;
                    float4 o4, d4;
                    uint32_t sid = 0;
                    if (item < nc) {
                        o4 = w.q_o[item];
                        d4 = w.q_d[item];
                        if (BATCH)
                            sid = __float_as_uint(w.q_t[item].w);
                        any = false;
                        tclose = IT_MAX_F;
                    } else {
                        o4 = w.s_o[item - nc];
                        d4 = w.s_d[item - nc];
                        if (BATCH)
                            sid = tag_scene(__float_as_uint(d4.w));
                        any = true;
                        tclose = fminf(o4.w, IT_MAX_F);
                    }
                    o = xyz(o4);
                    d = xyz(d4);
                    inv = v3(inv_dir(d.x), inv_dir(d.y), inv_dir(d.z));
                    hit = false;
                    hslot = HIT_NONE;
                    bu = bv = 0.0f;
                    sp = 0;
                    if (BATCH)
                        code = (int32_t)s_root[sid];
                    else
                        code = S.n_nodes ? (int32_t)S.root : DONE;
                }
                pos += min((uint32_t)__popcll(idle), avail);
                idle = __ballot(item == HIT_NONE);
            }
        }
        if (__ballot(item != HIT_NONE) == 0) {
            if (!more && pos >= end)
                break;
            continue;
        }
        // steps until enough lanes are idle to refill (or all are): the
        // loop holds only the traversal, so the ray state stays in place
        // (1,506 -> 1,547 Msamples/s against one loop with the refill at its
        // top)
        for (;;) {
            bool found = false;
            // one node step per lane at an inner node (a ray of a scene
            // without a BVH starts at DONE: only its planes are tested below)
            if (item != HIT_NONE && code >= 0) {
                float4 lx, hx, ly, hy, lz, hz, cw;
                const uint32_t rel = BATCH ? (uint32_t)code - top_base
                                           : (uint32_t)code;
                if (rel < ntop) {
                    const float4 *n = s_top + (size_t)rel * 7;
                    lx = n[0]; hx = n[1]; ly = n[2]; hy = n[3]; lz = n[4]; hz = n[5];
                    cw = n[6];
                } else {
                    const float4 *n = S.nodes + (size_t)code * 8;
                    lx = n[0]; hx = n[1]; ly = n[2]; hy = n[3]; lz = n[4]; hz = n[5];
                    cw = n[6];
                }
                if (COUNT) {
                    if (any)
                        nodes_s++;
                    else
                        nodes_c++;
                }
                float t0, t1, t2, t3;
                float k0 = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, o, inv,
                                  tclose, t0) ? t0 : INF;
                float k1 = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, o, inv,
                                  tclose, t1) ? t1 : INF;
                float k2 = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, o, inv,
                                  tclose, t2) ? t2 : INF;
                float k3 = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, o, inv,
                                  tclose, t3) ? t3 : INF;
                int32_t c0 = __float_as_int(cw.x), c1 = __float_as_int(cw.y),
                        c2 = __float_as_int(cw.z), c3 = __float_as_int(cw.w);
                // shadow rays take the hit children in slot order (any
                // hit ends them): keyed by slot, the same network puts
                // them first in that order, so a wave that holds both
                // kinds of ray runs one branch-free sort
                if (any) {
                    k0 = k0 != INF ? 0.0f : INF;
                    k1 = k1 != INF ? 1.0f : INF;
                    k2 = k2 != INF ? 2.0f : INF;
                    k3 = k3 != INF ? 3.0f : INF;
                }
                cx(k0, c0, k1, c1);
                cx(k2, c2, k3, c3);
                cx(k0, c0, k2, c2);
                cx(k1, c1, k3, c3);
                cx(k1, c1, k2, c2);
                if (k0 == INF) {
                    code = pop();
                } else {
                    // the hit children after the nearest, farthest
                    // first: hit keys are a prefix of the sorted four, so
                    // with np of them after the nearest, slot sp + j takes
                    // child np - j.  Three slots are written whatever np
                    // is (those at sp + np and above are free), without a
                    // branch per child, while they all lie in LDS
                    const uint32_t np = (uint32_t)(k1 != INF) + (uint32_t)(k2 != INF)
                                      + (uint32_t)(k3 != INF);
                    if (sp + 3 <= SS) {
                        lds[(uint32_t)sp * TRACE_BLOCK] =
                            (uint32_t)(np == 3u ? c3 : (np == 2u ? c2 : c1));
                        lds[(uint32_t)(sp + 1) * TRACE_BLOCK] = (uint32_t)(np == 3u ? c2 : c1);
                        lds[(uint32_t)(sp + 2) * TRACE_BLOCK] = (uint32_t)c1;
                        sp += (int)np;
                    } else {
                        if (k3 != INF)
                            push(c3);
                        if (k2 != INF)
                            push(c2);
                        if (k1 != INF)
                            push(c1);
                    }
                    code = c0;
                }
            }
            // Leaf step, after the node step of the same iteration: each lane
            // at a leaf hands its next one or two records to the wave's
            // lanes, one record per lane, so the triangle tests of the whole
            // wave run once instead of once per record slot of the lanes
            // that hold a leaf.  A test lane reads its owner's ray by a lane
            // shuffle; the owner reads the results back and takes them in
            // record order, as a test of the records one after the other
            // does (the same hit, the same counts, an any-hit query stopping
            // at the first hit).  Owners whose records do not fit in the 64
            // lanes keep their leaf for the next iteration.
            const bool at_leaf = item != HIT_NONE && code < 0 && code != DONE;
            const uint64_t lm = __ballot(at_leaf);
            if (lm) {
                const uint32_t v = ~(uint32_t)code, first = v >> 3, cnt = (v & 7u) + 1u;
                const bool two = at_leaf && cnt > 1u;
                const uint64_t m2 = __ballot(two);
                const uint64_t below = (1ull << lane) - 1ull;
                const uint32_t pre = (uint32_t)(__popcll(lm & below) + __popcll(m2 & below));
                const bool served = at_leaf && pre + 1u + (uint32_t)two <= 64u;
                const uint32_t J = (uint32_t)(__popcll(__ballot(served))
                                              + __popcll(__ballot(served && two)));
                uint32_t *tab = s_job + (threadIdx.x & ~63u);
                if (served) {
                    tab[pre] = lane;
                    if (two)
                        tab[pre + 1u] = lane | 64u;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const uint32_t job = tab[lane];
                const uint32_t ow = job & 63u;
                const V3 jo = v3(__shfl(o.x, ow, 64), __shfl(o.y, ow, 64), __shfl(o.z, ow, 64));
                const V3 jd = v3(__shfl(d.x, ow, 64), __shfl(d.y, ow, 64), __shfl(d.z, ow, 64));
                const uint32_t jrec = __shfl(first, ow, 64) + (job >> 6);
                float tt = INF, uu = 0.0f, vv = 0.0f;
                if (lane < J) {
                    const float4 *pr = S.prims + (size_t)jrec * 3;
                    const float4 A = pr[0], Bq = pr[1], Cq = pr[2];
                    float th;
                    bool h;
                    if (__float_as_uint(Bq.w) == MRT_PRIMKIND_SPHERE)
                        h = isect_sph(jo, jd, A, Bq, th);
                    else
                        h = isect_tri(jo, jd, A, Bq, Cq, S.one_sided != 0, th, uu, vv);
                    if (h)
                        tt = th;
                }
                const float t0 = __shfl(tt, pre, 64), u0 = __shfl(uu, pre, 64),
                            w0 = __shfl(vv, pre, 64);
                const float t1 = __shfl(tt, pre + 1u, 64), u1 = __shfl(uu, pre + 1u, 64),
                            w1 = __shfl(vv, pre + 1u, 64);
                if (served) {
                    if (COUNT) {
                        if (any)
                            prims_s++;
                        else
                            prims_c++;
                    }
                    if (t0 < tclose) {
                        tclose = t0;
                        hit = true;
                        hslot = first;
                        bu = u0;
                        bv = w0;
                        found = any;
                    }
                    if (!found && two) {
                        if (COUNT) {
                            if (any)
                                prims_s++;
                            else
                                prims_c++;
                        }
                        if (t1 < tclose) {
                            tclose = t1;
                            hit = true;
                            hslot = first + 1u;
                            bu = u1;
                            bv = w1;
                            found = any;
                        }
                    }
                    // the records after this pair: the same leaf code two on
                    if (!found)
                        code = cnt > 2 ? (int32_t)~(((first + 2u) << 3) | (cnt - 3u)) : pop();
                }
                // everything this step loaded or shuffled has been used by
                // now; saying so keeps the compiler from taking those
                // registers as pending in the next iteration, where it would
                // drain all memory operations in the middle of the node fetch
                __builtin_amdgcn_s_waitcnt(0);
            }
            if (item != HIT_NONE) {
                if (found || code == DONE) {
                    // the ray's planes: the scene's, or in a batch with
                    // planes those of the ray's scene (S.n_planes is then
                    // the batch's total)
                    const float4 *pl = S.planes;
                    uint32_t npl = S.n_planes;
                    if (BATCH && npl) {
                        const BatchScene &b =
                            bs[any ? tag_scene(__float_as_uint(w.s_d[item - nc].w))
                                   : __float_as_uint(w.q_t[item].w)];
                        pl = S.planes + 2 * (size_t)b.fp.plane0;
                        npl = b.n_planes;
                        // its loads end here: otherwise the compiler takes
                        // their registers as pending on every pass through
                        // the loop and drains all memory operations before
                        // the next node's last load (about 12 % of a batch
                        // launch)
                        __builtin_amdgcn_s_waitcnt(0);
                    }
                    if (!any) {
                        // closest_hit: the planes after the BVH
                        float tb = hit ? tclose : IT_MAX_F;
                        uint32_t slot = hit ? hslot : HIT_NONE;
                        bool hh = hit;
                        for (uint32_t k = 0; k < npl; k++) {
                            float tp;
                            if (isect_pln(o, d, pl[2 * k], pl[2 * k + 1], tp)
                                && (!hh || tp < tb)) {
                                hh = true;
                                tb = tp;
                                slot = HIT_PLANE | k;
                            }
                        }
                        w.hit[item] = make_float4(tb, __uint_as_float(hh ? slot : HIT_NONE),
                                                  bu, bv);
                    } else {
                        bool occ = hit;
                        // the ray's length is read only when planes are to
                        // be tested (a load left unused keeps its register
                        // pending for the compiler, as above)
                        if (!occ && npl) {
                            const float tlim = w.s_o[item - nc].w;
                            for (uint32_t k = 0; k < npl && !occ; k++) {
                                float tp;
                                if (isect_pln(o, d, pl[2 * k], pl[2 * k + 1], tp)
                                    && tp < tlim)
                                    occ = true;
                            }
                        }
                        w.vis[item - nc] = occ ? 0 : 1;
                    }
                    item = HIT_NONE;
                }
            }
            const uint64_t idl = __ballot(item == HIT_NONE);
            if (idl == __ballot(true)
                || ((more || pos < end) && __popcll(idl) >= REFILL_MIN))
                break;
        }
    }
    if (COUNT) {
        wave_add_u64(&stats[STAT_NODES_CLOSEST], nodes_c);
        wave_add_u64(&stats[STAT_PRIMS_CLOSEST], prims_c);
        wave_add_u64(&stats[STAT_NODES_SHADOW], nodes_s);
        wave_add_u64(&stats[STAT_PRIMS_SHADOW], prims_s);
    }
}

// With quadrics or alpha-tested textures: the longer leaf code keeps the
// registers it needs (8 waves would spill it).
template <bool COUNT, bool BATCH>
__global__ void __launch_bounds__(TRACE_BLOCK)
k_trace_ext(SceneDev S, TraceWork w, uint32_t *__restrict__ ctr,
            unsigned long long *__restrict__ stats,
            const BatchScene *__restrict__ bs, const uint2 *__restrict__ slot_info,
            uint32_t *__restrict__ ovf_buf)
{
    __shared__ uint32_t s_code[TRACE_SS * TRACE_BLOCK];
    trace_work<COUNT, BATCH, true>(S, w, ctr, stats, bs, slot_info, ovf_buf, s_code);
}

// ---- shading -----------------------------------------------------------------
struct Mat {
    float4 a, b, c;   // (color, metallic) (rough, ior, trans, ao) (emis, ecolor)
};

__device__ __forceinline__ float pow64(float x)
{
#pragma unroll
    for (int i = 0; i < 6; i++)
        x = x * x;
    return x;
}

// powf(1 - cos, 5) of fresnel_schlick_* (src/math/fresnel_schlick.c:19-38)
__device__ __forceinline__ float schlick_pow5(float c) { return powf(1.0f - c, 5.0f); }

// ggx_sample_hemisphere (src/material/utils.c:39-63)
__device__ __forceinline__ V3 ggx_sample(V3 n, float rough, float u2, float uphi)
{
    float phi = 6.28318530717958648f * uphi;
    float r2 = rough * rough;
    float cth = sqrtf((1.0f - u2) / (1.0f + (r2 * r2 - 1.0f) * u2));
    float sth = sqrtf(1.0f - cth * cth);
    V3 tg = unit(cross(n, v3(1.0f, 0.0f, 0.0f)));
    if (dot(tg, tg) < 0.0001f)
        tg = unit(cross(n, v3(0.0f, 1.0f, 0.0f)));
    V3 bt = unit(cross(n, tg));
    return add(mul(n, cth), add(mul(tg, cosf(phi) * sth), mul(bt, sinf(phi) * sth)));
}

// cos_weighted_sample_hemishphere (src/material/material_default.c:20-39)
__device__ __forceinline__ V3 cos_hemi(V3 n, float r1, float r2)
{
    float phi = 6.28318530717958648f * r1;
    V3 l = v3(cosf(phi) * sqrtf(r2), sqrtf(1.0f - r2), sinf(phi) * sqrtf(r2));
    n = unit(n);
    V3 u = fabsf(n.y) < 0.999f ? unit(cross(v3(0.f, 1.f, 0.f), n))
                               : unit(cross(v3(1.f, 0.f, 0.f), n));
    return unit(add(add(mul(u, l.x), mul(n, l.y)), mul(cross(n, u), l.z)));
}

// random_in_unit_sphere (src/math/random.c:42-50) on four counter draws
__device__ __forceinline__ V3 unit_sphere_pt(U4 u)
{
    V3 p = v3(u.x * 2.0f - 1.0f, u.y * 2.0f - 1.0f, u.z * 2.0f - 1.0f);
    if (dot(p, p) >= 1.0f)
        p = mul(unit(p), cbrtf(u.w));
    return p;
}

// metallic_color (src/material/metallic.c:25-48)
__device__ __forceinline__ void metal(const RngKey &rk, V3 n, const Mat &m,
                                      V3 &d, V3 &thr)
{
    float cth = -dot(unit(d), n);
    V3 mn = n;
    if (m.b.x > 0.0f) {
        U4 g = rk(2u);
        mn = ggx_sample(n, m.b.x, g.x, g.y);
    }
    d = unit(sub(d, mul(mn, 2.0f * dot(d, mn))));
    float p5 = schlick_pow5(cth);
    thr = had(thr, v3(m.a.x + (1.0f - m.a.x) * p5, m.a.y + (1.0f - m.a.y) * p5,
                      m.a.z + (1.0f - m.a.z) * p5));
}

// refracted_ray / refracted_value / refracted_v (src/material/refraction.c:16-76)
__device__ __forceinline__ void refract(const RngKey &rk, float ior_all, V3 n,
                                        bool front, const Mat &m, V3 &d,
                                        V3 &thr)
{
    float ior = m.b.y;
    float eta = front ? ior_all / ior : ior / ior_all;
    d = unit(d);
    float cth = -dot(d, n);
    if (cth > 1.0f)
        cth = 1.0f;
    float sth = sqrtf(1.0f - cth * cth);
    if (eta * sth > 1.0f) {
        metal(rk, n, m, d, thr);
        return;
    }
    float c2 = dot(neg(d), n);
    if (c2 > 1.0f)
        c2 = 1.0f;
    V3 perp = mul(add(d, mul(n, c2)), eta);
    V3 par = mul(n, -sqrtf(fabsf(1.0f - dot(perp, perp))));
    V3 out = unit(add(par, perp));
    if (m.b.x > 0.0f) {
        U4 g = rk(2u);
        out = ggx_sample(out, m.b.x, g.x, g.y);
    }
    d = out;
    thr = had(thr, v3(m.a.x, m.a.y, m.a.z));
}

// The scattering part of material_manager_v3 (src/ray_tracing/
// path_trace.c:15-35) and, for photons, material_manager_v3_phothon
// (src/caustic/trace_photon_path.c:18-35): metal, or dielectric_mat
// (src/material/dielectric.c:40-66).  Returns true when the material takes
// the diffuse branch (default_mat / default_mat_photon); otherwise d and
// thr are updated in place.
__device__ __forceinline__ bool scatter(const RngKey &rk, float ior_all, V3 nrm,
                                        bool front, const Mat &m, V3 &d,
                                        V3 &thr)
{
    bool is_metal;
    if (m.a.w == 1.0f) {
        is_metal = true;
    } else if (m.a.w == 0.0f) {
        is_metal = false;
    } else {
        U4 u = rk(1u);
        is_metal = m.a.w > u.x;
    }
    if (is_metal) {
        metal(rk, nrm, m, d, thr);
        return false;
    }
    if (!(m.b.y > 0.0f))
        return true;
    U4 u = rk(1u);
    float cth = -dot(unit(d), nrm);
    float q = (m.b.y - 1.0f) / (m.b.y + 1.0f);
    float f0 = powf(q, 2.0f);
    float refl = f0 + (1.0f - f0) * schlick_pow5(cth);
    if (refl >= u.y && m.b.y != 1.0f) {
        V3 mn = nrm;
        if (m.b.x > 0.0f) {
            U4 g = rk(2u);
            mn = ggx_sample(nrm, m.b.x, g.x, g.y);
        }
        d = unit(sub(d, mul(mn, 2.0f * dot(d, mn))));
        return false;
    }
    if (m.b.z == 1.0f) {
        refract(rk, ior_all, nrm, front, m, d, thr);
        return false;
    }
    if (m.b.z == 0.0f || m.b.z < u.z)
        return true;
    refract(rk, ior_all, nrm, front, m, d, thr);
    return false;
}

// d/d(roughness) of log p(m) for the GGX micro-normal that ggx_sample draws
// from u2: p(m) = D(m) cos(theta_m), alpha = roughness^2,
// d log D / d alpha = 2/alpha - 4 alpha / (alpha^2 + tan^2 theta).
__device__ __forceinline__ float ggx_score(float rough, float u2)
{
    float a = rough * rough, a2 = a * a;
    float c2 = (1.0f - u2) / (1.0f + (a2 - 1.0f) * u2);
    float t2 = (1.0f - c2) / fmaxf(c2, 1e-30f);
    return (2.0f / a - 4.0f * a / (a2 + t2)) * (2.0f * rough);
}

__device__ __forceinline__ void metal_adj(const RngKey &rk, V3 n, const Mat &m,
                                          V3 &d, V3 &f, V3 &df, float &score)
{
    float cth = -dot(unit(d), n);
    V3 mn = n;
    if (m.b.x > 0.0f) {
        U4 g = rk(2u);
        mn = ggx_sample(n, m.b.x, g.x, g.y);
        score = ggx_score(m.b.x, g.x);
    }
    d = unit(sub(d, mul(mn, 2.0f * dot(d, mn))));
    float p5 = schlick_pow5(cth);
    f = v3(m.a.x + (1.0f - m.a.x) * p5, m.a.y + (1.0f - m.a.y) * p5,
           m.a.z + (1.0f - m.a.z) * p5);
    df = v3(1.0f - p5, 1.0f - p5, 1.0f - p5);
}

__device__ __forceinline__ void refract_adj(const RngKey &rk, float ior_all, V3 n,
                                            bool front, const Mat &m, V3 &d,
                                            V3 &f, V3 &df, float &score)
{
    float ior = m.b.y;
    float eta = front ? ior_all / ior : ior / ior_all;
    d = unit(d);
    float cth = -dot(d, n);
    if (cth > 1.0f)
        cth = 1.0f;
    float sth = sqrtf(1.0f - cth * cth);
    if (eta * sth > 1.0f) {
        metal_adj(rk, n, m, d, f, df, score);
        return;
    }
    float c2 = dot(neg(d), n);
    if (c2 > 1.0f)
        c2 = 1.0f;
    V3 perp = mul(add(d, mul(n, c2)), eta);
    V3 par = mul(n, -sqrtf(fabsf(1.0f - dot(perp, perp))));
    V3 out = unit(add(par, perp));
    if (m.b.x > 0.0f) {
        U4 g = rk(2u);
        out = ggx_sample(out, m.b.x, g.x, g.y);
        score = ggx_score(m.b.x, g.x);
    }
    d = out;
    f = v3(m.a.x, m.a.y, m.a.z);
    df = v3(1.0f, 1.0f, 1.0f);
}

// scatter() with the factor, its albedo derivative and the GGX score.
__device__ __forceinline__ bool scatter_adj(const RngKey &rk, float ior_all, V3 nrm,
                                            bool front, const Mat &m, V3 &d,
                                            V3 &f, V3 &df, float &score)
{
    f = v3(1.0f, 1.0f, 1.0f);
    df = v3(0.0f, 0.0f, 0.0f);
    score = 0.0f;
    bool is_metal;
    if (m.a.w == 1.0f) {
        is_metal = true;
    } else if (m.a.w == 0.0f) {
        is_metal = false;
    } else {
        U4 u = rk(1u);
        is_metal = m.a.w > u.x;
    }
    if (is_metal) {
        metal_adj(rk, nrm, m, d, f, df, score);
        return false;
    }
    if (!(m.b.y > 0.0f))
        return true;
    U4 u = rk(1u);
    float cth = -dot(unit(d), nrm);
    float q = (m.b.y - 1.0f) / (m.b.y + 1.0f);
    float f0 = powf(q, 2.0f);
    float refl = f0 + (1.0f - f0) * schlick_pow5(cth);
    if (refl >= u.y && m.b.y != 1.0f) {
        V3 mn = nrm;
        if (m.b.x > 0.0f) {
            U4 g = rk(2u);
            mn = ggx_sample(nrm, m.b.x, g.x, g.y);
            score = ggx_score(m.b.x, g.x);
        }
        d = unit(sub(d, mul(mn, 2.0f * dot(d, mn))));
        return false;
    }
    if (m.b.z == 1.0f) {
        refract_adj(rk, ior_all, nrm, front, m, d, f, df, score);
        return false;
    }
    if (m.b.z == 0.0f || m.b.z < u.z)
        return true;
    refract_adj(rk, ior_all, nrm, front, m, d, f, df, score);
    return false;
}

// Hit point, normal facing the ray, side and material of a hit record
// h4 = (t, code, u, v): hit_plane (hit_plane.c:55-63), hit_sphere with
// set_normal_face (hit_sphere.c:37-46, normal.c:16-29), hit_triangle's
// interpolated vertex normal (hit_triangle.c:15-27, 81-91) and the
// quadrics' parts (u = part).
__device__ __forceinline__ void surface(const float4 *planes,
                                        const float4 *prims, const float *nrm9,
                                        bool one_sided, V3 o, V3 d, float4 h4,
                                        V3 &p, V3 &nrm, bool &front,
                                        uint32_t &mat)
{
    uint32_t code = __float_as_uint(h4.y);
    p = add(o, mul(d, h4.x));
    if (code & HIT_PLANE) {
        uint32_t k = code & ~HIT_PLANE;
        float4 P = planes[2 * k], N = planes[2 * k + 1];
        mat = __float_as_uint(P.w);
        front = (d.x * N.x + d.y * N.y + d.z * N.z) < 0.0f;
        nrm = front ? xyz(N) : neg(xyz(N));
        return;
    }
    const float4 *rec = prims + (size_t)code * 3;
    float4 A = rec[0], Bv = rec[1], C = rec[2];
    if (__float_as_uint(Bv.w) >= MRT_PRIMKIND_QUADRIC) {
        quadric_surface(A, Bv, C, nrm9 + (size_t)code * 9, d, p,
                        (uint32_t)h4.z, nrm, front, mat);
        return;
    }
    mat = __float_as_uint(C.w);
    if (__float_as_uint(Bv.w) == MRT_PRIMKIND_SPHERE) {
        float rad = Bv.x;
        V3 ow = sub(p, xyz(A));
        ow = v3(ow.x / rad, ow.y / rad, ow.z / rad);
        front = dot(d, ow) < 0.0f;
        nrm = front ? ow : neg(ow);
        return;
    }
    const float *N = nrm9 + (size_t)code * 9;
    float bu = h4.z, bv = h4.w, w = 1.0f - bu - bv;
    V3 nn = v3(w * N[0] + bu * N[3] + bv * N[6], w * N[1] + bu * N[4] + bv * N[7],
               w * N[2] + bu * N[5] + bv * N[8]);
    nn = unit(nn);
    if (!one_sided && dot(d, nn) > 0.0f) {
        front = false;
        nrm = neg(nn);
    } else {
        front = true;
        nrm = nn;
    }
}

// apply_all_map (apply_map.c:60-86) and get_hit_register_color
// (get_hit_color.c:17-36) at a hit with record h4 and point p: the normal
// map (apply_normal_map, :37-58; only image textures act as one), then the
// red channel of the transmission, roughness, metallic, ao and emission
// strength textures, the emission colour and the colour texture.
__device__ __forceinline__ void apply_textures(const TexDev &T,
                                               const float4 *planes,
                                               const float4 *prims, float4 h4,
                                               V3 p, uint32_t mat, V3 &nrm,
                                               Mat &m)
{
    uint4 M0 = T.mtex[3 * mat], M1 = T.mtex[3 * mat + 1], M2 = T.mtex[3 * mat + 2];
    if ((M0.x & M0.y & M0.z & M0.w & M1.x & M1.y & M1.z & M1.w) == MRT_NO_TEX)
        return;
    float sc = __uint_as_float(M2.y), u, v;
    uint32_t code = __float_as_uint(h4.y);
    if (code & HIT_PLANE) {
        uint32_t k = code & ~HIT_PLANE;
        plane_uv(planes[2 * k], planes[2 * k + 1], p, sc, u, v);
    } else {
        const float4 *rec = prims + (size_t)code * 3;
        prim_uv(T, rec[0], rec[1], rec[2], p, h4.z, h4.w, sc, u, v);
    }
    if (M1.w != MRT_NO_TEX) {
        float4 T0 = T.tex[3 * M1.w];
        if (__float_as_uint(T0.x) == MRT_TEX_IMAGE) {
            uint32_t c = tex_texel(T, T0, u, v);
            V3 nm = v3((float)(c & 255u) / 127.5f - 1.0f,
                       (float)((c >> 8) & 255u) / 127.5f - 1.0f,
                       (float)((c >> 16) & 255u) / 127.5f - 1.0f);
            V3 tg = cross(nrm, v3(0.f, -1.f, 0.f));
            if (dot(tg, tg) == 0.0f)
                tg = cross(nrm, v3(0.f, 0.f, 1.f));
            tg = unit(tg);
            V3 bt = unit(cross(nrm, tg));
            V3 mapped = unit(add(add(mul(tg, nm.x), mul(bt, nm.y)), mul(nrm, nm.z)));
            float k = __uint_as_float(M2.x);
            nrm = k == 1.0f ? mapped
                            : unit(add(mul(nrm, 1.0f - k), mul(mapped, k)));
        }
    }
    if (M0.w != MRT_NO_TEX)
        m.b.z = tex_rgb(T, M0.w, u, v, p).x;
    if (M0.z != MRT_NO_TEX)
        m.b.x = tex_rgb(T, M0.z, u, v, p).x;
    if (M0.y != MRT_NO_TEX)
        m.a.w = tex_rgb(T, M0.y, u, v, p).x;
    if (M1.x != MRT_NO_TEX)
        m.b.w = tex_rgb(T, M1.x, u, v, p).x;
    if (M1.y != MRT_NO_TEX)
        m.c.x = tex_rgb(T, M1.y, u, v, p).x;
    if (M1.z != MRT_NO_TEX) {
        V3 e = tex_rgb(T, M1.z, u, v, p);
        m.c.y = e.x;
        m.c.z = e.y;
        m.c.w = e.z;
    }
    if (M0.x != MRT_NO_TEX) {
        V3 c = tex_rgb(T, M0.x, u, v, p);
        m.a.x = c.x;
        m.a.y = c.y;
        m.a.z = c.z;
    }
}

enum : uint32_t { SC_NONE = 0, SC_DIFFUSE = 1 };

// ---- environment -------------------------------------------------------------
constexpr double PI_D = 3.14159265358979323846;

// Texel of get_background_color_value / get_hdr_color (background.c:15-57):
// v = 1/2 + asin(d.y)/pi, u = 1/2 - atan2(d.z, d.x)/(2 pi), texel
// ((int)(u (w-1)), (int)(v (h-1))), scaled by 2^(exposure + ratio - 1)
// (get_hdr_pixel_skybox, hit_register/get_hdr_value.c:43-70).
__device__ __forceinline__ V3 env_texel(const FrameParams &fp, V3 dir)
{
    V3 d = unit(dir);
    double v = 0.5 + asin((double)d.y) / PI_D;
    double u = 0.5 - atan2((double)d.z, (double)d.x) / (2.0 * PI_D);
    int x = (int)(u * (fp.env_w - 1)), y = (int)(v * (fp.env_h - 1));
    float4 t = fp.env_tex[(size_t)y * fp.env_w + x];
    return v3(t.x * fp.env_scale, t.y * fp.env_scale, t.z * fp.env_scale);
}

// Texel index of env_texel.
__device__ __forceinline__ uint32_t env_index(const FrameParams &fp, V3 dir)
{
    V3 d = unit(dir);
    double v = 0.5 + asin((double)d.y) / PI_D;
    double u = 0.5 - atan2((double)d.z, (double)d.x) / (2.0 * PI_D);
    int x = (int)(u * (fp.env_w - 1)), y = (int)(v * (fp.env_h - 1));
    return (uint32_t)y * fp.env_w + (uint32_t)x;
}

// Largest index i in [0, n-1] with cdf[i] <= u (cdf[0] = 0).
__device__ __forceinline__ uint32_t cdf_search(const float *cdf, uint32_t n, float u)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (cdf[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// cdf_search through a guide table g of G + 1 entries (G a power of two):
// g[k] = cdf_search(cdf, n, k / G) for k < G and g[G] = n - 1.  The answer
// is monotone in u, so for u in [k / G, (k + 1) / G) it lies in
// [g[k], g[k + 1]]; k = floor(u G) and k / G are exact in fp32.  The search
// over that range makes cdf_search's comparisons and returns its index,
// after one guide load and a few entries instead of log2(n) dependent loads.
__device__ __forceinline__ uint32_t cdf_search_guided(const float *cdf,
                                                      const uint32_t *g,
                                                      uint32_t G, float u)
{
    const uint32_t k = min((uint32_t)(u * (float)G), G - 1u);
    uint32_t lo = g[k], hi = g[k + 1];
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (cdf[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Solid-angle density of the MIS environment sampler in direction d.
__device__ __forceinline__ float env_pdf(const FrameParams &fp, V3 dir)
{
    V3 d = unit(dir);
    const uint32_t nw = fp.env_w - 1, nh = fp.env_h - 1;
    double v = 0.5 + asin((double)d.y) / PI_D;
    double u = 0.5 - atan2((double)d.z, (double)d.x) / (2.0 * PI_D);
    uint32_t y = (uint32_t)(v * nh), x = (uint32_t)(u * nw);
    y = y < nh ? y : nh - 1;
    x = x < nw ? x : nw - 1;
    float cose = sqrtf(fmaxf(0.0f, 1.0f - d.y * d.y));
    float pc = fp.mis_pcell[(size_t)y * nw + x];
    if (!(cose > 0.0f) || !(pc > 0.0f))
        return 0.0f;
    return pc * (float)nw * (float)nh / (19.7392088021787172f * cose);
}

// ENV: 0 compiles the HDRI code out, 1 keeps its MIS estimator, 2 its
// reference estimator (the launcher picks by the scene's); TEX = false
// compiles the texture code out.  Scenes without them, and every batch, get
// fewer registers and more resident waves.
// At least 4 waves per SIMD: the HDRI variants are held to 128 VGPRs
// (spilling a little) instead of 160-186 and 2-3 waves, 2 % more
// Msamples/s on the caustics workload; the others need fewer already.
// REC: the adjoint's recording replay (single scene, no textures, the MIS
// or no environment, exact light sums): the same paths and image, and per
// vertex a tape record for the backward pass (see k_adjoint_back).
template <bool BATCH, int ENV, bool TEX, bool REC = false>
__global__ void __launch_bounds__(SHADE_BLOCK)
__attribute__((amdgpu_waves_per_eu(4)))
k_shade(FrameParams fp, ShadeBufs B, uint32_t *__restrict__ ctr)
{
    const uint32_t n = ctr[CTR_CUR];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        uint32_t want_next = 0, n_rays = 0, want_pend = 0;
        uint32_t kind = SC_NONE, slot = 0, pixel = 0, sample = 0;
        V3 p = v3(0.f, 0.f, 0.f), nrm = v3(0.f, 0.f, 0.f), d = v3(0.f, 0.f, 0.f);
        V3 thr = v3(0.f, 0.f, 0.f), cp = v3(0.f, 0.f, 0.f);
        float ao = 1.0f, pb_next = 0.0f;
        bool env_ray = false;
        V3 env_d = v3(0.f, 0.f, 0.f), env_rad = v3(0.f, 0.f, 0.f);
        float env_omc = 1.0f;
        uint32_t sid = 0;
        const FrameParams *Fp = &fp;
        if (active) {
            float4 o4 = B.q_o[i], d4 = B.q_d[i], t4 = B.q_t[i], h4 = B.hit[i];
            slot = __float_as_uint(d4.w);
            uint32_t local = slot;
            if (BATCH) {
                // the scene comes with the path (q_t.w), so its frame loads
                // beside the slot's local index instead of after it
                sid = __float_as_uint(t4.w);
                local = B.slot_info[slot].y;
                Fp = &B.bs[sid].fp;
            }
            const FrameParams &F = *Fp;
            pixel = (local + F.slot0) % F.WH;
            sample = F.sample0 + (local + F.slot0) / F.WH;
            d = xyz(d4);
            thr = xyz(t4);
            uint32_t code = __float_as_uint(h4.y);
            if (code == HIT_NONE) {
                // add_skybox (src/ray_tracing/path_trace.c:37-48)
                float4 a = B.acc[slot];
                float ao_last = o4.w;
                V3 bg = v3(F.bg_clamped[0], F.bg_clamped[1], F.bg_clamped[2]);
                float wmis = 1.0f;
                if (ENV && F.has_env) {
                    bg = env_texel(fp, d);
                    if (ENV == 2) {
                        bg = v3(fminf(fmaxf(bg.x, 0.0f), 16.0f),
                                fminf(fmaxf(bg.y, 0.0f), 16.0f),
                                fminf(fmaxf(bg.z, 0.0f), 16.0f));
                    } else if (t4.w > 0.0f) {
                        // escape after a diffuse bounce: balance heuristic
                        float pl = env_pdf(fp, d);
                        wmis = t4.w / (t4.w + pl);
                    }
                }
                bg = mul(bg, ao_last);
                a.x = a.x + bg.x * thr.x * wmis;
                a.y = a.y + bg.y * thr.y * wmis;
                a.z = a.z + bg.z * thr.z * wmis;
                B.acc[slot] = a;
                if (REC) {
                    // the radiance the path escapes with (before thr), its
                    // texel and weight, and the number of vertices
                    uint32_t et = HIT_NONE;
                    float eg = 0.0f;
                    if (ENV && F.has_env) {
                        et = env_index(fp, d);
                        eg = fp.env_scale * ao_last * wmis;
                    }
                    float4 *te = B.tape;
                    te[slot] = make_float4(bg.x * wmis, bg.y * wmis, bg.z * wmis, eg);
                    te[B.tape_P + slot] = make_float4(__uint_as_float(et),
                                                      __uint_as_float(fp.bounce),
                                                      0.0f, 0.0f);
                }
            } else {
                uint32_t mat;
                bool front;
                surface(B.planes + 2 * (size_t)F.plane0, B.prims, B.nrm, F.one_sided != 0, xyz(o4), d,
                        h4, p, nrm, front, mat);
                Mat m;
                m.a = B.mats[3 * mat];
                m.b = B.mats[3 * mat + 1];
                m.c = B.mats[3 * mat + 2];
                if (TEX && F.textured)
                    apply_textures(B.tx, B.planes + 2 * (size_t)F.plane0, B.prims,
                                   h4, p, mat, nrm, m);
                ao = m.b.w;
                // material_manager_v3 (src/ray_tracing/path_trace.c:15-35)
                if (m.c.x > 0.0f) {
                    float4 a = B.acc[slot];
                    a.x = a.x + (m.c.y * m.c.x) * thr.x;
                    a.y = a.y + (m.c.z * m.c.x) * thr.y;
                    a.z = a.z + (m.c.w * m.c.x) * thr.z;
                    B.acc[slot] = a;
                }
                const RngKey rk{F.seed, pixel, sample, fp.bounce, 0u};
                const V3 thr_in = thr;
                V3 f_adj = v3(1.f, 1.f, 1.f), df_adj = v3(0.f, 0.f, 0.f);
                float sc_adj = 0.0f, ng_adj = 0.0f;
                uint32_t nt_adj = HIT_NONE;
                bool diffuse;
                if (REC) {
                    // scatter() with the factor it multiplies into thr, its
                    // albedo derivative and the GGX score
                    diffuse = scatter_adj(rk, F.ior_all, nrm, front, m, d, f_adj,
                                          df_adj, sc_adj);
                    if (diffuse) {
                        f_adj = v3(m.a.x, m.a.y, m.a.z);
                        df_adj = v3(1.f, 1.f, 1.f);
                    } else {
                        thr = had(thr, f_adj);
                    }
                } else {
                    diffuse = scatter(rk, F.ior_all, nrm, front, m, d, thr);
                }
                if (diffuse) {
                    // default_mat (src/material/material_default.c:71-92)
                    kind = SC_DIFFUSE;
                    V3 col = v3(m.a.x, m.a.y, m.a.z);
                    cp = had(col, thr);
                    thr = had(thr, col);
                    const float4 *L = B.lights + 3 * (size_t)F.light0;
                    for (uint32_t k = 0; k < F.n_lights; k++) {
                        float4 l0 = L[3 * k], l1 = L[3 * k + 1], l2 = L[3 * k + 2];
                        uint32_t ltype = __float_as_uint(l2.x);
                        int32_t lss = __float_as_int(l2.y);
                        V3 ld = ltype == MRT_LIGHT_POINT ? unit(sub(xyz(l0), p))
                                                         : neg(xyz(l0));
                        if (dot(nrm, ld) <= 0.0f)
                            continue;
                        bool soft = ltype == MRT_LIGHT_POINT && l1.w != 0.0f && lss > 1;
                        n_rays += soft ? (uint32_t)lss : 1u;
                    }
                    if (ENV && F.has_env) {
                        U4 e = rng4(F.seed, pixel, sample, fp.bounce, 3u);
                        if (ENV == 2) {
                            // importance_sampling (material_default.c:41-69)
                            const uint32_t ew = F.env_w, eh = F.env_h;
                            double vy = F.env_marg_inv[(int)((double)e.x * (eh - 1))];
                            double vx = F.env_cond_inv[(int)(vy * (eh - 1) * ew
                                                              + (double)e.y * (ew - 1))];
                            float pdf = F.env_pdf[(int)(vy * (eh - 1) * ew
                                                         + vx * (ew - 1))];
                            if (pdf > 0.0f) {
                                float th = (float)((vy - 0.5) * PI_D);
                                float ph = (float)((0.5 - vx) * 2.0 * PI_D);
                                env_d = v3((float)(cos((double)th) * cos((double)ph)),
                                           (float)sin((double)th),
                                           (float)(cos((double)th) * sin((double)ph)));
                                float cth = dot(nrm, env_d);
                                env_rad = had(mul(env_texel(fp, env_d), ao), thr);
                                env_rad = mul(env_rad, cth / pdf);
                                env_omc = 1.0f - cth;
                                env_ray = true;
                            }
                        } else {
                            // light-sample half of the MIS pair
                            const uint32_t nw = F.env_w - 1, nh = F.env_h - 1;
                            uint32_t y = cdf_search_guided(F.mis_marg,
                                                           F.mis_marg_guide,
                                                           F.guide_m, e.x);
                            uint32_t x = cdf_search_guided(
                                F.mis_cond + (size_t)y * F.env_w,
                                F.mis_cond_guide + (size_t)y * (F.guide_c + 1),
                                F.guide_c, e.y);
                            float pc = F.mis_pcell[(size_t)y * nw + x];
                            float vv = ((float)y + e.z) / (float)nh;
                            float uu = ((float)x + e.w) / (float)nw;
                            float el = (vv - 0.5f) * 3.14159265358979324f;
                            float ph = (0.5f - uu) * 6.28318530717958648f;
                            float ce = cosf(el);
                            env_d = v3(ce * cosf(ph), sinf(el), ce * sinf(ph));
                            float cn = dot(nrm, env_d);
                            if (pc > 0.0f && ce > 0.0f && cn > 0.0f) {
                                float pl = pc * (float)nw * (float)nh
                                           / (19.7392088021787172f * ce);
                                float pbs = cn / 3.14159265358979324f;
                                float wl = pl / (pl + pbs);
                                env_rad = had(mul(env_texel(fp, env_d), ao), thr);
                                env_rad = mul(env_rad, (cn / (3.14159265358979324f * pl)) * wl);
                                env_ray = true;
                                if (REC) {
                                    nt_adj = env_index(fp, env_d);
                                    ng_adj = fp.env_scale * ao
                                             * ((cn / (3.14159265358979324f * pl)) * wl);
                                }
                            }
                        }
                        n_rays += env_ray ? 1u : 0u;
                    }
                    U4 g = rng4(F.seed, pixel, sample, fp.bounce, 2u);
                    d = cos_hemi(nrm, g.z, g.w);
                    pb_next = dot(nrm, d) / 3.14159265358979324f;
                    want_pend = (n_rays || F.has_photons) ? 1u : 0u;
                }
                want_next = fp.last_bounce ? 0u : 1u;
                if (REC) {
                    // this vertex: throughput in, factor and score, albedo
                    // derivative and the environment light sample's weight;
                    // the next-event sum comes from k_resolve when the
                    // vertex has a pending record (else it is zero)
                    float4 *tv = B.tape + tape_row(fp.bounce, 0) * B.tape_P + slot;
                    tv[0] = make_float4(thr_in.x, thr_in.y, thr_in.z,
                                        __uint_as_float(mat));
                    tv[B.tape_P] = make_float4(f_adj.x, f_adj.y, f_adj.z, sc_adj);
                    tv[2 * (size_t)B.tape_P] = make_float4(df_adj.x, df_adj.y,
                                                           df_adj.z, ng_adj);
                    tv[3 * (size_t)B.tape_P] = make_float4(0.f, 0.f, 0.f,
                                                           __uint_as_float(HIT_NONE));
                    env_omc = __uint_as_float(nt_adj);   // to k_resolve
                }
            }
        }
        // allocations: every lane of the wave takes part
        uint32_t qn, j0, pr;
        block_alloc3(&ctr[CTR_NEXT], &ctr[ctr_shadow(fp.bounce)],
                     &ctr[ctr_pend(fp.bounce)],
                     want_next, n_rays, want_pend, qn, j0, pr);
        if (!active)
            continue;
        const FrameParams &F = *Fp;
        // the host sizes every queue for the worst case; never write past it
        if (qn >= B.cap_paths)
            want_next = 0;
        if (j0 + n_rays > B.cap_shadow || pr >= B.cap_paths)
            n_rays = want_pend = 0;
        if (want_next) {
            B.qn_o[qn] = make_float4(p.x, p.y, p.z, ao);
            B.qn_d[qn] = make_float4(d.x, d.y, d.z, __uint_as_float(slot));
            // a batch has no HDRI: .w carries the path's scene instead of
            // the BSDF pdf, for the batched refill trace kernel
            B.qn_t[qn] = make_float4(thr.x, thr.y, thr.z,
                                     BATCH ? __uint_as_float(sid) : pb_next);
        }
        if (kind == SC_DIFFUSE && n_rays) {
            V3 view = unit(sub(v3(F.cam_pos[0], F.cam_pos[1], F.cam_pos[2]), p));
            uint32_t j = j0;
            const float4 *L = B.lights + 3 * (size_t)F.light0;
            for (uint32_t k = 0; k < F.n_lights; k++) {
                float4 l0 = L[3 * k], l1 = L[3 * k + 1], l2 = L[3 * k + 2];
                uint32_t ltype = __float_as_uint(l2.x);
                int32_t lss = __float_as_int(l2.y);
                bool point = ltype == MRT_LIGHT_POINT;
                V3 lpos = xyz(l0);
                V3 ld = point ? unit(sub(lpos, p)) : neg(lpos);
                float ndl = dot(nrm, ld);
                if (ndl <= 0.0f)
                    continue;
                V3 nl = neg(ld);
                V3 refl = sub(nl, mul(nrm, 2.0f * dot(nl, nrm)));
                float rdv = dot(refl, view);
                if (rdv < 0.0f)
                    rdv = 0.0f;
                float sp64 = pow64(rdv);
                bool soft = point && l1.w != 0.0f && lss > 1;
                uint32_t cnt = soft ? (uint32_t)lss : 1u;
                for (uint32_t s = 0; s < cnt; s++, j++) {
                    V3 sd;
                    float tmax;
                    if (!point) {
                        sd = unit(ld);
                        tmax = 2.0f * IT_MAX_F;
                    } else {
                        V3 tgt = lpos;
                        if (soft) {
                            U4 u = rng4(F.seed, pixel, sample, fp.bounce,
                                        16u + k * 4096u + s);
                            tgt = add(lpos, mul(unit_sphere_pt(u), l1.w));
                        }
                        V3 dl = sub(tgt, p);
                        sd = unit(dl);
                        tmax = sqrtf(dot(dl, dl));
                    }
                    B.s_o[j] = make_float4(p.x, p.y, p.z, tmax);
                    B.s_d[j] = make_float4(sd.x, sd.y, sd.z,
                                           __uint_as_float((sid << 18) | (k << 12) | s));
                    B.s_c[j] = make_float2(ndl, sp64);
                }
            }
            if (env_ray) {
                B.s_o[j] = make_float4(p.x, p.y, p.z, 2.0f * IT_MAX_F);
                B.s_d[j] = make_float4(env_d.x, env_d.y, env_d.z,
                                       __uint_as_float(TAG_ENV));
                B.s_c[j] = make_float2(0.0f, 0.0f);
            }
        }
        if (want_pend) {
            B.p_a[pr] = make_float4(cp.x, cp.y, cp.z, __uint_as_float(slot));
            // .w: bit 0 the environment ray, bits 1.. a batch path's scene
            B.p_b[pr] = make_uint4(want_next ? qn : HIT_NONE, j0, n_rays,
                                   (env_ray ? 1u : 0u) | (sid << 1));
            B.p_c[pr] = make_float4(env_rad.x, env_rad.y, env_rad.z, env_omc);
            if (F.has_photons)
                B.p_x[pr] = make_float4(p.x, p.y, p.z, 0.0f);
        }
    }
}

// ---- direct-light resolve (light_manager.c:20-70, p_light.c:67-91,
//      d_light.c:35-57): the quantised sums of the visible lights ----------
__device__ __forceinline__ float lq(bool exact, float acc, float add)
{
    float s = acc + add;
    return exact ? s : floorf(s);
}

template <bool BATCH, bool REC = false>
__global__ void __launch_bounds__(256)
k_resolve(FrameParams fp, ShadeBufs B, uint32_t *__restrict__ ctr)
{
    const uint32_t n = min(ctr[ctr_pend(fp.bounce)], B.cap_paths);
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += gridDim.x * blockDim.x) {
        float4 a = B.p_a[r];
        uint4 b = B.p_b[r];
        uint32_t slot = __float_as_uint(a.w);
        // a batch path's scene is in p_b.w (bits 1..); its light range and
        // flags sit in the first line of its record.  Batches have no HDRI
        // and no photon map.
        uint32_t flags = fp.flags, light0 = fp.light0;
        if (BATCH) {
            const BatchScene &bsc = B.bs[b.w >> 1];
            flags = bsc.flags;
            light0 = bsc.light0;
        }
        const bool env = !BATCH && (b.w & 1u);
        const bool exact = (flags & MRT_SCENE_LIGHT_EXACT) != 0;
        const float4 *L = B.lights + 3 * (size_t)light0;
        uint32_t j = b.y, end = b.y + b.z;
        float lr = 0.0f, lg = 0.0f, lb = 0.0f;
        if (env)
            end--;                      // the environment ray comes last
        while (j < end) {
            uint32_t tag = __float_as_uint(B.s_d[j].w);
            uint32_t k = tag_light(tag);
            float4 l0 = L[3 * k], l1 = L[3 * k + 1], l2 = L[3 * k + 2];
            uint32_t ltype = __float_as_uint(l2.x);
            int32_t lss = __float_as_int(l2.y);
            bool point = ltype == MRT_LIGHT_POINT;
            bool soft = point && l1.w != 0.0f && lss > 1;
            uint32_t cnt = soft ? (uint32_t)lss : 1u;
            uint32_t nvis = 0;
            for (uint32_t s = 0; s < cnt; s++)
                nvis += B.vis[j + s];
            float2 c = B.s_c[j];
            j += cnt;
            float br = l0.w;
            if (point) {
                float factor = soft ? (float)nvis / (float)cnt : 1.0f;
                if (!(nvis > 0))
                    continue;
                float diff = br * c.x * factor;
                lr = lq(exact, lr, l1.x * diff);
                lg = lq(exact, lg, l1.y * diff);
                lb = lq(exact, lb, l1.z * diff);
                float spec = c.y * 1.0f * factor;
                lr = lq(exact, lr, l1.x * br * spec);
                lg = lq(exact, lg, l1.y * br * spec);
                lb = lq(exact, lb, l1.z * br * spec);
            } else {
                if (!nvis)
                    continue;
                lr = lq(exact, lr, l1.x * br * c.x);
                lg = lq(exact, lg, l1.y * br * c.x);
                lb = lq(exact, lb, l1.z * br * c.x);
                lr = lq(exact, lr, l1.x * br * 1.0f * c.y);
                lg = lq(exact, lg, l1.y * br * 1.0f * c.y);
                lb = lq(exact, lb, l1.z * br * 1.0f * c.y);
            }
        }
        float4 s = B.acc[slot];
        // default_mat: total lighting = direct + caustics
        V3 dl = v3(lr / 255.0f, lg / 255.0f, lb / 255.0f);
        if (!BATCH && fp.has_photons)
            dl = add(dl, xyz(B.p_k[r]));
        s.x = s.x + a.x * dl.x;
        s.y = s.y + a.y * dl.y;
        s.z = s.z + a.z * dl.z;
        const bool env_vis = env && B.vis[end];
        if (env_vis) {
            float4 c = B.p_c[r];
            s.x = s.x + c.x;
            s.y = s.y + c.y;
            s.z = s.z + c.z;
            if (!REC && fp.estimator == MRT_EST_REFERENCE && b.x != HIT_NONE) {
                float4 t = B.qn_t[b.x];
                t.x = t.x * c.w;
                t.y = t.y * c.w;
                t.z = t.z * c.w;
                B.qn_t[b.x] = t;
            }
        }
        B.acc[slot] = s;
        if (REC) {
            // the vertex's next-event radiance (before thr and albedo): the
            // exact light sum and the visible environment sample
            float4 *tv = B.tape + tape_row(fp.bounce, 2) * B.tape_P + slot;
            uint32_t nt = HIT_NONE;
            V3 N = dl;
            if (env_vis) {
                nt = __float_as_uint(B.p_c[r].w);
                if (nt != HIT_NONE) {
                    const float ng = tv[0].w;
                    const float4 tx = fp.env_tex[nt];
                    N = add(N, v3(tx.x * ng, tx.y * ng, tx.z * ng));
                }
            }
            tv[B.tape_P] = make_float4(N.x, N.y, N.z, __uint_as_float(nt));
        }
    }
}

__global__ void k_advance(uint32_t *__restrict__ ctr,
                          unsigned long long *__restrict__ stats, uint32_t cap,
                          uint32_t bounce)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t prev = bounce + 1;      // parity of bounce - 1
        // atomics: the advance kernels of two render lanes may run together
        atomicAdd(&stats[STAT_RAYS_CLOSEST], (unsigned long long)ctr[CTR_CUR]);
        atomicAdd(&stats[STAT_RAYS_SHADOW],
                  (unsigned long long)ctr[ctr_shadow(prev)]);
        ctr[CTR_CUR] = min(ctr[CTR_NEXT], cap);
        ctr[CTR_NEXT] = 0;
        ctr[ctr_shadow(prev)] = 0;
        ctr[ctr_pend(prev)] = 0;
        reset_sets(ctr);
    }
}

// Per pixel, in sample order: accum += sum of the batch's finite samples.
__global__ void __launch_bounds__(256)
k_accumulate(const float4 *__restrict__ acc, float *__restrict__ accum,
             uint32_t WH, uint32_t nb, unsigned long long *__restrict__ stats)
{
    uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bad = 0;
    if (px < WH) {
        float r = 0.0f, g = 0.0f, b = 0.0f;
        for (uint32_t k = 0; k < nb; k++) {
            float4 v = acc[(size_t)k * WH + px];
            if (isfinite(v.x) && isfinite(v.y) && isfinite(v.z)) {
                r += v.x;
                g += v.y;
                b += v.z;
            } else {
                bad++;
            }
        }
        accum[(size_t)px * 3] += r;
        accum[(size_t)px * 3 + 1] += g;
        accum[(size_t)px * 3 + 2] += b;
    }
    wave_add_u64(&stats[STAT_NONFINITE], bad);
}

// ---- path-replay adjoint ------------------------------------------------------
// The forward estimator of a path, with E_v the emission at vertex v, f_v its
// scatter factor (albedo, Schlick colour of a metal, 1 for a dielectric
// reflection), N_v its next-event terms in local units (lights, HDRI light
// sample) and X_v the radiance arriving along the sampled direction, is
//   R_v = E_v + f_v (N_v + X_v),  X_{v-1} = R_v,  C = R_0,
// with throughput thr_v = f_0 ... f_{v-1}.  Replaying the path with the same
// counter-RNG keys and walking it backward gives, without division,
//   dC/dE_v = thr_v,  dC/df_v = thr_v (N_v + X_v),
//   dC/dtexel = thr_v f_v g (for the texel a light sample or an escape saw),
// and for the GGX roughness of a specular vertex the score-function term
//   dC/dr = d log p(m)/dr * thr_v f_v X_v
// (the sampled micro-normal m moves with r; everything else is fixed).
// The HDRI sampling tables stay those of the uploaded map.

__device__ __forceinline__ void lds_add3(float *g, V3 v)
{
    if (v.x != 0.0f)
        atomicAdd(g, v.x);
    if (v.y != 0.0f)
        atomicAdd(g + 1, v.y);
    if (v.z != 0.0f)
        atomicAdd(g + 2, v.z);
}

__global__ void __launch_bounds__(256)
k_tape_init(float4 *__restrict__ tape, uint32_t P, uint32_t max_depth)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P)
        return;
    tape[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    tape[(size_t)P + i] = make_float4(__uint_as_float(HIT_NONE),
                                      __uint_as_float(max_depth), 0.0f, 0.0f);
}

// The backward pass of the path-replay adjoint, one path per thread over
// the tape its recorded replay (k_shade / k_resolve with REC) wrote: from
// the escape radiance X after the last vertex, vertex v = n-1 .. 0 gets
//   dC/dalbedo_v = dl thr_v df_v (N_v + X),   dC/dE_v = dl thr_v,
//   dC/drough_v  = score_v dl . (thr_v f_v X)  (score-function term),
//   dC/dtexel    = dl thr_v f_v w_v           (environment sample),
// then X = E_v + f_v (N_v + X).  Material sums go through LDS, texel
// gradients to HBM with float atomics.
// Texel gradients of a block are summed in an LDS hash table first (open
// addressing on the texel index, a few probes, then the HBM atomic as the
// fallback) and added to HBM once per entry at the block's end: the
// environment samples concentrate on the few texels of the sun, whose
// HBM atomics from every path would otherwise queue on the same addresses.
constexpr uint32_t TEX_HASH = 2048;       // entries per block (24 KB of sums)
constexpr uint32_t TEX_PROBES = 8;

__device__ __forceinline__ void texel_add(float *__restrict__ g_tex,
                                          uint32_t *s_key, float *s_val,
                                          uint32_t t, V3 v)
{
    uint32_t h = (t * 2654435761u) >> (32 - 11);   // log2(TEX_HASH) = 11
    for (uint32_t p = 0; p < TEX_PROBES; p++, h = (h + 1) & (TEX_HASH - 1)) {
        uint32_t k = s_key[h];
        if (k == HIT_NONE)
            k = atomicCAS(&s_key[h], HIT_NONE, t);
        if (k == HIT_NONE || k == t) {
            atomicAdd(&s_val[3 * h], v.x);
            atomicAdd(&s_val[3 * h + 1], v.y);
            atomicAdd(&s_val[3 * h + 2], v.z);
            return;
        }
    }
    atomicAdd(&g_tex[3 * (size_t)t], v.x);
    atomicAdd(&g_tex[3 * (size_t)t + 1], v.y);
    atomicAdd(&g_tex[3 * (size_t)t + 2], v.z);
}

__global__ void __launch_bounds__(256)
k_adjoint_back(FrameParams fp, ShadeBufs B, AdjBufs A, uint32_t P)
{
    __shared__ float s_g[ADJ_MAXMAT * ADJ_NG];
    __shared__ uint32_t s_key[TEX_HASH];
    __shared__ float s_val[3 * TEX_HASH];
    for (uint32_t k = threadIdx.x; k < ADJ_MAXMAT * ADJ_NG; k += blockDim.x)
        s_g[k] = 0.0f;
    for (uint32_t k = threadIdx.x; k < TEX_HASH; k += blockDim.x) {
        s_key[k] = HIT_NONE;
        s_val[3 * k] = s_val[3 * k + 1] = s_val[3 * k + 2] = 0.0f;
    }
    __syncthreads();
    const size_t TP = B.tape_P;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P;
         i += gridDim.x * blockDim.x) {
        const uint32_t pixel = i % fp.WH;
        const V3 dl = v3(A.dimg[3 * (size_t)pixel] * A.inv_spp,
                         A.dimg[3 * (size_t)pixel + 1] * A.inv_spp,
                         A.dimg[3 * (size_t)pixel + 2] * A.inv_spp);
        if (dl.x == 0.0f && dl.y == 0.0f && dl.z == 0.0f)
            continue;
        const float4 e0 = B.tape[i], e1 = B.tape[TP + i];
        V3 X = xyz(e0);
        const float eg = e0.w;
        const uint32_t et = __float_as_uint(e1.x), nv = __float_as_uint(e1.y);
        if (nv == 0) {
            if (A.g_tex && et != HIT_NONE)
                texel_add(A.g_tex, s_key, s_val, et, mul(dl, eg));
            continue;
        }
        for (int v = (int)nv - 1; v >= 0; v--) {
            const float4 *tv = B.tape + tape_row((uint32_t)v, 0) * TP + i;
            const float4 t0 = tv[0], t1 = tv[TP], t2 = tv[2 * TP], t3 = tv[3 * TP];
            const uint32_t mt = __float_as_uint(t0.w);
            const V3 thr = xyz(t0), f = xyz(t1), df = xyz(t2), N = xyz(t3);
            const float sc = t1.w, ng = t2.w;
            const uint32_t nt = __float_as_uint(t3.w);
            const V3 Av = add(N, X);
            const V3 tf = had(thr, f);
            lds_add3(s_g + mt * ADJ_NG + 4, had(dl, thr));   // dC/dE = thr
            lds_add3(s_g + mt * ADJ_NG, had(had(dl, thr), had(df, Av)));
            if (sc != 0.0f) {
                float r = sc * dot(dl, had(tf, X));
                if (r != 0.0f)
                    atomicAdd(s_g + mt * ADJ_NG + 3, r);
            }
            if (A.g_tex) {
                if (nt != HIT_NONE)
                    texel_add(A.g_tex, s_key, s_val, nt,
                              v3(dl.x * tf.x * ng, dl.y * tf.y * ng, dl.z * tf.z * ng));
                if (v == (int)nv - 1 && et != HIT_NONE)
                    texel_add(A.g_tex, s_key, s_val, et,
                              v3(dl.x * tf.x * eg, dl.y * tf.y * eg, dl.z * tf.z * eg));
            }
            const float4 mc = B.mats[3 * mt + 2];
            const V3 Ev = mc.x > 0.0f ? v3(mc.y * mc.x, mc.z * mc.x, mc.w * mc.x)
                                      : v3(0.f, 0.f, 0.f);
            X = add(Ev, had(f, Av));
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < A.n_mat * ADJ_NG; k += blockDim.x)
        if (s_g[k] != 0.0f)
            atomicAdd(&A.g_mat[k], s_g[k]);
    if (A.g_tex)
        for (uint32_t k = threadIdx.x; k < TEX_HASH; k += blockDim.x) {
            const uint32_t t = s_key[k];
            if (t == HIT_NONE)
                continue;
            atomicAdd(&A.g_tex[3 * (size_t)t], s_val[3 * k]);
            atomicAdd(&A.g_tex[3 * (size_t)t + 1], s_val[3 * k + 1]);
            atomicAdd(&A.g_tex[3 * (size_t)t + 2], s_val[3 * k + 2]);
        }
}

hipError_t launch_tape_init(hipStream_t st, float4 *tape, uint32_t P,
                            uint32_t max_depth)
{
    hipLaunchKernelGGL(k_tape_init, dim3((P + 255) / 256), dim3(256), 0, st, tape,
                       P, max_depth);
    return hipGetLastError();
}

hipError_t launch_adjoint_back(hipStream_t st, uint32_t blocks,
                               const FrameParams &fp, const ShadeBufs &B,
                               const AdjBufs &A, uint32_t P)
{
    hipLaunchKernelGGL(k_adjoint_back, dim3(blocks), dim3(256), 0, st, fp, B, A, P);
    return hipGetLastError();
}

// ---- batches: many scenes in one wavefront -----------------------------------
// The global path index space is the scenes' W*H*spp ranges one after the
// other (sample-major inside a scene); a pass takes P consecutive indices.

// Scene holding global path g: the last scene with path0 <= g, searched
// in scenes [lo, hi] (those of the pass).
__device__ __forceinline__ uint32_t scene_at(const BatchScene *bs, uint32_t lo,
                                             uint32_t hi, unsigned long long g)
{
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (bs[mid].path0 <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256)
k_braygen(const BatchScene *__restrict__ bs, const BatchCtl *__restrict__ ctl,
          unsigned long long base, uint32_t s_lo, uint32_t s_hi,
          float4 *__restrict__ q_o, float4 *__restrict__ q_d,
          float4 *__restrict__ q_t, float4 *__restrict__ acc,
          uint2 *__restrict__ slot_info, uint32_t o0, uint32_t Pl, uint32_t P,
          uint32_t *__restrict__ ctr)
{
    if (base == BASE_FROM_CTL)
        base = ctl->base;
    const unsigned long long total = ctl->total;
    const uint32_t n = base >= total ? 0u
                       : (uint32_t)min((unsigned long long)P, total - base);
    const uint32_t nl = n > o0 ? min(Pl, n - o0) : 0u;   // this lane's paths
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        reset_queues(ctr, nl);
        if (o0 == 0)
            ctr[CTR_PASS] = n;
    }
    if (i >= nl)
        return;
    const unsigned long long g = base + o0 + i;
    const uint32_t sid = scene_at(bs, s_lo, min(s_hi, ctl->n_scenes - 1), g);
    const FrameParams &F = bs[sid].fp;
    const uint32_t local = (uint32_t)(g - bs[sid].path0);
    V3 o, d;
    primary_ray(F, local % F.WH, F.sample0 + local / F.WH, o, d);
    q_o[i] = make_float4(o.x, o.y, o.z, 1.0f);
    q_d[i] = make_float4(d.x, d.y, d.z, __uint_as_float(i));
    q_t[i] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(sid));
    acc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    slot_info[i] = make_uint2(sid, local);
}

// Per (scene, pixel) of the pass, in sample order: the thread of a pixel's
// first slot in the pass adds that pixel's samples of the pass (slots
// i, i + W*H, ... inside the scene's part of the pass).
__global__ void __launch_bounds__(256)
k_baccumulate(const BatchScene *__restrict__ bs, const BatchCtl *__restrict__ ctl,
              unsigned long long base, const float4 *__restrict__ acc,
              const uint2 *__restrict__ slot_info, float *__restrict__ accum,
              const uint32_t *__restrict__ ctr, unsigned long long *__restrict__ stats)
{
    const uint32_t n = ctr[CTR_PASS];
    if (base == BASE_FROM_CTL)
        base = ctl->base;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bad = 0;
    if (i < n) {
        const uint2 si = slot_info[i];
        const BatchScene &b = bs[si.x];
        const uint32_t WH = b.fp.WH;
        // the scene's part of this pass, in local indices: [l0, l1)
        const unsigned long long g0 = base > b.path0 ? base : b.path0;
        const uint32_t l0 = (uint32_t)(g0 - b.path0);
        const unsigned long long pend = base + n, send = b.path0 + b.paths;
        const uint32_t l1 = (uint32_t)((pend < send ? pend : send) - b.path0);
        if (si.y - l0 < WH) {
            float r = 0.0f, g = 0.0f, bl = 0.0f;
            uint32_t slot = i;
            for (uint32_t l = si.y; l < l1; l += WH, slot += WH) {
                float4 v = acc[slot];
                if (isfinite(v.x) && isfinite(v.y) && isfinite(v.z)) {
                    r += v.x;
                    g += v.y;
                    bl += v.z;
                } else {
                    bad++;
                }
            }
            float *dst = accum + b.accum0 + (size_t)(si.y % WH) * 3;
            dst[0] += r;
            dst[1] += g;
            dst[2] += bl;
        }
    }
    wave_add_u64(&stats[STAT_NONFINITE], bad);
}

__global__ void k_bnext(BatchCtl *__restrict__ ctl, uint32_t P)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        ctl->base += P;
}

// ---- primary-ray AOVs and explicit rays ------------------------------------
__global__ void __launch_bounds__(TRACE_BLOCK)
k_aov(FrameParams fp, SceneDev S, const float4 *__restrict__ prims,
      const float *__restrict__ nrm, uint32_t sample, int aov,
      void *__restrict__ out)
{
    __shared__ uint32_t s_code[TRACE_SS * TRACE_BLOCK];
    uint32_t ovf_l[TRACE_OVF];
    uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= fp.WH)
        return;
    V3 o, d;
    primary_ray(fp, px, sample, o, d);
    float t, bu, bv;
    uint32_t code, nodes = 0, np = 0;
    bool hit = closest_hit<true, TRACE_SS>(S, o, d, s_code + threadIdx.x,
                                           TRACE_BLOCK, ovf_l, 1u, t,
                                           code, bu, bv, nodes, np);
    switch (aov) {
    case MRT_AOV_PRIM_ID: {
        uint32_t id = MRT_PRIM_NONE;
        if (hit)
            id = (code & HIT_PLANE) ? (MRT_PRIM_PLANE | (code & ~HIT_PLANE))
                                    : __float_as_uint(prims[(size_t)code * 3].w);
        ((uint32_t *)out)[px] = id;
        break;
    }
    case MRT_AOV_T:
        ((float *)out)[px] = hit ? t : -1.0f;
        break;
    case MRT_AOV_NODES:
        ((uint32_t *)out)[px] = nodes;
        break;
    case MRT_AOV_PRIMS:
        ((uint32_t *)out)[px] = np;
        break;
    case MRT_AOV_NORMAL: {
        V3 n = v3(0.f, 0.f, 0.f);
        if (hit) {
            V3 p;
            bool front;
            uint32_t mat;
            float4 h4 = make_float4(t, __uint_as_float(code), bu, bv);
            surface(S.planes, prims, nrm, fp.one_sided != 0, o, d, h4, p, n,
                    front, mat);
            if (fp.textured) {
                // the normal render mode shows the normal-mapped normal
                // (get_hit_register_color after apply_all_map)
                Mat m;
                m.a = m.b = m.c = make_float4(0.f, 0.f, 0.f, 0.f);
                apply_textures(S.tx, S.planes, prims, h4, p, mat, n, m);
            }
        }
        float *o3 = (float *)out + (size_t)px * 3;
        o3[0] = n.x;
        o3[1] = n.y;
        o3[2] = n.z;
        break;
    }
    default:
        break;
    }
}

__global__ void __launch_bounds__(TRACE_BLOCK)
k_trace_rays(SceneDev S, const float4 *__restrict__ prims, uint32_t n,
             const float *__restrict__ orig, const float *__restrict__ dir,
             uint32_t *__restrict__ prim_id, float *__restrict__ tout)
{
    __shared__ uint32_t s_code[TRACE_SS * TRACE_BLOCK];
    uint32_t ovf_l[TRACE_OVF];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    V3 o = v3(orig[3 * (size_t)i], orig[3 * (size_t)i + 1], orig[3 * (size_t)i + 2]);
    V3 d = v3(dir[3 * (size_t)i], dir[3 * (size_t)i + 1], dir[3 * (size_t)i + 2]);
    float t, bu, bv;
    uint32_t code, nodes = 0, np = 0;
    bool hit = closest_hit<false, TRACE_SS>(S, o, d, s_code + threadIdx.x,
                                            TRACE_BLOCK, ovf_l, 1u, t,
                                            code, bu, bv, nodes, np);
    uint32_t id = MRT_PRIM_NONE;
    if (hit)
        id = (code & HIT_PLANE) ? (MRT_PRIM_PLANE | (code & ~HIT_PLANE))
                                : __float_as_uint(prims[(size_t)code * 3].w);
    prim_id[i] = id;
    tout[i] = hit ? t : -1.0f;
}

// ---- photons (src/caustic) -------------------------------------------------------
// hash_cell (src/gpu/convert_caustic_utils.c:15-24) on integer cell coords
__device__ __forceinline__ uint32_t hash_cell(int ix, int iy, int iz,
                                              uint32_t table)
{
    uint32_t n = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u)
                 ^ ((uint32_t)iz * 83492791u);
    return n % table;
}

// One photon per thread: emission from a point light (isotropic, power
// colour * brightness / N, caustic_manager.c:30-53) or from the environment
// through the disk of the scene's bounding sphere, then at most 10 segments
// of trace_single_photon_path (trace_photon_path.c:47-75); a photon is kept
// where it first reaches a diffuse surface after a specular event.
// At least 4 waves per SIMD (128 VGPRs, spilling, instead of 184 and 2):
// 4 M photons emitted in 5.5 instead of 8.0 ms on the caustics workload.
__global__ void __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(4)))
k_photon_emit(FrameParams fp, SceneDev S, PhotonEmit pe, ShadeBufs B,
              float4 *__restrict__ out_pos, float4 *__restrict__ out_pow,
              uint32_t *__restrict__ counter, uint32_t cap)
{
    __shared__ uint32_t s_code[TRACE_SS * TRACE_BLOCK];
    uint32_t ovf_l[TRACE_OVF];
    uint32_t nodes = 0, np = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < pe.n_total;
         base += stride) {
        const uint32_t gid = base + threadIdx.x;
        bool stored = false;
        V3 sp = v3(0.f, 0.f, 0.f), spow = v3(0.f, 0.f, 0.f);
        if (gid < pe.n_total) {
            RngKey rk{fp.seed, gid, 0u, 0u, 1u};
            V3 o, d, pw;
            bool ok = true;
            const uint32_t npl = pe.n_point * pe.n_per;
            if (gid < npl) {
                uint32_t k = pe.point_idx[gid / pe.n_per];
                float4 l0 = B.lights[3 * k], l1 = B.lights[3 * k + 1];
                float nf = (float)pe.n_per;
                pw = v3(l1.x * l0.w / nf, l1.y * l0.w / nf, l1.z * l0.w / nf);
                // vec3_random_unit's law (uniform on the sphere)
                U4 u = rk(0u);
                float z = 1.0f - 2.0f * u.x;
                float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
                float ph = 6.28318530717958648f * u.y;
                d = v3(rr * cosf(ph), rr * sinf(ph), z);
                o = xyz(l0);
            } else {
                const uint32_t nw = fp.env_w - 1, nh = fp.env_h - 1;
                U4 e = rk(4u), u = rk(0u);
                uint32_t y = cdf_search_guided(fp.mis_marg, fp.mis_marg_guide,
                                               fp.guide_m, e.x);
                uint32_t x = cdf_search_guided(
                    fp.mis_cond + (size_t)y * fp.env_w,
                    fp.mis_cond_guide + (size_t)y * (fp.guide_c + 1), fp.guide_c, e.y);
                float pc = fp.mis_pcell[(size_t)y * nw + x];
                float vv = ((float)y + e.z) / (float)nh;
                float uu = ((float)x + e.w) / (float)nw;
                float el = (vv - 0.5f) * 3.14159265358979324f;
                float ph = (0.5f - uu) * 6.28318530717958648f;
                float ce = cosf(el);
                ok = pc > 0.0f && ce > 0.0f;
                V3 w = v3(ce * cosf(ph), sinf(el), ce * sinf(ph));
                float pl = pc * (float)nw * (float)nh / (19.7392088021787172f * ce);
                V3 a = fabsf(w.y) >= 0.999f ? v3(1.f, 0.f, 0.f) : v3(0.f, 1.f, 0.f);
                V3 c = cross(a, w);
                float l = sqrtf((c.x * c.x + c.y * c.y) + c.z * c.z);
                V3 t1 = v3(c.x / l, c.y / l, c.z / l);
                V3 t2 = cross(w, t1);
                float dx = 0.0f, dy = 0.0f;
                float sx = 2.0f * u.x - 1.0f, sy = 2.0f * u.y - 1.0f;
                if (sx != 0.0f || sy != 0.0f) {
                    float r, phi;
                    if (fabsf(sx) > fabsf(sy)) {
                        r = sx;
                        phi = 0.78539816339744831f * (sy / sx);
                    } else {
                        r = sy;
                        phi = 1.57079632679489662f - 0.78539816339744831f * (sx / sy);
                    }
                    dx = r * cosf(phi);
                    dy = r * sinf(phi);
                }
                const float R = pe.sr;
                o = v3(pe.sc[0] + (w.x * (R * 1.01f) + (t1.x * dx + t2.x * dy) * R),
                       pe.sc[1] + (w.y * (R * 1.01f) + (t1.y * dx + t2.y * dy) * R),
                       pe.sc[2] + (w.z * (R * 1.01f) + (t1.z * dx + t2.z * dy) * R));
                d = neg(w);
                // planes are unbounded: one that the line meets on the
                // environment side of the disk blocks the photon
                for (uint32_t k = 0; ok && k < S.n_planes; k++) {
                    float tb;
                    if (isect_pln(o, w, S.planes[2 * k], S.planes[2 * k + 1], tb))
                        ok = false;
                }
                float flux = 3.14159265358979324f * R * R / (pl * (float)pe.n_per);
                pw = ok ? mul(env_texel(fp, w), flux) : v3(0.f, 0.f, 0.f);
            }
            for (uint32_t depth = 0; ok && depth < 10; depth++) {
                rk.bounce = depth;
                float t, bu, bv;
                uint32_t code;
                if (!closest_hit<false, TRACE_SS>(S, o, d, s_code + threadIdx.x,
                                                  TRACE_BLOCK, ovf_l, 1u,
                                                  t, code, bu, bv, nodes, np))
                    break;
                V3 p, nrm;
                bool front;
                uint32_t mat;
                surface(B.planes, B.prims, B.nrm, fp.one_sided != 0, o, d,
                        make_float4(t, __uint_as_float(code), bu, bv), p, nrm,
                        front, mat);
                o = p;
                Mat m;
                m.a = B.mats[3 * mat];
                m.b = B.mats[3 * mat + 1];
                m.c = B.mats[3 * mat + 2];
                if (fp.textured)
                    apply_textures(B.tx, B.planes, B.prims,
                                   make_float4(t, __uint_as_float(code), bu, bv),
                                   p, mat, nrm, m);
                if (scatter(rk, fp.ior_all, nrm, front, m, d, pw)) {
                    pw = had(pw, v3(m.a.x, m.a.y, m.a.z));
                    if (depth != 0) {
                        stored = true;
                        sp = p;
                        spow = pw;
                    }
                    break;
                }
            }
        }
        uint32_t slot = wave_alloc(counter, stored ? 1u : 0u);
        if (stored && slot < cap) {
            out_pos[slot] = make_float4(sp.x, sp.y, sp.z, __uint_as_float(gid));
            out_pow[slot] = make_float4(spow.x, spow.y, spow.z, 0.0f);
        }
    }
}

// Sort key of a photon: (bucket << 32) | photon id, so the sort is
// deterministic; value = its slot.
__global__ void __launch_bounds__(256)
k_photon_keys(const float4 *__restrict__ pos, uint32_t n, float cell,
              uint32_t table, unsigned long long *__restrict__ keys,
              uint32_t *__restrict__ vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float4 p = pos[i];
    uint32_t b = hash_cell((int)floorf(p.x / cell), (int)floorf(p.y / cell),
                           (int)floorf(p.z / cell), table);
    keys[i] = ((unsigned long long)b << 32) | __float_as_uint(p.w);
    vals[i] = i;
}

__global__ void __launch_bounds__(256)
k_photon_reorder(const float4 *__restrict__ pos, const float4 *__restrict__ pow,
                 const unsigned long long *__restrict__ keys,
                 const uint32_t *__restrict__ vals, uint32_t n,
                 float4 *__restrict__ rec,
                 uint2 *__restrict__ range, uint32_t *__restrict__ occ)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t s = vals[i];
    rec[2 * (size_t)i] = pos[s];
    rec[2 * (size_t)i + 1] = pow[s];
    uint32_t b = (uint32_t)(keys[i] >> 32);
    if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != b) {
        range[b].x = i;
        atomicOr(&occ[b >> 5], 1u << (b & 31u));
    }
    if (i + 1 == n || (uint32_t)(keys[i + 1] >> 32) != b)
        range[b].y = i + 1;
}

__global__ void __launch_bounds__(256)
k_fill_u32(uint32_t *__restrict__ p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// estimate_caustics_physically_based / calculate_caustic_radiance
// (src/caustic/get_caustic.c:37-88) over the cells around x: the k
// nearest photons within squared distance `radius` (k > 0; nothing unless
// k are found) or all photons within distance `radius` (k == 0), summed
// and divided by pi r^2 and the norm factor.  A photon counts only from
// its own cell, so a hash collision between neighbours cannot count it
// twice.  kd / ki are this thread's k-NN lists in LDS (stride `stride`).
constexpr int KNN_MAX = 32;
template <bool KNN>
__device__ __forceinline__ V3 caustic_at(const PhotonMapDev &pm, V3 x,
                                         float *kd, uint32_t *ki, int stride,
                                         uint32_t *examined = nullptr,
                                         uint32_t *cells = nullptr,
                                         uint32_t *lookups = nullptr)
{
    // the cells meeting the cube [x - s, x + s]^3, s = the search distance
    // (the cell edge, widened by 1e-4 against rounding): two or three per
    // axis instead of 27 cells
    const float sd = pm.cell * 1.0001f;
    const int x0 = (int)floorf((x.x - sd) / pm.cell), x1 = (int)floorf((x.x + sd) / pm.cell);
    const int y0 = (int)floorf((x.y - sd) / pm.cell), y1 = (int)floorf((x.y + sd) / pm.cell);
    const int z0 = (int)floorf((x.z - sd) / pm.cell), z1 = (int)floorf((x.z + sd) / pm.cell);
    const bool knn = KNN;
    const uint32_t K = pm.k < (uint32_t)KNN_MAX ? pm.k : (uint32_t)KNN_MAX;
    const float cut = knn ? pm.radius : pm.radius * pm.radius;
    V3 sum = v3(0.f, 0.f, 0.f);
    uint32_t cnt = 0;
    // the search reaches at most distance rmax from x (k-NN: squared
    // distances up to the radius; by radius: the radius): a cell whose box
    // lies farther holds no photon the tests below would take, so it is
    // skipped (about a quarter of the 27 cells; the sums are unchanged)
    const float rmax2 = cut * 1.001f;
    for (int iz = z0; iz <= z1; iz++)
        for (int iy = y0; iy <= y1; iy++)
            for (int ix = x0; ix <= x1; ix++) {
                {
                    const float bx = fmaxf(fmaxf((float)ix * pm.cell - x.x,
                                                 x.x - (float)(ix + 1) * pm.cell), 0.0f);
                    const float by = fmaxf(fmaxf((float)iy * pm.cell - x.y,
                                                 x.y - (float)(iy + 1) * pm.cell), 0.0f);
                    const float bz = fmaxf(fmaxf((float)iz * pm.cell - x.z,
                                                 x.z - (float)(iz + 1) * pm.cell), 0.0f);
                    if ((bx * bx + by * by) + bz * bz > rmax2)
                        continue;
                }
                uint32_t b = hash_cell(ix, iy, iz, pm.table);
                // one photon of the bucket (P its position, W its power when
                // gathering by radius), in bucket order
                auto take = [&](float4 P, float4 W, uint32_t j) {
                    if ((int)floorf(P.x / pm.cell) != ix
                        || (int)floorf(P.y / pm.cell) != iy
                        || (int)floorf(P.z / pm.cell) != iz)
                        return;
                    float ddx = P.x - x.x, ddy = P.y - x.y, ddz = P.z - x.z;
                    float d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                    if (!(d2 <= cut))
                        return;
                    if (!knn) {
                        sum = add(sum, xyz(W));
                        return;
                    }
                    if (cnt == K && !(d2 < kd[(K - 1) * stride]))
                        return;
                    uint32_t pos = cnt < K ? cnt++ : K - 1;
                    while (pos > 0 && d2 < kd[(pos - 1) * stride]) {
                        kd[pos * stride] = kd[(pos - 1) * stride];
                        ki[pos * stride] = ki[(pos - 1) * stride];
                        pos--;
                    }
                    kd[pos * stride] = d2;
                    ki[pos * stride] = j;
                };
                // the occupancy bits (2 bits per photon, L2-resident) turn
                // most empty cells away before their bucket bounds are read
                if (examined)
                    *lookups += 1;
                if (!((pm.occ[b >> 5] >> (b & 31u)) & 1u))
                    continue;
                const uint2 br = pm.range[b];
                const uint32_t je = br.y;
                uint32_t j = br.x;
                if (examined) {
                    *examined += je - j;
                    *cells += 1;
                }
                // four records in flight at a time instead of one
                for (; j + 4 <= je; j += 4) {
                    float4 P[4], W[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        P[u] = pm.rec[2 * (size_t)(j + u)];
                        W[u] = knn ? make_float4(0.f, 0.f, 0.f, 0.f)
                                   : pm.rec[2 * (size_t)(j + u) + 1];
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        take(P[u], W[u], j + u);
                }
                for (; j < je; j++)
                    take(pm.rec[2 * (size_t)j],
                         knn ? make_float4(0.f, 0.f, 0.f, 0.f)
                             : pm.rec[2 * (size_t)j + 1], j);
            }
    float area;
    if (knn) {
        if (cnt < K)
            return v3(0.f, 0.f, 0.f);
        area = 3.14159265358979324f * kd[(K - 1) * stride];
        if (area < 1e-9f)
            return v3(0.f, 0.f, 0.f);
        for (uint32_t j = 0; j < K; j++)
            sum = add(sum, xyz(pm.rec[2 * (size_t)ki[j * stride] + 1]));
    } else {
        area = 3.14159265358979324f * cut;
    }
    return v3(sum.x / area / pm.norm, sum.y / area / pm.norm,
              sum.z / area / pm.norm);
}

constexpr int CAUSTIC_BLOCK = 128;

// The k-NN lists live in LDS; the fixed-radius gather needs none, so its
// instantiation keeps the CU's LDS free for more resident waves.
// COUNT: also count the gathers and the photon records they examine.
template <bool KNN, bool COUNT>
__global__ void __launch_bounds__(CAUSTIC_BLOCK)
k_caustic(FrameParams fp, ShadeBufs B, uint32_t *__restrict__ ctr,
          unsigned long long *__restrict__ stats)
{
    __shared__ float s_d[KNN ? KNN_MAX * CAUSTIC_BLOCK : 1];
    __shared__ uint32_t s_i[KNN ? KNN_MAX * CAUSTIC_BLOCK : 1];
    const uint32_t n = min(ctr[ctr_pend(fp.bounce)], B.cap_paths);
    uint32_t queries = 0, examined = 0, cells = 0, lookups = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += gridDim.x * blockDim.x) {
        V3 c = caustic_at<KNN>(fp.pm, xyz(B.p_x[r]), s_d + (KNN ? threadIdx.x : 0),
                               s_i + (KNN ? threadIdx.x : 0), CAUSTIC_BLOCK,
                               COUNT ? &examined : nullptr, &cells, &lookups);
        B.p_k[r] = make_float4(c.x, c.y, c.z, 0.0f);
        queries++;
    }
    if (COUNT) {
        wave_add_u64(&stats[STAT_GATHER_QUERIES], queries);
        wave_add_u64(&stats[STAT_GATHER_PHOTONS], examined);
        wave_add_u64(&stats[STAT_GATHER_CELLS], cells);
        wave_add_u64(&stats[STAT_GATHER_LOOKUPS], lookups);
    }
}

template <bool KNN>
__global__ void __launch_bounds__(CAUSTIC_BLOCK)
k_caustic_at(PhotonMapDev pm, uint32_t n, const float *__restrict__ pts,
             float *__restrict__ out)
{
    __shared__ float s_d[KNN ? KNN_MAX * CAUSTIC_BLOCK : 1];
    __shared__ uint32_t s_i[KNN ? KNN_MAX * CAUSTIC_BLOCK : 1];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    V3 c = caustic_at<KNN>(pm, v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]),
                           s_d + (KNN ? threadIdx.x : 0),
                           s_i + (KNN ? threadIdx.x : 0), CAUSTIC_BLOCK);
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
}

// ---- launchers -----------------------------------------------------------------
hipError_t launch_photon_emit(hipStream_t st, const FrameParams &fp,
                              const SceneDev &S, const PhotonEmit &pe,
                              const ShadeBufs &B, float4 *pos, float4 *pow,
                              uint32_t *counter, uint32_t cap)
{
    uint32_t blocks = (pe.n_total + TRACE_BLOCK - 1) / TRACE_BLOCK;
    if (blocks > 65535u)
        blocks = 65535u;
    hipLaunchKernelGGL(k_photon_emit, dim3(blocks), dim3(TRACE_BLOCK), 0, st, fp,
                       S, pe, B, pos, pow, counter, cap);
    return hipGetLastError();
}

hipError_t launch_photon_keys(hipStream_t st, const float4 *pos, uint32_t n,
                              float cell, uint32_t table,
                              unsigned long long *keys, uint32_t *vals)
{
    hipLaunchKernelGGL(k_photon_keys, dim3((n + 255) / 256), dim3(256), 0, st,
                       pos, n, cell, table, keys, vals);
    return hipGetLastError();
}

hipError_t launch_photon_reorder(hipStream_t st, const float4 *pos,
                                 const float4 *pow,
                                 const unsigned long long *keys,
                                 const uint32_t *vals, uint32_t n,
                                 float4 *rec, uint2 *range, uint32_t *occ)
{
    hipLaunchKernelGGL(k_photon_reorder, dim3((n + 255) / 256), dim3(256), 0,
                       st, pos, pow, keys, vals, n, rec, range, occ);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, size_t n)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 65535)
        blocks = 65535;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, st, p,
                       v, n);
    return hipGetLastError();
}

hipError_t launch_caustic(hipStream_t st, uint32_t blocks,
                          const FrameParams &fp, const ShadeBufs &B,
                          uint32_t *ctr, bool count, unsigned long long *stats)
{
#define MRT_GATHER(K, C)                                                      \
    hipLaunchKernelGGL((k_caustic<K, C>), dim3(blocks), dim3(CAUSTIC_BLOCK), 0, \
                       st, fp, B, ctr, stats)
    if (fp.pm.k) {
        if (count) MRT_GATHER(true, true); else MRT_GATHER(true, false);
    } else {
        if (count) MRT_GATHER(false, true); else MRT_GATHER(false, false);
    }
#undef MRT_GATHER
    return hipGetLastError();
}

hipError_t launch_caustic_at(hipStream_t st, const PhotonMapDev &pm,
                             uint32_t n, const float *pts, float *out)
{
    dim3 g((n + CAUSTIC_BLOCK - 1) / CAUSTIC_BLOCK);
    if (pm.k)
        hipLaunchKernelGGL(k_caustic_at<true>, g, dim3(CAUSTIC_BLOCK), 0, st, pm, n,
                           pts, out);
    else
        hipLaunchKernelGGL(k_caustic_at<false>, g, dim3(CAUSTIC_BLOCK), 0, st, pm, n,
                           pts, out);
    return hipGetLastError();
}
// ---- batch merge ----------------------------------------------------------------
// One scene's 4-wide nodes and records copied into the shared arrays at
// node0 / prim0: inner links move by node0, leaf ranges by prim0, material
// indices (record word 11, or a quadric's shading words 1-3) by mat0.
__global__ void __launch_bounds__(256)
k_rebase(float4 *__restrict__ nodes, uint32_t n_nodes, float4 *__restrict__ prims,
         float *__restrict__ nrm, uint32_t n_prims, uint32_t node0, uint32_t prim0,
         uint32_t mat0)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_nodes) {
        float4 c = nodes[(size_t)i * 8 + 6];
        float *cf = &c.x;
        for (int j = 0; j < 4; j++) {
            int32_t ch = __float_as_int(cf[j]);
            if (ch >= 0) {
                ch += (int32_t)node0;
            } else {
                uint32_t v = ~(uint32_t)ch;
                v = (((v >> 3) + prim0) << 3) | (v & 7u);
                ch = (int32_t)~v;
            }
            cf[j] = __int_as_float(ch);
        }
        nodes[(size_t)i * 8 + 6] = c;
    }
    if (i < n_prims) {
        uint32_t kind = __float_as_uint(prims[(size_t)i * 3 + 1].w);
        if (kind >= MRT_PRIMKIND_QUADRIC) {
            float *a = nrm + (size_t)i * 9;
            for (int j = 1; j <= 3; j++)
                a[j] = __uint_as_float(__float_as_uint(a[j]) + mat0);
        } else {
            float4 &cw = prims[(size_t)i * 3 + 2];
            cw.w = __uint_as_float(__float_as_uint(cw.w) + mat0);
        }
    }
}

hipError_t launch_rebase(hipStream_t st, float4 *nodes, uint32_t n_nodes,
                         float4 *prims, float *nrm, uint32_t n_prims,
                         uint32_t node0, uint32_t prim0, uint32_t mat0)
{
    uint32_t n = n_nodes > n_prims ? n_nodes : n_prims;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_rebase, dim3((n + 255) / 256), dim3(256), 0, st, nodes,
                       n_nodes, prims, nrm, n_prims, node0, prim0, mat0);
    return hipGetLastError();
}

// ---- HDRI importance tables on the device -----------------------------------
// The same double-precision operations, in the same order, as the host
// builders (hdri_tables.c: make_importance_sampling_map's tables and the MIS
// tables), so the results are bit for bit the host's: every sequential sum
// stays one thread's loop, divisions are IEEE, and the inverse-CDF entries
// come from a binary search for the first CDF value >= i / (n-1), which is
// the index the host's forward scan stops at (the CDFs never decrease).

// get_grey_value (calc_pdf.c:16-35) of texel i; also the MIS luminance.
__device__ __forceinline__ double env_gray(const float4 *tex, size_t i, double sc)
{
    float4 t = tex[i];
    return 0.2126 * ((double)t.x * sc) + 0.7152 * ((double)t.y * sc)
           + 0.0722 * ((double)t.z * sc);
}

// One thread per row: the row average (calc_raw_average) and, for the
// h - 1 MIS rows, the sum of the cell weights lum * band[y].
__global__ void __launch_bounds__(64)
k_env_rows(const float4 *__restrict__ tex, uint32_t w, uint32_t h, double sc,
           const double *__restrict__ band, double *__restrict__ rowavg,
           double *__restrict__ mrow)
{
    uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h)
        return;
    double s = 0.0;
    for (uint32_t x = 0; x < w; x++)
        s += env_gray(tex, (size_t)y * w + x, sc);
    rowavg[y] = s / w;
    if (y + 1 < h) {
        double b = band[y], m = 0.0;
        for (uint32_t x = 0; x + 1 < w; x++) {
            double lum = env_gray(tex, (size_t)y * w + x, sc);
            m += lum > 0.0 ? lum * b : 0.0;
        }
        mrow[y] = m;
    }
}

// One thread: the column average, the normalised marginal CDF of the
// reference tables, and the MIS row CDF.  scal = {colavg, total}.
__global__ void k_env_marginal(uint32_t h, const double *__restrict__ rowavg,
                               const double *__restrict__ mrow,
                               double *__restrict__ mcdf,
                               float *__restrict__ mis_marg,
                               double *__restrict__ scal)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    double colavg = 0.0;
    for (uint32_t y = 0; y < h; y++)
        colavg += rowavg[y];
    colavg = colavg / h;
    double sum = 0.0;
    for (uint32_t y = 0; y < h; y++) {
        sum += rowavg[y] / (colavg / 255);
        mcdf[y] = sum;
    }
    if (sum > 0.0)
        for (uint32_t y = 0; y < h; y++)
            mcdf[y] /= sum;
    const uint32_t nh = h - 1;
    double total = 0.0;
    for (uint32_t y = 0; y < nh; y++)
        total += mrow[y];
    double acc = 0.0;
    mis_marg[0] = 0.0f;
    for (uint32_t y = 0; y < nh; y++) {
        acc += mrow[y];
        mis_marg[y + 1] = total > 0.0 ? (float)(acc / total) : (float)(y + 1) / nh;
    }
    mis_marg[nh] = 1.0f;
    scal[0] = colavg;
    scal[1] = total;
}

// One thread per row: the normalised conditional CDF of the reference
// tables (process_conditional_row) and the MIS column CDF of the row.
__global__ void __launch_bounds__(64)
k_env_cond(const float4 *__restrict__ tex, uint32_t w, uint32_t h, double sc,
           const double *__restrict__ band, const double *__restrict__ rowavg,
           const double *__restrict__ mrow, double *__restrict__ ccdf,
           float *__restrict__ mis_cond)
{
    uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h)
        return;
    double *c = ccdf + (size_t)y * w;
    double sum = 0.0;
    for (uint32_t x = 0; x < w; x++) {
        sum += env_gray(tex, (size_t)y * w + x, sc) / (rowavg[y] / 255);
        c[x] = sum;
    }
    if (sum > 0.0)
        for (uint32_t x = 0; x < w; x++)
            c[x] /= sum;
    if (y + 1 < h) {
        const uint32_t nw = w - 1;
        float *mc = mis_cond + (size_t)y * w;
        double b = band[y], r = mrow[y], ra = 0.0;
        mc[0] = 0.0f;
        for (uint32_t x = 0; x < nw; x++) {
            double lum = env_gray(tex, (size_t)y * w + x, sc);
            ra += lum > 0.0 ? lum * b : 0.0;
            mc[x + 1] = r > 0.0 ? (float)(ra / r) : (float)(x + 1) / nw;
        }
        mc[nw] = 1.0f;
    }
}

// First index x with cdf[x] >= u, or n when none does.
__device__ __forceinline__ uint32_t cdf_first_ge(const double *cdf, uint32_t n,
                                                 double u)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] >= u)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// One thread per texel: pdf_joint, the inverse conditional CDF entry, the
// MIS cell probability, and (first h threads) the inverse marginal entry.
// Entries whose target no CDF value reaches stay 0, as the host's calloc.
__global__ void __launch_bounds__(256)
k_env_point(const float4 *__restrict__ tex, uint32_t w, uint32_t h, double sc,
            const double *__restrict__ band, const double *__restrict__ mcdf,
            const double *__restrict__ ccdf, const double *__restrict__ scal,
            float *__restrict__ pdf, float *__restrict__ marg_inv,
            float *__restrict__ cond_inv, float *__restrict__ pcell)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h)
        return;
    uint32_t y = (uint32_t)(i / w), x = (uint32_t)(i % w);
    double g = env_gray(tex, i, sc);
    pdf[i] = (float)(g / scal[0]);
    uint32_t k = cdf_first_ge(ccdf + (size_t)y * w, w, (double)x / (double)(w - 1));
    cond_inv[i] = k < w ? (float)((double)k / (double)(w - 1)) : 0.0f;
    if (y + 1 < h && x + 1 < w) {
        double v = g > 0.0 ? g * band[y] : 0.0;
        pcell[(size_t)y * (w - 1) + x] = scal[1] > 0.0 ? (float)(v / scal[1]) : 0.0f;
    }
    if (i < h) {
        uint32_t m = cdf_first_ge(mcdf, h, (double)i / (double)(h - 1));
        marg_inv[i] = m < h ? (float)((double)m / (double)(h - 1)) : 0.0f;
    }
}

__global__ void __launch_bounds__(256)
k_rgb_to_float4(const float *__restrict__ rgb, float4 *__restrict__ tex, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        tex[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.0f);
}

hipError_t launch_rgb_to_float4(hipStream_t st, const float *rgb, float4 *tex,
                                size_t n)
{
    hipLaunchKernelGGL(k_rgb_to_float4, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, st, rgb, tex, n);
    return hipGetLastError();
}

// One thread per guide entry: the marginal's Gm + 1, then every
// conditional row's Gc + 1.
__global__ void __launch_bounds__(256)
k_env_guides(const float *__restrict__ mis_marg, const float *__restrict__ mis_cond,
             uint32_t w, uint32_t h, uint32_t Gm, uint32_t Gc,
             uint32_t *__restrict__ gm, uint32_t *__restrict__ gc)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nh = h - 1, nw = w - 1;
    const size_t nm = (size_t)Gm + 1, nc = (size_t)nh * (Gc + 1);
    if (i < nm) {
        const uint32_t k = (uint32_t)i;
        gm[k] = k < Gm ? cdf_search(mis_marg, nh, (float)k / (float)Gm) : nh - 1;
    } else if (i < nm + nc) {
        const size_t j = i - nm;
        const uint32_t y = (uint32_t)(j / (Gc + 1)), k = (uint32_t)(j % (Gc + 1));
        gc[j] = k < Gc ? cdf_search(mis_cond + (size_t)y * w, nw,
                                    (float)k / (float)Gc)
                       : nw - 1;
    }
}

hipError_t launch_env_guides(hipStream_t st, const float *mis_marg,
                             const float *mis_cond, uint32_t w, uint32_t h,
                             uint32_t Gm, uint32_t Gc, uint32_t *gm, uint32_t *gc)
{
    const size_t n = (size_t)Gm + 1 + (size_t)(h - 1) * (Gc + 1);
    hipLaunchKernelGGL(k_env_guides, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, st, mis_marg, mis_cond, w, h, Gm, Gc, gm, gc);
    return hipGetLastError();
}

hipError_t launch_env_tables(hipStream_t st, const float4 *tex, uint32_t w,
                             uint32_t h, double sc, const double *band,
                             double *scratch, float *pdf, float *marg_inv,
                             float *cond_inv, float *mis_marg, float *mis_cond,
                             float *mis_pcell)
{
    double *rowavg = scratch, *mrow = rowavg + h, *mcdf = mrow + h,
           *scal = mcdf + h, *ccdf = scal + 2;
    hipLaunchKernelGGL(k_env_rows, dim3((h + 63) / 64), dim3(64), 0, st, tex, w,
                       h, sc, band, rowavg, mrow);
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(1), 0, st, h, rowavg, mrow,
                       mcdf, mis_marg, scal);
    hipLaunchKernelGGL(k_env_cond, dim3((h + 63) / 64), dim3(64), 0, st, tex, w,
                       h, sc, band, rowavg, mrow, ccdf, mis_cond);
    size_t n = (size_t)w * h;
    hipLaunchKernelGGL(k_env_point, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, st, tex, w, h, sc, band, mcdf, ccdf, scal, pdf,
                       marg_inv, cond_inv, mis_pcell);
    return hipGetLastError();
}

hipError_t launch_raygen(hipStream_t st, const FrameParams &fp, float4 *q_o,
                         float4 *q_d, float4 *q_t, float4 *acc, uint32_t P,
                         uint32_t *ctr)
{
    dim3 grid((P + 255) / 256);
    hipLaunchKernelGGL(k_raygen, grid, dim3(256), 0, st, fp, q_o, q_d, q_t,
                       acc, P, ctr);
    return hipGetLastError();
}

// Launches MRT_GO(C, Q) with the template flags the two run-time flags pick.
#define MRT_DISPATCH2(c, q)                                                   \
    do {                                                                      \
        if (c) {                                                              \
            if (q) MRT_GO(true, true); else MRT_GO(true, false);              \
        } else {                                                              \
            if (q) MRT_GO(false, true); else MRT_GO(false, false);            \
        }                                                                     \
    } while (0)

// The persistent trace grids take their stack overflow columns from ovf.
static bool ovf_fits(const StackOvf &ovf, uint32_t blocks)
{
    return ovf.buf != nullptr && blocks <= ovf.blocks;
}

hipError_t launch_trace(hipStream_t st, uint32_t blocks, bool count,
                        const SceneDev &S, const StackOvf &ovf,
                        const TraceWork &w, uint32_t *ctr,
                        unsigned long long *stats, bool top, bool refill,
                        uint32_t refill_blocks)
{
    if (!ovf_fits(ovf, blocks))
        return hipErrorInvalidValue;
    const BatchScene *nb = nullptr;
    const uint2 *ni = nullptr;
    // the tree's top levels from LDS: needs the root at node 0 and the
    // plain (triangle / sphere) leaf code; a scene without a BVH (planes
    // only) takes k_trace_top
    if (top && refill && S.n_nodes > 0 && S.root == 0 && !(S.quad | S.alpha)
        && refill_blocks > 0) {
        if (!ovf_fits(ovf, refill_blocks))
            return hipErrorInvalidValue;
        // the grid: the blocks that stay resident at the kernel's 6 waves
        // per SIMD (refill_blocks_per_cu), not the 8-wave kernels' count
        const uint32_t ntop = min((uint32_t)TOP_NODES, S.n_nodes);
        if (count)
            hipLaunchKernelGGL((k_trace_refill<true, false>), dim3(refill_blocks),
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats, ovf.buf,
                               (const BatchScene *)nullptr, 0u, 0u, ntop);
        else
            hipLaunchKernelGGL((k_trace_refill<false, false>), dim3(refill_blocks),
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats, ovf.buf,
                               (const BatchScene *)nullptr, 0u, 0u, ntop);
        return hipGetLastError();
    }
    if (top && S.root == 0 && !(S.quad | S.alpha)) {
        if (count)
            hipLaunchKernelGGL(k_trace_top<true>, dim3(blocks), dim3(TRACE_BLOCK),
                               0, st, S, w, ctr, stats, ovf.buf);
        else
            hipLaunchKernelGGL(k_trace_top<false>, dim3(blocks), dim3(TRACE_BLOCK),
                               0, st, S, w, ctr, stats, ovf.buf);
        return hipGetLastError();
    }
#define MRT_GO(C, Q)                                                          \
    do {                                                                      \
        if (Q)                                                                \
            hipLaunchKernelGGL((k_trace_ext<C, false>), dim3(blocks),         \
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats, nb, \
                               ni, ovf.buf);                                  \
        else                                                                  \
            hipLaunchKernelGGL((k_trace<C, false>), dim3(blocks),             \
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats, nb, \
                               ni, ovf.buf);                                  \
    } while (0)
    MRT_DISPATCH2(count, (S.quad | S.alpha) != 0);
#undef MRT_GO
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t st, uint32_t blocks, const FrameParams &fp,
                        const ShadeBufs &B, uint32_t *ctr)
{
#define MRT_SHADE(E, T)                                                       \
    hipLaunchKernelGGL((k_shade<false, E, T>), dim3(blocks), dim3(SHADE_BLOCK), 0, \
                       st, fp, B, ctr)
    const int env = !fp.has_env ? 0 : fp.estimator == MRT_EST_REFERENCE ? 2 : 1;
    if (B.tape) {
        // the adjoint's recording replay: no textures, MIS or no environment
        if (fp.textured || env == 2)
            return hipErrorInvalidValue;
        if (env == 1)
            hipLaunchKernelGGL((k_shade<false, 1, false, true>), dim3(blocks),
                               dim3(SHADE_BLOCK), 0, st, fp, B, ctr);
        else
            hipLaunchKernelGGL((k_shade<false, 0, false, true>), dim3(blocks),
                               dim3(SHADE_BLOCK), 0, st, fp, B, ctr);
        return hipGetLastError();
    }
    if (fp.textured) {
        if (env == 2) MRT_SHADE(2, true); else if (env == 1) MRT_SHADE(1, true);
        else MRT_SHADE(0, true);
    } else {
        if (env == 2) MRT_SHADE(2, false); else if (env == 1) MRT_SHADE(1, false);
        else MRT_SHADE(0, false);
    }
#undef MRT_SHADE
    return hipGetLastError();
}

hipError_t launch_resolve(hipStream_t st, uint32_t blocks,
                          const FrameParams &fp, const ShadeBufs &B,
                          uint32_t *ctr)
{
    if (B.tape)
        hipLaunchKernelGGL((k_resolve<false, true>), dim3(blocks), dim3(256), 0, st,
                           fp, B, ctr);
    else
        hipLaunchKernelGGL((k_resolve<false, false>), dim3(blocks), dim3(256), 0,
                           st, fp, B, ctr);
    return hipGetLastError();
}

hipError_t launch_braygen(hipStream_t st, const ShadeBufs &B,
                          const BatchCtl *ctl, unsigned long long base,
                          uint32_t s_lo, uint32_t s_hi, float4 *q_o, float4 *q_d,
                          float4 *q_t, uint32_t o0, uint32_t Pl, uint32_t P,
                          uint32_t *ctr)
{
    if (s_lo > s_hi)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_braygen, dim3((Pl + 255) / 256), dim3(256), 0, st, B.bs,
                       ctl, base, s_lo, s_hi, q_o, q_d, q_t, B.acc, B.slot_info,
                       o0, Pl, P, ctr);
    return hipGetLastError();
}

hipError_t launch_btrace(hipStream_t st, uint32_t blocks, bool count,
                         const SceneDev &S, const StackOvf &ovf,
                         const ShadeBufs &B, const TraceWork &w,
                         uint32_t *ctr, unsigned long long *stats,
                         uint32_t refill_blocks, uint32_t n_scenes,
                         uint32_t top_base, uint32_t ntop)
{
    if (refill_blocks > 0) {
        if (!ovf_fits(ovf, refill_blocks) || n_scenes > BATCH_LDS_SCENES
            || (S.quad | S.alpha) || w.q_t == nullptr || ntop > (uint32_t)TOP_NODES)
            return hipErrorInvalidValue;
        if (count)
            hipLaunchKernelGGL((k_trace_refill<true, true>), dim3(refill_blocks),
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats, ovf.buf,
                               (const BatchScene *)B.bs, n_scenes, top_base, ntop);
        else
            hipLaunchKernelGGL((k_trace_refill<false, true>), dim3(refill_blocks),
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats, ovf.buf,
                               (const BatchScene *)B.bs, n_scenes, top_base, ntop);
        return hipGetLastError();
    }
    if (!ovf_fits(ovf, blocks))
        return hipErrorInvalidValue;
#define MRT_GO(C, Q)                                                          \
    do {                                                                      \
        if (Q)                                                                \
            hipLaunchKernelGGL((k_trace_ext<C, true>), dim3(blocks),          \
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats,    \
                               B.bs, (const uint2 *)B.slot_info, ovf.buf);    \
        else                                                                  \
            hipLaunchKernelGGL((k_trace<C, true>), dim3(blocks),              \
                               dim3(TRACE_BLOCK), 0, st, S, w, ctr, stats,    \
                               B.bs, (const uint2 *)B.slot_info, ovf.buf);    \
    } while (0)
    MRT_DISPATCH2(count, (S.quad | S.alpha) != 0);
#undef MRT_GO
    return hipGetLastError();
}

hipError_t launch_bshade(hipStream_t st, uint32_t blocks, const FrameParams &fp,
                         const ShadeBufs &B, uint32_t *ctr)
{
    // batches take no HDRI and no textured material
    hipLaunchKernelGGL((k_shade<true, 0, false>), dim3(blocks), dim3(SHADE_BLOCK), 0, st, fp,
                       B, ctr);
    return hipGetLastError();
}

hipError_t launch_bresolve(hipStream_t st, uint32_t blocks,
                           const FrameParams &fp, const ShadeBufs &B,
                           uint32_t *ctr)
{
    hipLaunchKernelGGL((k_resolve<true, false>), dim3(blocks), dim3(256), 0, st, fp, B,
                       ctr);
    return hipGetLastError();
}

hipError_t launch_baccumulate(hipStream_t st, const ShadeBufs &B,
                              const BatchCtl *ctl, unsigned long long base,
                              float *accum, uint32_t P, uint32_t *ctr,
                              unsigned long long *stats)
{
    hipLaunchKernelGGL(k_baccumulate, dim3((P + 255) / 256), dim3(256), 0, st,
                       B.bs, ctl, base, (const float4 *)B.acc,
                       (const uint2 *)B.slot_info, accum, (const uint32_t *)ctr,
                       stats);
    return hipGetLastError();
}

hipError_t launch_bnext(hipStream_t st, BatchCtl *ctl, uint32_t P)
{
    hipLaunchKernelGGL(k_bnext, dim3(1), dim3(64), 0, st, ctl, P);
    return hipGetLastError();
}

hipError_t launch_advance(hipStream_t st, uint32_t *ctr,
                          unsigned long long *stats, uint32_t cap,
                          uint32_t bounce)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, ctr, stats, cap,
                       bounce);
    return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t st, const float4 *acc, float *accum,
                             uint32_t WH, uint32_t nb,
                             unsigned long long *stats)
{
    hipLaunchKernelGGL(k_accumulate, dim3((WH + 255) / 256), dim3(256), 0, st,
                       acc, accum, WH, nb, stats);
    return hipGetLastError();
}

hipError_t launch_aov(hipStream_t st, const FrameParams &fp, const SceneDev &S,
                      const float4 *prims, const float *nrm, uint32_t sample,
                      int aov, void *out)
{
    hipLaunchKernelGGL(k_aov, dim3((fp.WH + TRACE_BLOCK - 1) / TRACE_BLOCK),
                       dim3(TRACE_BLOCK), 0, st, fp, S, prims, nrm, sample, aov,
                       out);
    return hipGetLastError();
}

hipError_t launch_trace_rays(hipStream_t st, const SceneDev &S,
                             const float4 *prims, uint32_t n,
                             const float *orig, const float *dir,
                             uint32_t *prim_id, float *t)
{
    hipLaunchKernelGGL(k_trace_rays, dim3((n + TRACE_BLOCK - 1) / TRACE_BLOCK),
                       dim3(TRACE_BLOCK), 0, st, S, prims, n, orig, dir,
                       prim_id, t);
    return hipGetLastError();
}

int refill_blocks_per_cu(bool batch)
{
    int nb = 0;
    hipError_t e = batch
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, k_trace_refill<false, true>, TRACE_BLOCK, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, k_trace_refill<false, false>, TRACE_BLOCK, 0);
    if (e != hipSuccess || nb < 1)
        nb = 4;
    return nb;
}

int trace_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_trace<false, false>, TRACE_BLOCK, 0) != hipSuccess
        || nb < 1)
        nb = 4;
    return nb;
}

}  // namespace mrt
