// BVH construction on the device (gfx950): replaces the host build of the
// reference (src/ray_tracing/bvh/bvh_build.c:17-124, qshort_axis.c:88-111)
// for large scenes.  Pipeline:
//   1. primitive boxes and centroids (the host builder's bounds: triangle
//      vertex boxes, spheres one ulp outward, quadric solids widened);
//   2. 63-bit Morton codes of the centroids, rocPRIM radix sort;
//   3. PLOC (Meister and Bittner, "Parallel Locally-Ordered Clustering for
//      Bounding Volume Hierarchy Construction", TVCG 2018): each cluster
//      looks for the cluster in a window of +-R positions whose union box
//      has the smallest surface area; mutual nearest neighbours merge; the
//      cluster list is compacted; repeat until one cluster is left;
//   4. subtrees of at most MRT_BVH_MAX_LEAF primitives become leaves (the
//      host builder's rule), primitives are laid out in leaf order;
//   5. the binary tree is collapsed to 4-wide nodes level by level, opening
//      the largest-area inner child until a node has four children (the
//      host collapse's rule).
// Every step is deterministic: ties in the neighbour search go to the lower
// index, and new nodes take their indices from prefix sums, not atomics.
#include "mrt_kernels.h"
#include "mrt_build.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cstring>
#include <vector>

namespace mrt {

namespace {

constexpr int PLOC_BLOCK = 256;
// Search radius in positions: on the 1080p mesh workload 16 traces fastest
// (404.6 Msamples/s against 392.6 for 8, 386.8 for 32, 378.3 for 64; wider
// windows merge into fewer but looser nodes).
constexpr int PLOC_RADIUS = 16;

__device__ __forceinline__ float box_area(float4 lo, float4 hi)
{
    float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Ordered-integer image of a float (monotonic for all finite values).
__device__ __forceinline__ uint32_t f2ord(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t u)
{
    uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

// Box of a cylinder / cone / hyperboloid, as quadric_box in bvh_build.c.
__device__ void quad_box(const mrt_quadric &q, float4 &blo, float4 &bhi)
{
    double A[3] = {q.axis[0], q.axis[1], q.axis[2]};
    double P[3] = {q.pos[0], q.pos[1], q.pos[2]};
    double h = q.height, lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        double side = sqrt(fmax(0.0, 1.0 - A[k] * A[k]));
        if (q.kind == MRT_QUADRIC_CONE) {
            double base = P[k] + A[k] * h, rr = (double)q.radius * side;
            lo[k] = fmin(P[k], base - rr);
            hi[k] = fmax(P[k], base + rr);
        } else {
            double rad = q.radius;
            if (q.kind == MRT_QUADRIC_HYPERBOLOID) {
                double hc = 0.5 * h / q.c;
                rad = fmax(q.a, q.b) * sqrt(fmax(0.0, q.shape + hc * hc));
            }
            double e = fabs(A[k]) * 0.5 * h + rad * side;
            lo[k] = P[k] - e;
            hi[k] = P[k] + e;
        }
    }
    double scale = 0.0;
    for (int k = 0; k < 3; k++)
        scale = fmax(scale, fmax(fabs(lo[k]), fabs(hi[k])));
    double pad = 1e-5 * (scale + h + q.radius + q.a + q.b + q.c);
    blo = make_float4(nextafterf((float)(lo[0] - pad), -INFINITY),
                      nextafterf((float)(lo[1] - pad), -INFINITY),
                      nextafterf((float)(lo[2] - pad), -INFINITY), 0.0f);
    bhi = make_float4(nextafterf((float)(hi[0] + pad), INFINITY),
                      nextafterf((float)(hi[1] + pad), INFINITY),
                      nextafterf((float)(hi[2] + pad), INFINITY), 0.0f);
}

struct SceneSrc {
    const float       *tri_pos, *tri_nrm, *sph;
    const uint32_t    *tri_mat, *sph_mat;
    const mrt_quadric *quad;
    uint32_t           n_tri, n_sph, n_quad;
};

__device__ void prim_box(const SceneSrc &S, uint32_t i, float4 &lo, float4 &hi)
{
    if (i < S.n_tri) {
        const float *v = S.tri_pos + (size_t)i * 9;
        lo = make_float4(fminf(v[0], fminf(v[3], v[6])), fminf(v[1], fminf(v[4], v[7])),
                         fminf(v[2], fminf(v[5], v[8])), 0.0f);
        hi = make_float4(fmaxf(v[0], fmaxf(v[3], v[6])), fmaxf(v[1], fmaxf(v[4], v[7])),
                         fmaxf(v[2], fmaxf(v[5], v[8])), 0.0f);
    } else if (i < S.n_tri + S.n_sph) {
        const float *q = S.sph + (size_t)(i - S.n_tri) * 4;
        lo = make_float4(nextafterf(q[0] - q[3], -INFINITY), nextafterf(q[1] - q[3], -INFINITY),
                         nextafterf(q[2] - q[3], -INFINITY), 0.0f);
        hi = make_float4(nextafterf(q[0] + q[3], INFINITY), nextafterf(q[1] + q[3], INFINITY),
                         nextafterf(q[2] + q[3], INFINITY), 0.0f);
    } else {
        quad_box(S.quad[i - S.n_tri - S.n_sph], lo, hi);
    }
}

// Boxes of every primitive and the block-reduced centroid bounds
// (ordered-integer atomics, one per block and axis).
__global__ void __launch_bounds__(256)
k_prim_boxes(SceneSrc S, uint32_t n, float4 *__restrict__ plo,
             float4 *__restrict__ phi, uint32_t *__restrict__ cb)
{
    __shared__ uint32_t red[6][256];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    if (i < n) {
        float4 lo, hi;
        prim_box(S, i, lo, hi);
        plo[i] = lo;
        phi[i] = hi;
        if (!(isfinite(lo.x) && isfinite(lo.y) && isfinite(lo.z) && isfinite(hi.x)
              && isfinite(hi.y) && isfinite(hi.z)))
            atomicOr(&cb[6], 1u);
        float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
        for (int k = 0; k < 3; k++) {
            m[k] = f2ord(c[k]);
            m[3 + k] = f2ord(c[k]);
        }
    }
    for (int k = 0; k < 6; k++)
        red[k][threadIdx.x] = m[k];
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 6; k++) {
                uint32_t a = red[k][threadIdx.x], b = red[k][threadIdx.x + s];
                red[k][threadIdx.x] = k < 3 ? min(a, b) : max(a, b);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) {
            if (k < 3)
                atomicMin(&cb[k], red[k][0]);
            else
                atomicMax(&cb[k], red[k][0]);
        }
}

__device__ __forceinline__ unsigned long long spread21(uint32_t x)
{
    unsigned long long v = x & 0x1FFFFFu;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void __launch_bounds__(256)
k_morton(uint32_t n, const float4 *__restrict__ plo, const float4 *__restrict__ phi,
         const uint32_t *__restrict__ cb, unsigned long long *__restrict__ keys,
         uint32_t *__restrict__ vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float4 lo = plo[i], hi = phi[i];
    float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    unsigned long long key = 0;
    for (int k = 0; k < 3; k++) {
        float a = ord2f(cb[k]), b = ord2f(cb[3 + k]);
        float t = b > a ? (c[k] - a) / (b - a) : 0.0f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        uint32_t q = (uint32_t)(t * 2097151.0f);
        key |= spread21(q) << k;
    }
    keys[i] = key;
    vals[i] = i;
}

// Leaf node i (sorted position) of the binary tree: the box of primitive
// src[i]; left = ~src, right = -1.
__global__ void __launch_bounds__(256)
k_leaves(uint32_t n, const uint32_t *__restrict__ src, const float4 *__restrict__ plo,
         const float4 *__restrict__ phi, float4 *__restrict__ nlo,
         float4 *__restrict__ nhi, uint32_t *__restrict__ clus)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t p = src[i];
    float4 lo = plo[p], hi = phi[p];
    lo.w = __int_as_float(~(int32_t)p);
    hi.w = __int_as_float(-1);
    nlo[i] = lo;
    nhi[i] = hi;
    clus[i] = i;
}

// Nearest neighbour of every cluster in [i-R, i+R] by the area of the union
// box; ties to the lower position.  Boxes of the block's window in LDS.
template <int PLOC_R>
__global__ void __launch_bounds__(PLOC_BLOCK)
k_ploc_nn(uint32_t m, const uint32_t *__restrict__ clus,
          const float4 *__restrict__ nlo, const float4 *__restrict__ nhi,
          uint32_t *__restrict__ nn)
{
    __shared__ float4 slo[PLOC_BLOCK + 2 * PLOC_R], shi[PLOC_BLOCK + 2 * PLOC_R];
    const int base = (int)(blockIdx.x * PLOC_BLOCK) - PLOC_R;
    for (int k = threadIdx.x; k < PLOC_BLOCK + 2 * PLOC_R; k += PLOC_BLOCK) {
        int j = base + k;
        if (j >= 0 && j < (int)m) {
            uint32_t c = clus[j];
            slo[k] = nlo[c];
            shi[k] = nhi[c];
        }
    }
    __syncthreads();
    int i = (int)(blockIdx.x * PLOC_BLOCK + threadIdx.x);
    if (i >= (int)m)
        return;
    int li = i - base;
    float4 alo = slo[li], ahi = shi[li];
    unsigned long long best = ~0ull;
    int j0 = max(0, i - PLOC_R), j1 = min((int)m - 1, i + PLOC_R);
    for (int j = j0; j <= j1; j++) {
        if (j == i)
            continue;
        float4 blo = slo[j - base], bhi = shi[j - base];
        float4 ulo = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.f);
        float4 uhi = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.f);
        unsigned long long key = ((unsigned long long)__float_as_uint(box_area(ulo, uhi)) << 32)
                                 | (uint32_t)j;
        best = key < best ? key : best;
    }
    nn[i] = (uint32_t)best;
}

// Flags of position i: low word 1 when i merges with its mutual neighbour
// (i < nn), high word 1 when i survives (is not the partner of a merge).
__global__ void __launch_bounds__(256)
k_ploc_flags(uint32_t m, const uint32_t *__restrict__ nn,
             unsigned long long *__restrict__ flags)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m)
        return;
    uint32_t j = nn[i];
    bool mutual = m > 1 && nn[j] == i;
    unsigned long long merge = (mutual && i < j) ? 1ull : 0ull;
    unsigned long long keep = (mutual && i > j) ? 0ull : 1ull;
    flags[i] = merge | (keep << 32);
}

__global__ void __launch_bounds__(256)
k_ploc_merge(uint32_t m, uint32_t node0, const uint32_t *__restrict__ clus,
             const uint32_t *__restrict__ nn,
             const unsigned long long *__restrict__ flags,
             const unsigned long long *__restrict__ rank,
             float4 *__restrict__ nlo, float4 *__restrict__ nhi,
             uint32_t *__restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m)
        return;
    unsigned long long f = flags[i], r = rank[i];
    if (!(f >> 32))
        return;
    uint32_t pos = (uint32_t)(r >> 32);
    if (f & 1ull) {
        uint32_t a = clus[i], b = clus[nn[i]];
        uint32_t node = node0 + (uint32_t)(r & 0xFFFFFFFFull);
        float4 alo = nlo[a], ahi = nhi[a], blo = nlo[b], bhi = nhi[b];
        nlo[node] = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z),
                                __int_as_float((int32_t)a));
        nhi[node] = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z),
                                __int_as_float((int32_t)b));
        out[pos] = node;
    } else {
        out[pos] = clus[i];
    }
}

__device__ __forceinline__ bool is_leaf(const float4 *nlo, uint32_t x)
{
    return __float_as_int(nlo[x].w) < 0;
}

// Primitive counts of the inner nodes [a, b), children first (they were
// created in earlier iterations).
__global__ void __launch_bounds__(256)
k_counts(uint32_t a, uint32_t b, const float4 *__restrict__ nlo,
         const float4 *__restrict__ nhi, uint32_t *__restrict__ cnt)
{
    uint32_t x = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= b)
        return;
    uint32_t l = (uint32_t)__float_as_int(nlo[x].w), r = (uint32_t)__float_as_int(nhi[x].w);
    cnt[x] = cnt[l] + cnt[r];
}

// Writes the records of primitive src at leaf-order slot k (the layout of
// mrt_bvh_build: mrt_host.h).
__device__ void write_prim(const SceneSrc &S, uint32_t p, uint32_t k,
                           float4 *__restrict__ prims, float *__restrict__ nrm)
{
    float4 *r = prims + (size_t)k * 3;
    float *o = nrm + (size_t)k * 9;
    if (p < S.n_tri) {
        const float *v = S.tri_pos + (size_t)p * 9;
        float b[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        float c[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        r[0] = make_float4(v[0], v[1], v[2], __uint_as_float(MRT_PRIM_TRIANGLE | p));
        r[1] = make_float4(b[0], b[1], b[2], __uint_as_float(MRT_PRIMKIND_TRIANGLE));
        r[2] = make_float4(c[0], c[1], c[2], __uint_as_float(S.tri_mat[p]));
        if (S.tri_nrm) {
            for (int j = 0; j < 9; j++)
                o[j] = S.tri_nrm[(size_t)p * 9 + j];
        } else {
            double cx = (double)b[1] * c[2] - (double)b[2] * c[1];
            double cy = (double)b[2] * c[0] - (double)b[0] * c[2];
            double cz = (double)b[0] * c[1] - (double)b[1] * c[0];
            double l = sqrt(cx * cx + cy * cy + cz * cz);
            l = l > 0.0 ? 1.0 / l : 0.0;
            for (int j = 0; j < 3; j++) {
                o[3 * j] = (float)(cx * l);
                o[3 * j + 1] = (float)(cy * l);
                o[3 * j + 2] = (float)(cz * l);
            }
        }
    } else if (p < S.n_tri + S.n_sph) {
        uint32_t j = p - S.n_tri;
        const float *q = S.sph + (size_t)j * 4;
        r[0] = make_float4(q[0], q[1], q[2], __uint_as_float(MRT_PRIM_SPHERE | j));
        r[1] = make_float4(q[3], q[3] * q[3], 0.0f, __uint_as_float(MRT_PRIMKIND_SPHERE));
        r[2] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(S.sph_mat[j]));
        for (int t = 0; t < 9; t++)
            o[t] = 0.0f;
    } else {
        uint32_t j = p - S.n_tri - S.n_sph;
        const mrt_quadric &q = S.quad[j];
        bool hyp = q.kind == MRT_QUADRIC_HYPERBOLOID;
        r[0] = make_float4(q.pos[0], q.pos[1], q.pos[2], __uint_as_float(MRT_PRIM_QUADRIC | j));
        r[1] = make_float4(q.axis[0], q.axis[1], q.axis[2],
                           __uint_as_float(MRT_PRIMKIND_QUADRIC + q.kind));
        r[2] = make_float4(hyp ? q.a : q.radius, hyp ? q.b : 0.0f, hyp ? q.c : 0.0f, q.height);
        for (int t = 0; t < 9; t++)
            o[t] = 0.0f;
        o[0] = hyp ? q.shape : 0.0f;
        o[1] = __uint_as_float(q.mat);
        o[2] = __uint_as_float(q.mat_top);
        o[3] = __uint_as_float(q.mat_bot);
    }
}

// Writes the primitives of the subtree at x (at most MRT_BVH_MAX_LEAF, so
// at most 7 nodes) from leaf-order slot off on, left subtree first.
__device__ void gather_leaf(const SceneSrc &S, const float4 *nlo, const float4 *nhi,
                            uint32_t x, uint32_t off, float4 *prims, float *nrm)
{
    uint32_t st[8];
    int sp = 0;
    st[sp++] = x;
    while (sp) {
        uint32_t y = st[--sp];
        int32_t l = __float_as_int(nlo[y].w);
        if (l < 0) {
            write_prim(S, (uint32_t)~l, off++, prims, nrm);
            continue;
        }
        st[sp++] = (uint32_t)__float_as_int(nhi[y].w);
        st[sp++] = (uint32_t)l;
    }
}

// Top-down leaf-order offsets of the inner nodes [a, b) (parents first):
// a node with an offset hands its children theirs; a child of at most
// MRT_BVH_MAX_LEAF primitives becomes a leaf and its records are written.
__global__ void __launch_bounds__(256)
k_offsets(uint32_t a, uint32_t b, SceneSrc S, const float4 *__restrict__ nlo,
          const float4 *__restrict__ nhi, const uint32_t *__restrict__ cnt,
          int32_t *__restrict__ off, float4 *__restrict__ prims,
          float *__restrict__ nrm)
{
    uint32_t x = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= b)
        return;
    int32_t o = off[x];
    if (o < 0 || cnt[x] <= MRT_BVH_MAX_LEAF)
        return;
    uint32_t l = (uint32_t)__float_as_int(nlo[x].w), r = (uint32_t)__float_as_int(nhi[x].w);
    int32_t ol = o, orr = o + (int32_t)cnt[l];
    off[l] = ol;
    off[r] = orr;
    if (cnt[l] <= MRT_BVH_MAX_LEAF)
        gather_leaf(S, nlo, nhi, l, (uint32_t)ol, prims, nrm);
    if (cnt[r] <= MRT_BVH_MAX_LEAF)
        gather_leaf(S, nlo, nhi, r, (uint32_t)orr, prims, nrm);
}

// 4-wide collapse, one level: entry e = (binary node, 4-wide index).  The
// children: the two of the node, then the largest-area inner child opened
// until four (ties to the first); inner children get 4-wide indices from
// the prefix sum of the per-entry counts (pass 2).
struct Kids {
    uint32_t n[4];
    int      nc;
};

__device__ Kids collapse_kids(const float4 *nlo, const float4 *nhi, const uint32_t *cnt,
                              uint32_t x)
{
    Kids K;
    K.n[0] = (uint32_t)__float_as_int(nlo[x].w);
    K.n[1] = (uint32_t)__float_as_int(nhi[x].w);
    K.nc = 2;
    while (K.nc < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int i = 0; i < K.nc; i++) {
            uint32_t y = K.n[i];
            if (cnt[y] > MRT_BVH_MAX_LEAF) {
                float ar = box_area(nlo[y], nhi[y]);
                if (ar > ba) {
                    ba = ar;
                    best = i;
                }
            }
        }
        if (best < 0)
            break;
        uint32_t y = K.n[best];
        for (int i = K.nc; i > best + 1; i--)
            K.n[i] = K.n[i - 1];
        K.n[best] = (uint32_t)__float_as_int(nlo[y].w);
        K.n[best + 1] = (uint32_t)__float_as_int(nhi[y].w);
        K.nc++;
    }
    return K;
}

__global__ void __launch_bounds__(256)
k_c4_count(uint32_t ne, const uint2 *__restrict__ ent, const float4 *__restrict__ nlo,
           const float4 *__restrict__ nhi, const uint32_t *__restrict__ cnt,
           uint32_t *__restrict__ ninner)
{
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    Kids K = collapse_kids(nlo, nhi, cnt, ent[e].x);
    uint32_t c = 0;
    for (int i = 0; i < K.nc; i++)
        c += cnt[K.n[i]] > MRT_BVH_MAX_LEAF ? 1u : 0u;
    ninner[e] = c;
}

__global__ void __launch_bounds__(256)
k_c4_write(uint32_t ne, uint32_t next0, const uint2 *__restrict__ ent,
           const float4 *__restrict__ nlo, const float4 *__restrict__ nhi,
           const uint32_t *__restrict__ cnt, const int32_t *__restrict__ off,
           const uint32_t *__restrict__ rank, float4 *__restrict__ nodes4,
           uint2 *__restrict__ ent_next)
{
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne)
        return;
    Kids K = collapse_kids(nlo, nhi, cnt, ent[e].x);
    float lx[4], hx[4], ly[4], hy[4], lz[4], hz[4];
    int32_t code[4];
    uint32_t r = rank[e];
    for (int i = 0; i < 4; i++) {
        if (i >= K.nc) {
            lx[i] = hx[i] = ly[i] = hy[i] = lz[i] = hz[i] = 1e30f;
            code[i] = ~(int32_t)0;
            continue;
        }
        uint32_t y = K.n[i];
        float4 lo = nlo[y], hi = nhi[y];
        lx[i] = lo.x;
        hx[i] = hi.x;
        ly[i] = lo.y;
        hy[i] = hi.y;
        lz[i] = lo.z;
        hz[i] = hi.z;
        if (cnt[y] <= MRT_BVH_MAX_LEAF) {
            code[i] = ~(int32_t)(((uint32_t)off[y] << 3) | (cnt[y] - 1u));
        } else {
            uint32_t k = next0 + r++;
            code[i] = (int32_t)k;
            ent_next[r - 1] = make_uint2(y, k);
        }
    }
    float4 *n = nodes4 + (size_t)ent[e].y * 8;
    n[0] = make_float4(lx[0], lx[1], lx[2], lx[3]);
    n[1] = make_float4(hx[0], hx[1], hx[2], hx[3]);
    n[2] = make_float4(ly[0], ly[1], ly[2], ly[3]);
    n[3] = make_float4(hy[0], hy[1], hy[2], hy[3]);
    n[4] = make_float4(lz[0], lz[1], lz[2], lz[3]);
    n[5] = make_float4(hz[0], hz[1], hz[2], hz[3]);
    n[6] = make_float4(__int_as_float(code[0]), __int_as_float(code[1]),
                       __int_as_float(code[2]), __int_as_float(code[3]));
    n[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_set_i32(int32_t *p, uint32_t n, int32_t v)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ void k_set_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Exclusive scan of e[0..n) into r, and the total in *total (device).
template <typename T>
hipError_t scan_total(hipStream_t st, const T *e, T *r, uint32_t n, T *total,
                      void *&tmp, size_t &tmp_bytes)
{
    size_t need = 0;
    hipError_t err = rocprim::exclusive_scan(nullptr, need, e, r, T(0), n,
                                             rocprim::plus<T>(), st);
    if (err != hipSuccess)
        return err;
    if (need > tmp_bytes) {
        if (tmp)
            (void)hipFree(tmp);
        tmp = nullptr;
        if ((err = hipMalloc(&tmp, need)) != hipSuccess)
            return err;
        tmp_bytes = need;
    }
    if ((err = rocprim::exclusive_scan(tmp, need, e, r, T(0), n, rocprim::plus<T>(),
                                       st)) != hipSuccess)
        return err;
    // total = last rank + last element
    T last_r, last_e;
    if ((err = hipMemcpyAsync(&last_r, r + n - 1, sizeof(T), hipMemcpyDeviceToHost, st))
            != hipSuccess
        || (err = hipMemcpyAsync(&last_e, e + n - 1, sizeof(T), hipMemcpyDeviceToHost, st))
               != hipSuccess
        || (err = hipStreamSynchronize(st)) != hipSuccess)
        return err;
    *total = last_r + last_e;
    return hipSuccess;
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf()
    {
        if (p)
            (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)); }
};

inline uint32_t grid(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

int device_bvh_build(hipStream_t st, const mrt_scene *s, DeviceBvh *out,
                     const char **why)
{
    memset(out, 0, sizeof(*out));
    *why = "";
    const uint32_t n = s->n_triangles + s->n_spheres + s->n_quadrics;
    if (n < 64) {
        *why = "fewer than 64 primitives";
        return MRT_E_UNSUPPORTED;
    }
    hipError_t e = hipSuccess;
    bool soft = false;       // the tree is unusable: build on the host
#define HE(x)                                                                 \
    do {                                                                      \
        if ((e = (x)) != hipSuccess) {                                        \
            *why = hipGetErrorString(e);                                      \
            goto fail;                                                        \
        }                                                                     \
    } while (0)
    {
        DevBuf<float> tri_pos, tri_nrm, sph;
        DevBuf<uint32_t> tri_mat, sph_mat, cb, src_in, src, clus, clus2, nn, cnt, ninner,
            rank4;
        DevBuf<mrt_quadric> quad;
        DevBuf<float4> plo, phi, nlo, nhi, prims, nodes4;
        DevBuf<unsigned long long> kin, kout, flags, frank, tot64;
        DevBuf<int32_t> off;
        DevBuf<float> nrm;
        DevBuf<uint2> ent, ent2;
        void *tmp = nullptr;
        size_t tmp_bytes = 0;
        std::vector<uint2> ranges;      // inner nodes created per iteration
        uint32_t m = n, node0 = n, root = 0, n4 = 0, depth = 0;
        const uint32_t cbinit[7] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u, 0u};
        uint32_t nonfinite = 0;
        SceneSrc S;

        HE(tri_pos.alloc((size_t)s->n_triangles * 9));
        HE(tri_mat.alloc(s->n_triangles));
        HE(sph.alloc((size_t)s->n_spheres * 4));
        HE(sph_mat.alloc(s->n_spheres));
        HE(quad.alloc(s->n_quadrics));
        if (s->tri_nrm)
            HE(tri_nrm.alloc((size_t)s->n_triangles * 9));
        if (s->n_triangles) {
            HE(hipMemcpyAsync(tri_pos.p, s->tri_pos, (size_t)s->n_triangles * 36,
                              hipMemcpyHostToDevice, st));
            HE(hipMemcpyAsync(tri_mat.p, s->tri_mat, (size_t)s->n_triangles * 4,
                              hipMemcpyHostToDevice, st));
            if (s->tri_nrm)
                HE(hipMemcpyAsync(tri_nrm.p, s->tri_nrm, (size_t)s->n_triangles * 36,
                                  hipMemcpyHostToDevice, st));
        }
        if (s->n_spheres) {
            HE(hipMemcpyAsync(sph.p, s->sph, (size_t)s->n_spheres * 16,
                              hipMemcpyHostToDevice, st));
            HE(hipMemcpyAsync(sph_mat.p, s->sph_mat, (size_t)s->n_spheres * 4,
                              hipMemcpyHostToDevice, st));
        }
        if (s->n_quadrics)
            HE(hipMemcpyAsync(quad.p, s->quadrics, (size_t)s->n_quadrics * sizeof(mrt_quadric),
                              hipMemcpyHostToDevice, st));
        S.tri_pos = tri_pos.p;
        S.tri_nrm = s->tri_nrm ? tri_nrm.p : nullptr;
        S.sph = sph.p;
        S.tri_mat = tri_mat.p;
        S.sph_mat = sph_mat.p;
        S.quad = quad.p;
        S.n_tri = s->n_triangles;
        S.n_sph = s->n_spheres;
        S.n_quad = s->n_quadrics;

        // 1-2: boxes, Morton codes, sort
        HE(plo.alloc(n));
        HE(phi.alloc(n));
        HE(cb.alloc(7));
        HE(hipMemcpyAsync(cb.p, cbinit, sizeof(cbinit), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_prim_boxes, dim3(grid(n, 256)), dim3(256), 0, st, S, n, plo.p,
                           phi.p, cb.p);
        HE(hipGetLastError());
        HE(hipMemcpyAsync(&nonfinite, cb.p + 6, 4, hipMemcpyDeviceToHost, st));
        HE(hipStreamSynchronize(st));
        if (nonfinite) {
            *why = "non-finite primitive";
            soft = true;
            goto fail;
        }
        HE(kin.alloc(n));
        HE(kout.alloc(n));
        HE(src_in.alloc(n));
        HE(src.alloc(n));
        hipLaunchKernelGGL(k_morton, dim3(grid(n, 256)), dim3(256), 0, st, n, plo.p, phi.p,
                           cb.p, kin.p, src_in.p);
        HE(hipGetLastError());
        {
            size_t need = 0;
            HE(rocprim::radix_sort_pairs(nullptr, need, kin.p, kout.p, src_in.p, src.p, n, 0,
                                         63, st));
            HE(hipMalloc(&tmp, need ? need : 1));
            tmp_bytes = need ? need : 1;
            HE(rocprim::radix_sort_pairs(tmp, need, kin.p, kout.p, src_in.p, src.p, n, 0, 63,
                                         st));
        }

        // 3: PLOC
        HE(nlo.alloc(2 * (size_t)n));
        HE(nhi.alloc(2 * (size_t)n));
        HE(clus.alloc(n));
        HE(clus2.alloc(n));
        HE(nn.alloc(n));
        HE(flags.alloc(n));
        HE(frank.alloc(n));
        hipLaunchKernelGGL(k_leaves, dim3(grid(n, 256)), dim3(256), 0, st, n, src.p, plo.p,
                           phi.p, nlo.p, nhi.p, clus.p);
        HE(hipGetLastError());
        while (m > 1) {
            if (ranges.size() >= 4096) {       // pathological input
                *why = "clustering needs too many passes";
                soft = true;
                goto fail;
            }
            hipLaunchKernelGGL(k_ploc_nn<PLOC_RADIUS>, dim3(grid(m, PLOC_BLOCK)),
                               dim3(PLOC_BLOCK), 0, st, m, clus.p, nlo.p, nhi.p, nn.p);
            hipLaunchKernelGGL(k_ploc_flags, dim3(grid(m, 256)), dim3(256), 0, st, m, nn.p,
                               flags.p);
            HE(hipGetLastError());
            unsigned long long t = 0;
            HE(scan_total<unsigned long long>(st, flags.p, frank.p, m, &t, tmp, tmp_bytes));
            uint32_t merges = (uint32_t)(t & 0xFFFFFFFFull), keep = (uint32_t)(t >> 32);
            if (merges == 0 || keep + merges != m) {
                *why = "clustering made no progress";
                soft = true;
                goto fail;
            }
            hipLaunchKernelGGL(k_ploc_merge, dim3(grid(m, 256)), dim3(256), 0, st, m, node0,
                               clus.p, nn.p, flags.p, frank.p, nlo.p, nhi.p, clus2.p);
            HE(hipGetLastError());
            ranges.push_back(make_uint2(node0, node0 + merges));
            node0 += merges;
            m = keep;
            std::swap(clus.p, clus2.p);
        }
        HE(hipMemcpyAsync(&root, clus.p, 4, hipMemcpyDeviceToHost, st));
        HE(hipStreamSynchronize(st));
        if (root != 2 * n - 2 || node0 != 2 * n - 1) {
            *why = "clustering did not end in one tree";
            soft = true;
            goto fail;
        }

        // 4: counts (children first), leaf offsets and records (parents first)
        HE(cnt.alloc(2 * (size_t)n));
        hipLaunchKernelGGL(k_set_u32, dim3(grid(n, 256)), dim3(256), 0, st, cnt.p, n, 1u);
        for (const uint2 &r : ranges)
            hipLaunchKernelGGL(k_counts, dim3(grid(r.y - r.x, 256)), dim3(256), 0, st, r.x,
                               r.y, nlo.p, nhi.p, cnt.p);
        HE(hipGetLastError());
        HE(off.alloc(2 * (size_t)n));
        HE(prims.alloc((size_t)n * 3));
        HE(nrm.alloc((size_t)n * 9));
        hipLaunchKernelGGL(k_set_i32, dim3(grid(2 * (uint64_t)n, 256)), dim3(256), 0, st,
                           off.p, 2 * n, -1);
        HE(hipMemsetAsync(off.p + root, 0, 4, st));
        for (size_t k = ranges.size(); k-- > 0;)
            hipLaunchKernelGGL(k_offsets, dim3(grid(ranges[k].y - ranges[k].x, 256)),
                               dim3(256), 0, st, ranges[k].x, ranges[k].y, S, nlo.p, nhi.p,
                               cnt.p, off.p, prims.p, nrm.p);
        HE(hipGetLastError());

        // 5: 4-wide collapse, level by level
        HE(nodes4.alloc((size_t)n * 8));
        HE(ent.alloc(n));
        HE(ent2.alloc(n));
        HE(ninner.alloc(n));
        HE(rank4.alloc(n));
        {
            uint2 e0 = make_uint2(root, 0u);
            HE(hipMemcpyAsync(ent.p, &e0, sizeof(e0), hipMemcpyHostToDevice, st));
            uint32_t ne = 1;
            n4 = 1;
            while (ne) {
                depth++;
                hipLaunchKernelGGL(k_c4_count, dim3(grid(ne, 256)), dim3(256), 0, st, ne,
                                   ent.p, nlo.p, nhi.p, cnt.p, ninner.p);
                HE(hipGetLastError());
                uint32_t tot = 0;
                HE(scan_total<uint32_t>(st, ninner.p, rank4.p, ne, &tot, tmp, tmp_bytes));
                if ((uint64_t)n4 + tot > n) {
                    *why = "4-wide node count out of range";
                    soft = true;
                    goto fail;
                }
                hipLaunchKernelGGL(k_c4_write, dim3(grid(ne, 256)), dim3(256), 0, st, ne, n4,
                                   ent.p, nlo.p, nhi.p, cnt.p, off.p, rank4.p, nodes4.p,
                                   ent2.p);
                HE(hipGetLastError());
                n4 += tot;
                ne = tot;
                std::swap(ent.p, ent2.p);
            }
        }
        HE(hipStreamSynchronize(st));
        if (depth > MRT_BVH4_MAX_DEPTH) {
            *why = "4-wide depth above the traversal stack's bound";
            soft = true;
            goto fail;
        }
        // hand the arrays over (nodes copied to an exact-size buffer)
        HE(hipMalloc((void **)&out->nodes, (size_t)n4 * 128));
        HE(hipMemcpyAsync(out->nodes, nodes4.p, (size_t)n4 * 128, hipMemcpyDeviceToDevice, st));
        HE(hipStreamSynchronize(st));
        out->prims = prims.p;
        prims.p = nullptr;
        out->nrm = nrm.p;
        nrm.p = nullptr;
        out->n_nodes = n4;
        out->n_prims = n;
        out->depth = depth;
        out->iterations = (uint32_t)ranges.size();
        if (tmp)
            (void)hipFree(tmp);
        return MRT_OK;
    fail:
        if (tmp)
            (void)hipFree(tmp);
        if (out->nodes)
            (void)hipFree(out->nodes);
        memset(out, 0, sizeof(*out));
        if (soft)
            return MRT_E_UNSUPPORTED;
        return e == hipErrorOutOfMemory ? MRT_E_NOMEM : MRT_E_HIP;
    }
#undef HE
}

}  // namespace mrt
