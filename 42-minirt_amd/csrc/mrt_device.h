// Device-side building blocks of the wavefront path tracer (gfx950):
// counter RNG, primitive tests and the ordered BVH traversal.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mrt.h"
#include "mrt_host.h"

namespace mrt {

constexpr float IT_MIN_F = 0.001f;    // IT_MIN, includes/minirt.h:44
constexpr float IT_MAX_F = 1000.0f;   // IT_MAX, includes/minirt.h:43
constexpr float BOX_WIDEN = 1.0000004f;  // 1 + 2 gamma_3 (conservative slab)
constexpr uint32_t HIT_NONE = 0xFFFFFFFFu;
constexpr uint32_t HIT_PLANE = 0x80000000u;   // hit code of plane k: flag | k

// ---- counter-based RNG -----------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter = {pixel, sample,
// bounce << 20 | block, stream}, key = the 64-bit scene seed.
struct U4 {
    float x, y, z, w;
};

__device__ __forceinline__ U4 philox_u4(uint32_t c0, uint32_t c1, uint32_t c2,
                                        uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    const float s = 5.9604644775390625e-08f;   // 2^-24
    return U4{(float)(c0 >> 8) * s, (float)(c1 >> 8) * s,
              (float)(c2 >> 8) * s, (float)(c3 >> 8) * s};
}

__device__ __forceinline__ U4 rng4(uint64_t seed, uint32_t pixel,
                                   uint32_t sample, uint32_t bounce,
                                   uint32_t block)
{
    return philox_u4(pixel, sample, (bounce << 20) | block, 0u,
                     (uint32_t)seed, (uint32_t)(seed >> 32));
}

// The counter of one path vertex: stream 0 camera paths, 1 photons.
struct RngKey {
    uint64_t seed;
    uint32_t pixel, sample, bounce, stream;
    __device__ __forceinline__ U4 operator()(uint32_t block) const
    {
        return philox_u4(pixel, sample, (bounce << 20) | block, stream,
                         (uint32_t)seed, (uint32_t)(seed >> 32));
    }
};

// ---- small vector helpers (explicit evaluation order) ---------------------
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 had(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// vec3_unit: v / |v| per component, zero stays zero
__device__ __forceinline__ V3 unit(V3 a)
{
    float l = sqrtf(dot(a, a));
    if (l > 0.0f)
        return v3(a.x / l, a.y / l, a.z / l);
    return v3(0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ V3 xyz(float4 f) { return v3(f.x, f.y, f.z); }

// ---- primitive tests ------------------------------------------------------
// Moller-Trumbore as hit_triangle (src/math/triangle/hit_triangle.c:68-96);
// B = v1 - v0, C = v2 - v0 stored in the record.
__device__ __forceinline__ bool isect_tri(V3 o, V3 d, float4 A, float4 B,
                                          float4 C, bool one_sided, float &t,
                                          float &bu, float &bv)
{
    V3 e1 = xyz(B), e2 = xyz(C);
    V3 h = cross(d, e2);
    float a = dot(e1, h);
    if (one_sided ? (a < 1e-6f) : (fabsf(a) < 1e-6f))
        return false;
    float f = 1.0f / a;
    V3 s = sub(o, xyz(A));
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f)
        return false;
    V3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f)
        return false;
    float tt = f * dot(e2, q);
    if (tt < IT_MIN_F || tt > IT_MAX_F)
        return false;
    t = tt;
    bu = u;
    bv = v;
    return true;
}

// hit_sphere's roots (src/math/sphere/hit_sphere.c:31-47, valid_t.c:16-26)
// in the cancellation-free form (Haines et al., Ray Tracing Gems ch. 7).
__device__ __forceinline__ bool isect_sph(V3 o, V3 d, float4 A, float4 B,
                                          float &t)
{
    V3 oc = sub(o, xyz(A));
    float r2 = B.x * B.x;
    float bq = -dot(oc, d);
    V3 l = add(oc, mul(d, bq));
    float disc = r2 - dot(l, l);
    if (disc < 0.0f)
        return false;
    float sq = sqrtf(disc);
    float qv = bq + copysignf(sq, bq);
    float cc = dot(oc, oc) - r2;
    float ta = qv != 0.0f ? cc / qv : 0.0f;
    float tb = qv;
    float t0 = fminf(ta, tb), t1 = fmaxf(ta, tb);
    float tt = t0;
    if (tt < IT_MIN_F || tt > IT_MAX_F) {
        tt = t1;
        if (tt < IT_MIN_F || tt > IT_MAX_F)
            return false;
    }
    t = tt;
    return true;
}

// hit_plane (src/math/plane/hit_plane.c:43-67); P = point | mat, N = normal.
__device__ __forceinline__ bool isect_pln(V3 o, V3 d, float4 P, float4 N, float &t)
{
    float denom = N.x * d.x + N.y * d.y + N.z * d.z;
    float tt = ((P.x - o.x) * N.x + (P.y - o.y) * N.y + (P.z - o.z) * N.z) / denom;
    if (fabsf(denom) < 1e-8f)
        return false;
    if (tt < IT_MIN_F || tt > IT_MAX_F)
        return false;
    t = tt;
    return true;
}

// ---- cylinders, cones, hyperboloids ---------------------------------------
// Record (mrt_host.h): A = pos | src id, B = axis | kind, C = (radius or a,
// b, c, height); the shading record holds shape, mat, mat_top, mat_bot.
// Hits report the part in bu: 0 body, 1 top cap, 2 bottom cap.
enum : uint32_t { QPART_BODY = 0, QPART_TOP = 1, QPART_BOTTOM = 2 };

// Roots of a t^2 + b t + c labelled as solve_quadratic labels them
// (src/math/quadratique/quadratique.c:17-30): t0 = (-b - sqrt(disc)) / 2a,
// t1 = (-b + sqrt(disc)) / 2a, also when a < 0, evaluated without
// cancellation (q = -(b + sign(b) sqrt(disc)) / 2; roots q / a and c / q).
// a == 0 (a ray along a cylinder's axis or a cone's side) has no root here;
// the reference divides by zero there.
__device__ __forceinline__ bool quad_roots(float a, float b, float c,
                                           float &t0, float &t1)
{
    float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f || a == 0.0f)
        return false;
    float sq = sqrtf(disc);
    float q = -0.5f * (b + copysignf(sq, b));
    if (q == 0.0f) {
        t0 = t1 = 0.0f;
        return true;
    }
    float r1 = q / a, r2 = c / q;
    t0 = b >= 0.0f ? r1 : r2;
    t1 = b >= 0.0f ? r2 : r1;
    return true;
}

// valid_t (src/math/valid_t.c:16-26): t0 if in [IT_MIN, IT_MAX], else t1.
__device__ __forceinline__ bool valid_root(float t0, float t1, float &t)
{
    t = t0;
    if (t < IT_MIN_F || t > IT_MAX_F) {
        t = t1;
        if (t < IT_MIN_F || t > IT_MAX_F)
            return false;
    }
    return true;
}

// A cap disk of centre pc and normal n (cylinder_cap_bottom / _top,
// hit_cylinder.c:39-100; cone_cap, hit_cone.c:81-108).
__device__ __forceinline__ bool cap_disk(V3 o, V3 d, V3 pc, V3 n, float rad,
                                         float &t)
{
    float denom = dot(d, n);
    if (fabsf(denom) < 1e-6f)
        return false;
    float tt = dot(sub(pc, o), n) / denom;
    if (tt < IT_MIN_F || tt > IT_MAX_F)
        return false;
    V3 e = sub(add(o, mul(d, tt)), pc);
    if (sqrtf(dot(e, e)) > rad)
        return false;
    t = tt;
    return true;
}

// hit_cylinder (hit_cylinder.c:107-134): the body root valid_t picks, kept
// when its projection on the axis from the bottom centre lies in [0, h]
// (the other root is not tried), then each cap when nearer.
__device__ __forceinline__ bool isect_cyl(V3 o, V3 d, float4 A, float4 B,
                                          float4 C, float &t, float &part)
{
    V3 P = xyz(A), ax = xyz(B);
    float rad = C.x, h = C.w;
    V3 oc = sub(o, P);
    float dd = dot(d, ax), oo = dot(oc, ax);
    float qa = dot(d, d) - dd * dd;
    float qb = 2.0f * (dot(d, oc) - dd * oo);
    float qc = dot(oc, oc) - oo * oo - rad * rad;
    bool hit = false;
    float t0, t1, th;
    V3 base = sub(P, mul(ax, h * 0.5f));
    if (quad_roots(qa, qb, qc, t0, t1) && valid_root(t0, t1, th)) {
        float proj = dot(sub(add(o, mul(d, th)), base), ax);
        if (!(proj < 0.0f || proj > h)) {
            t = th;
            part = (float)QPART_BODY;
            hit = true;
        }
    }
    float tc;
    if (cap_disk(o, d, base, ax, rad, tc) && (!hit || tc < t)) {
        t = tc;
        part = (float)QPART_BOTTOM;
        hit = true;
    }
    if (cap_disk(o, d, add(P, mul(ax, h * 0.5f)), ax, rad, tc) && (!hit || tc < t)) {
        t = tc;
        part = (float)QPART_TOP;
        hit = true;
    }
    return hit;
}

// hit_cone (hit_cone.c:111-129) with init_cone_quadratic (quadratique.c:
// 55-66), k = (r / h)^2: the body root valid_t picks when its projection
// from the apex lies in [0, h]; only otherwise the base cap.
__device__ __forceinline__ bool isect_cone(V3 o, V3 d, float4 A, float4 B,
                                           float4 C, float &t, float &part)
{
    V3 P = xyz(A), ax = xyz(B);
    float rad = C.x, h = C.w;
    float kr = rad / h, k1 = 1.0f + kr * kr;
    V3 oc = sub(o, P);
    float dd = dot(d, ax), oo = dot(oc, ax);
    float qa = dot(d, d) - k1 * dd * dd;
    float qb = 2.0f * (dot(d, oc) - k1 * dd * oo);
    float qc = dot(oc, oc) - k1 * oo * oo;
    float t0, t1, th;
    if (!quad_roots(qa, qb, qc, t0, t1) || !valid_root(t0, t1, th))
        return false;
    float proj = dot(sub(add(o, mul(d, th)), P), ax);
    if (!(proj < 0.0f || proj > h)) {
        t = th;
        part = (float)QPART_BODY;
        return true;
    }
    float tc;
    if (cap_disk(o, d, add(P, mul(ax, h)), ax, rad, tc)) {
        t = tc;
        part = (float)QPART_TOP;
        return true;
    }
    return false;
}

// The hyperboloid's quadric frame (get_basis_vectors,
// quadratique_hyperboloid.c:15-26): z = axis, x = unit(cross(t, z)) with
// t = +Y when |z.x| > 0.9 else +X, y = cross(z, x).
__device__ __forceinline__ void hyp_frame(V3 z, V3 &x, V3 &y)
{
    V3 tmp = fabsf(z.x) > 0.9f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
    x = unit(cross(tmp, z));
    y = cross(z, x);
}

// hit_hyperboloid (hit_hyperboloid.c:79-104): roots in (IT_MIN, IT_MAX),
// t0 first, then t1 when nearer, each kept when |proj on the axis| <= h/2.
__device__ __forceinline__ bool isect_hyp(V3 o, V3 d, float4 A, float4 B,
                                          float4 C, float shape, float &t)
{
    V3 P = xyz(A), z = xyz(B), x, y;
    hyp_frame(z, x, y);
    V3 oc = sub(o, P);
    V3 dl = v3(dot(d, x), dot(d, y), dot(d, z));
    V3 ol = v3(dot(oc, x), dot(oc, y), dot(oc, z));
    float aa = C.x * C.x, bb = C.y * C.y, cc = C.z * C.z;
    float qa = (dl.x * dl.x) / aa + (dl.y * dl.y) / bb - (dl.z * dl.z) / cc;
    float qb = 2.0f * ((dl.x * ol.x) / aa + (dl.y * ol.y) / bb - (dl.z * ol.z) / cc);
    float qc = (ol.x * ol.x) / aa + (ol.y * ol.y) / bb - (ol.z * ol.z) / cc - shape;
    float t0, t1;
    if (!quad_roots(qa, qb, qc, t0, t1))
        return false;
    float hh = C.w * 0.5f, tmax = IT_MAX_F;
    bool hit = false;
    if (t0 > IT_MIN_F && t0 < tmax
        && fabsf(dot(sub(add(o, mul(d, t0)), P), z)) <= hh) {
        t = t0;
        tmax = t0;
        hit = true;
    }
    if (t1 > IT_MIN_F && t1 < tmax
        && fabsf(dot(sub(add(o, mul(d, t1)), P), z)) <= hh) {
        t = t1;
        hit = true;
    }
    return hit;
}

__device__ __forceinline__ bool isect_quadric(V3 o, V3 d, float4 A, float4 B,
                                              float4 C, const float *aux,
                                              float &t, float &part)
{
    uint32_t kind = __float_as_uint(B.w) - MRT_PRIMKIND_QUADRIC;
    part = (float)QPART_BODY;
    if (kind == MRT_QUADRIC_CYLINDER)
        return isect_cyl(o, d, A, B, C, t, part);
    if (kind == MRT_QUADRIC_CONE)
        return isect_cone(o, d, A, B, C, t, part);
    return isect_hyp(o, d, A, B, C, aux[0], t);
}

// Normal facing the ray, side and material at a quadric hit p:
// handle_cylinder_hit (hit_cylinder.c:15-37; hit_cylinder sets front_face
// again from the already flipped body normal, so the body is always the
// front side, :128-131), handle_cone_hit (hit_cone.c:58-79), the caps, and
// compute_normal (hit_hyperboloid.c:40-56, whose frame differs from the
// quadric's: u = unit(cross(t, axis)), t = +Y when |axis.y| < 0.99 else
// +X); materials per part as update_hit_cylinder / update_hit_cone
// (hit_obj.c:17-53).
__device__ __forceinline__ void quadric_surface(float4 A, float4 B, float4 C,
                                                const float *aux, V3 d, V3 p,
                                                uint32_t part, V3 &nrm,
                                                bool &front, uint32_t &mat)
{
    uint32_t kind = __float_as_uint(B.w) - MRT_PRIMKIND_QUADRIC;
    V3 P = xyz(A), ax = xyz(B);
    float h = C.w;
    mat = __float_as_uint(aux[part == QPART_TOP ? 2 : part == QPART_BOTTOM ? 3 : 1]);
    if (part != QPART_BODY) {
        V3 n = part == QPART_TOP ? ax : neg(ax);
        front = dot(d, n) < 0.0f;
        nrm = front ? n : neg(n);
        return;
    }
    V3 ow;
    if (kind == MRT_QUADRIC_CYLINDER) {
        V3 rel = sub(p, sub(P, mul(ax, h * 0.5f)));
        ow = unit(sub(rel, mul(ax, dot(rel, ax))));
    } else if (kind == MRT_QUADRIC_CONE) {
        V3 rel = sub(p, P);
        float kr = C.x / h;
        ow = unit(sub(rel, mul(ax, (1.0f + kr * kr) * dot(rel, ax))));
    } else {
        V3 tmp = fabsf(ax.y) < 0.99f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
        V3 u = unit(cross(tmp, ax)), v = cross(ax, u), rel = sub(p, P);
        ow = unit(add(add(mul(u, 2.0f * dot(rel, u) / (C.x * C.x)),
                          mul(v, 2.0f * dot(rel, v) / (C.y * C.y))),
                      mul(ax, -2.0f * dot(rel, ax) / (C.z * C.z))));
    }
    bool f = dot(d, ow) < 0.0f;
    nrm = f ? ow : neg(ow);
    front = kind == MRT_QUADRIC_CYLINDER ? true : f;
}

// ---- textures ----------------------------------------------------------------
// Per texture 3 float4: (type, width, height, texel0 as uint32 bits),
// (scale, c1 rgb 0..255), (c2 rgb, 0); texels RGBA8; per material 3 uint4:
// (color, metallic, roughness, transmission), (ao, emission_strength,
// emission_color, normal) texture indices, (normal_intensity bits, scale
// bits, 0, 0); tri_uv 6 floats per source triangle.
struct TexDev {
    const float4   *tex;
    const uint32_t *texels;
    const uint4    *mtex;
    const float    *tri_uv;
};

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// modf's fractional part (keeps the sign)
__device__ __forceinline__ float frac_part(float x) { return x - truncf(x); }

// The reference's in-plane "right" vector of an axis: unit(cross(+Y, axis)),
// or unit(cross(+X, axis)) when |axis.y| >= 0.99.
__device__ __forceinline__ V3 axis_right(V3 ax)
{
    return unit(cross(fabsf(ax.y) < 0.99f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f), ax));
}

// Nearest texel of an image ((int)(u (w-1)), (int)(v (h-1))), kept inside
// the image (the reference reads outside it for (u, v) outside [0, 1]).
__device__ __forceinline__ uint32_t tex_texel(const TexDev &T, float4 T0, float u,
                                              float v)
{
    uint32_t w = __float_as_uint(T0.y), h = __float_as_uint(T0.z);
    int x = (int)(u * (float)(w - 1)), y = (int)(v * (float)(h - 1));
    x = min(max(x, 0), (int)w - 1);
    y = min(max(y, 0), (int)h - 1);
    return T.texels[__float_as_uint(T0.w) + (uint32_t)y * w + (uint32_t)x];
}

// get_tex_color (get_tex_color.c:27-58) with get_solid_texture
// (get_solid_texture.c:30-43): colour in 0..1.
__device__ __forceinline__ V3 tex_rgb(const TexDev &T, uint32_t tid, float u,
                                      float v, V3 p)
{
    float4 T0 = T.tex[3 * tid], T1 = T.tex[3 * tid + 1], T2 = T.tex[3 * tid + 2];
    uint32_t type = __float_as_uint(T0.x);
    if (type == MRT_TEX_IMAGE) {
        uint32_t c = tex_texel(T, T0, u, v);
        return v3((float)(c & 255u) / 255.0f, (float)((c >> 8) & 255u) / 255.0f,
                  (float)((c >> 16) & 255u) / 255.0f);
    }
    if (type == MRT_TEX_BLACK)
        return v3(0.f, 0.f, 0.f);
    int parity;
    float sc = T1.x;
    if (type == MRT_TEX_CHECKER_LOCAL) {
        parity = (int)floorf(u * sc) + (int)floorf(v * sc);
    } else {
        V3 q = mul(p, sc);
        parity = (int)floorf(q.x) + (int)floorf(q.y + 0.5f) + (int)floorf(q.z);
    }
    return (parity & 1) ? v3(T2.x / 255.0f, T2.y / 255.0f, T2.z / 255.0f)
                        : v3(T1.y / 255.0f, T1.z / 255.0f, T1.w / 255.0f);
}

// get_tex_transparency (get_tex_color.c:60-70): 1 - alpha of an image.
__device__ __forceinline__ float tex_transparency(const TexDev &T, uint32_t tid,
                                                  float u, float v)
{
    float4 T0 = T.tex[3 * tid];
    if (__float_as_uint(T0.x) != MRT_TEX_IMAGE)
        return 0.0f;
    return 1.0f - (float)(tex_texel(T, T0, u, v) >> 24) / 255.0f;
}

// (u, v) at p on a BVH primitive (record A, B, C, shading record aux),
// scaled by the material scale sc of the part hit: get_sphere_uv
// (hit_sphere.c:15-29), hit_triangle's interpolation (hit_triangle.c:
// 91-92), get_cylinder_lateral_uv / get_cylinder_cap_uv (cylinder_uv.c),
// get_cone_lateral_uv / get_cone_cap_uv (hit_cone.c:14-56, whose cap
// centre is taken at height h / 2) and get_hyperboloid_uv
// (hit_hyperboloid.c:15-38).
__device__ __forceinline__ void prim_uv(const TexDev &T, float4 A, float4 B,
                                        float4 C, V3 p, float bu, float bv,
                                        float sc, float &u, float &v)
{
    const float PI = 3.14159265358979323846f;
    uint32_t kind = __float_as_uint(B.w);
    if (kind == MRT_PRIMKIND_TRIANGLE) {
        u = v = 0.0f;
        if (T.tri_uv) {
            const float *q = T.tri_uv + (size_t)(__float_as_uint(A.w) & ~MRT_PRIM_KIND_MASK) * 6;
            float w = 1.0f - bu - bv;
            u = w * q[0] + bu * q[2] + bv * q[4];
            v = w * q[1] + bu * q[3] + bv * q[5];
        }
        return;
    }
    if (kind == MRT_PRIMKIND_SPHERE) {
        float r = B.x;
        u = clamp01((atan2f((-p.z + A.z) / r, (p.x - A.x) / r) + PI) / (2.0f * PI));
        v = clamp01(acosf((-p.y + A.y) / r) / PI);
    } else {
        uint32_t q = kind - MRT_PRIMKIND_QUADRIC, part = (uint32_t)bu;
        V3 P = xyz(A), ax = xyz(B), right = axis_right(ax);
        float h = C.w;
        if (part == QPART_BODY) {
            V3 rel = sub(p, P);
            u = clamp01((atan2f(dot(rel, cross(ax, right)), dot(rel, right)) + PI)
                        / (2.0f * PI));
            float z = dot(sub(p, P), ax);
            v = q == MRT_QUADRIC_CYLINDER ? clamp01(z / h + 0.5f)
                : q == MRT_QUADRIC_CONE   ? clamp01(z / h)
                                          : clamp01(z / (h * 0.5f) * 0.5f + 0.5f);
        } else {
            float off = q == MRT_QUADRIC_CONE ? h * 0.5f
                        : part == QPART_TOP  ? h * 0.5f : -(h * 0.5f);
            V3 rel = sub(p, add(P, mul(ax, off)));
            float dia = 2.0f * C.x;
            u = clamp01(dot(rel, right) / dia + 0.5f);
            v = clamp01(dot(rel, cross(ax, right)) / dia + 0.5f);
        }
    }
    if (sc != 1.0f) {
        u = frac_part(u * sc);
        v = frac_part(v * sc);
    }
}

// get_uv_plane (hit_plane.c:18-41): coordinates along the plane's tangent
// frame, times 0.1 * scale, wrapped by v - floor(v).
__device__ __forceinline__ void plane_uv(float4 P, float4 N, V3 p, float sc,
                                         float &u, float &v)
{
    V3 n = xyz(N);
    V3 tg = unit(cross(fabsf(n.y) < 0.99f ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f), n));
    V3 local = sub(p, xyz(P));
    u = dot(local, tg) * (sc * 0.1f);
    v = dot(local, cross(n, tg)) * (sc * 0.1f);
    u = u - floorf(u);
    v = v - floorf(v);
}

// Counter of the alpha test's draw: the primitive, its (u, v) on a 2^-12
// grid and the ray direction on a 2^-7 grid, so a decision depends on where
// and from where a primitive is hit, not on the traversal order, and a
// rounding difference in the ray changes it only at a grid line.
__device__ __forceinline__ U4 alpha_draw(float u, float v, V3 d, uint32_t src)
{
    uint32_t qd = (uint32_t)((d.x + 1.0f) * 127.5f)
                  | ((uint32_t)((d.y + 1.0f) * 127.5f) << 8)
                  | ((uint32_t)((d.z + 1.0f) * 127.5f) << 16);
    return philox_u4((uint32_t)(int32_t)floorf(u * 4096.0f),
                     (uint32_t)(int32_t)floorf(v * 4096.0f), qd, src,
                     0xA17A7E57u, 0x5EEDA1FAu);
}

// The stochastic alpha test of hit_register_bvh (hit_register.c:57-69): a
// hit on a texel of transparency tr is skipped when tr == 1, or with
// probability tr (alpha_draw).
__device__ __forceinline__ bool alpha_skips(const TexDev &T, const float *aux,
                                            float4 A, float4 B, float4 C, V3 o,
                                            V3 d, float t, float bu, float bv)
{
    uint32_t kind = __float_as_uint(B.w), mat;
    if (kind >= MRT_PRIMKIND_QUADRIC) {
        uint32_t part = (uint32_t)bu;
        mat = __float_as_uint(aux[part == QPART_TOP ? 2 : part == QPART_BOTTOM ? 3 : 1]);
    } else {
        mat = __float_as_uint(C.w);
    }
    uint4 M0 = T.mtex[3 * mat];
    if (M0.x == MRT_NO_TEX)
        return false;
    if (__float_as_uint(T.tex[3 * M0.x].x) != MRT_TEX_IMAGE)
        return false;
    float sc = __uint_as_float(T.mtex[3 * mat + 2].y), u, v;
    prim_uv(T, A, B, C, add(o, mul(d, t)), bu, bv, sc, u, v);
    float tr = tex_transparency(T, M0.x, u, v);
    if (tr == 1.0f)
        return true;
    if (tr == 0.0f)
        return false;
    return alpha_draw(u, v, d, __float_as_uint(A.w)).x < tr;
}

__device__ __forceinline__ float inv_dir(float d)
{
    return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d));
}

__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx,
                                     float hy, float hz, V3 o, V3 inv,
                                     float tclose, float &tnear)
{
    float tx1 = (lx - o.x) * inv.x, tx2 = (hx - o.x) * inv.x;
    float ty1 = (ly - o.y) * inv.y, ty2 = (hy - o.y) * inv.y;
    float tz1 = (lz - o.z) * inv.z, tz2 = (hz - o.z) * inv.z;
    float tmin = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)),
                       fmaxf(fminf(tz1, tz2), 0.0f));
    float tmax = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)),
                       fminf(fmaxf(tz1, tz2), tclose));
    tnear = tmin;
    return tmin <= tmax * BOX_WIDEN;
}

struct SceneDev {
    const float4 *nodes;      // 8 per node
    const float4 *prims;      // 3 per primitive (leaf order)
    const float  *aux;        // 9 per primitive: shading record
    const float4 *planes;     // 2 per plane: point | mat, normal | 0
    uint32_t      n_nodes;    // 0: no BVH
    uint32_t      n_planes;
    uint32_t      one_sided;
    uint32_t      root;       // root node (scenes of a batch share one array)
    uint32_t      quad;       // nonzero: the primitives include quadrics
    uint32_t      alpha;      // nonzero: some material has an image colour
                              // texture with transparent texels
    TexDev        tx;
};

constexpr int STACK4_MAX = 3 * (int)MRT_BVH4_MAX_DEPTH + 1;

// Compare-exchange of (entry distance, child code) pairs.
// Compare-exchange of (key, code) pairs: the smaller key first, ties kept
// in place.  Written as selects, so a divergent wave runs four v_cndmask
// instead of a branch around the swap.
__device__ __forceinline__ void cx(float &ka, int32_t &ca, float &kb, int32_t &cb)
{
    const bool s = ka > kb;
    const float lo = s ? kb : ka, hi = s ? ka : kb;
    const int32_t cl = s ? cb : ca, ch = s ? ca : cb;
    ka = lo;
    kb = hi;
    ca = cl;
    cb = ch;
}

// Ordered traversal of the 4-wide BVH: one 128-byte node gives the four
// child boxes; the hits are sorted by entry distance (network (0,1) (2,3)
// (0,2) (1,3) (1,2)), the nearest is entered and the others are pushed
// farthest first.  The first SS stack entries live in LDS (lds_code,
// stride `stride`), the next STACK4_MAX - SS at ovf[k * ovf_stride]: a
// private array (stride 1) or this lane's column of a device buffer, which
// keeps the kernel free of scratch memory.  ANY: stop at the first hit
// below tlimit.
// EXT = false drops the quadric and alpha tests from the leaf loop (scenes
// without them keep the traversal's register count and code size).
// TOP: nodes 0 .. ntop-1 (the top levels, laid out breadth-first at the
// front of the node array) are read from `top`, the block's copy in LDS
// (7 float4 per node), instead of from global memory.
template <bool ANY, bool COUNT, int SS, bool EXT = true, bool TOP = false>
__device__ __forceinline__ bool traverse(const SceneDev &S, V3 o, V3 d,
                                         float tlimit, uint32_t *lds_code,
                                         int stride, uint32_t *ovf,
                                         uint32_t ovf_stride,
                                         float &t_out, uint32_t &hit_slot,
                                         float &bu_out, float &bv_out,
                                         uint32_t &nodes, uint32_t &prims,
                                         const float4 *top = nullptr,
                                         uint32_t ntop = 0)
{
    if (S.n_nodes == 0)
        return false;
    // end of the walk: not a node index (>= 0) nor a leaf code (~(first << 3
    // | count - 1) with count <= MRT_BVH_MAX_LEAF has low bits 4..7)
    const int32_t DONE = INT32_MIN;
    const float INF = __int_as_float(0x7f800000);
    V3 inv = v3(inv_dir(d.x), inv_dir(d.y), inv_dir(d.z));
    float tclose = tlimit;
    bool hit = false;
    int sp = 0;
    int32_t code = (int32_t)S.root;

    auto push = [&](int32_t cc) {
        if (sp < SS)
            lds_code[sp * stride] = (uint32_t)cc;
        else if (sp < STACK4_MAX)
            ovf[(uint32_t)(sp - SS) * ovf_stride] = (uint32_t)cc;
        sp = sp < STACK4_MAX ? sp + 1 : sp;
    };
    auto pop = [&]() -> int32_t {
        if (sp == 0)
            return DONE;
        sp--;
        return (int32_t)(sp < SS ? lds_code[sp * stride]
                                 : ovf[(uint32_t)(sp - SS) * ovf_stride]);
    };
    // Inner node: push the hit children but the nearest, return the nearest
    // (or the next stack entry when no child is hit).
    auto node = [&](int32_t nc) -> int32_t {
        float4 lx, hx, ly, hy, lz, hz, cw;
        if (TOP && (uint32_t)nc < ntop) {
            const float4 *n = top + (size_t)nc * 7;
            lx = n[0]; hx = n[1]; ly = n[2]; hy = n[3]; lz = n[4]; hz = n[5];
            cw = n[6];
        } else {
            const float4 *n = S.nodes + (size_t)nc * 8;
            lx = n[0]; hx = n[1]; ly = n[2]; hy = n[3]; lz = n[4]; hz = n[5];
            cw = n[6];
        }
        if (COUNT)
            nodes++;
        float t0, t1, t2, t3;
        float k0 = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, o, inv, tclose, t0) ? t0 : INF;
        float k1 = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, o, inv, tclose, t1) ? t1 : INF;
        float k2 = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, o, inv, tclose, t2) ? t2 : INF;
        float k3 = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, o, inv, tclose, t3) ? t3 : INF;
        int32_t c0 = __float_as_int(cw.x), c1 = __float_as_int(cw.y),
                c2 = __float_as_int(cw.z), c3 = __float_as_int(cw.w);
        if (ANY) {
            // any hit: no order needed; the hit children move to the front
            // in slot order (3 % less shadow time than sorting them on the
            // 1080p mesh workload)
            if (k0 == INF) { k0 = k1; c0 = c1; k1 = k2; c1 = c2; k2 = k3; c2 = c3; k3 = INF; }
            if (k0 == INF) { k0 = k1; c0 = c1; k1 = k2; c1 = c2; k2 = INF; }
            if (k0 == INF) { k0 = k1; c0 = c1; k1 = INF; }
            if (k1 == INF) { k1 = k2; c1 = c2; k2 = k3; c2 = c3; k3 = INF; }
            if (k1 == INF) { k1 = k2; c1 = c2; k2 = INF; }
            if (k2 == INF) { k2 = k3; c2 = c3; k3 = INF; }
        } else {
            cx(k0, c0, k1, c1);
            cx(k2, c2, k3, c3);
            cx(k0, c0, k2, c2);
            cx(k1, c1, k3, c3);
            cx(k1, c1, k2, c2);
        }
        if (k0 == INF)
            return pop();
        // the hit children after the nearest, farthest first: they are a
        // prefix of the four (sorted, or moved to the front), so with np of
        // them after the nearest, slot sp + j takes child np - j.  Three
        // slots are written whatever np is (those at sp + np and above are
        // free), without a branch per child, while they all lie in LDS.
        const uint32_t np = (uint32_t)(k1 != INF) + (uint32_t)(k2 != INF)
                          + (uint32_t)(k3 != INF);
        if (sp + 3 <= SS) {
            lds_code[sp * stride] = (uint32_t)(np == 3u ? c3 : (np == 2u ? c2 : c1));
            lds_code[(sp + 1) * stride] = (uint32_t)(np == 3u ? c2 : c1);
            lds_code[(sp + 2) * stride] = (uint32_t)c1;
            sp += (int)np;
        } else {
            if (k3 != INF)
                push(c3);
            if (k2 != INF)
                push(c2);
            if (k1 != INF)
                push(c1);
        }
        return c0;
    };
    // One primitive record; true when an any-hit query is answered.
    auto test = [&](float4 A, float4 B, float4 C, uint32_t slot) -> bool {
        if (COUNT)
            prims++;
        float tt, uu = 0.0f, vv = 0.0f;
        bool h;
        uint32_t kind = __float_as_uint(B.w);
        if (kind == MRT_PRIMKIND_SPHERE)
            h = isect_sph(o, d, A, B, tt);
        else if (!EXT || kind == MRT_PRIMKIND_TRIANGLE)
            // (without EXT every other record is a triangle)
            h = isect_tri(o, d, A, B, C, S.one_sided != 0, tt, uu, vv);
        else
            h = isect_quadric(o, d, A, B, C, S.aux + (size_t)slot * 9, tt, uu);
        if (EXT && h && tt < tclose && S.alpha
            && alpha_skips(S.tx, S.aux + (size_t)slot * 9, A, B, C, o, d, tt, uu, vv))
            h = false;
        if (h && tt < tclose) {
            tclose = tt;
            hit = true;
            hit_slot = slot;
            bu_out = uu;
            bv_out = vv;
            return ANY;
        }
        return false;
    };
    // A leaf's next two records, fetched together: the six loads of a pair
    // are in flight at once, where one record at a time waits a full memory
    // latency per record (8 % more Msamples/s on the 1080p mesh workload;
    // all four at once gains nothing more and costs 18 VGPRs).  One pair per
    // loop iteration: the records after it stay the lane's code (the same
    // leaf code two records on), so no lane of a wave waits in an inner loop
    // over another lane's long leaf (k_trace_top on that workload: 1,168 ->
    // 1,224 Msamples/s).  True when an any-hit query is answered.
    auto leaf_pair = [&](int32_t &lc) -> bool {
        const uint32_t v = ~(uint32_t)lc, first = v >> 3, cnt = (v & 7u) + 1u;
        const float4 *p = S.prims + (size_t)first * 3;
        float4 A0 = p[0], B0 = p[1], C0 = p[2], A1, B1, C1;
        const bool two = cnt > 1;
        if (two) {
            A1 = p[3];
            B1 = p[4];
            C1 = p[5];
        }
        if (test(A0, B0, C0, first))
            return true;
        if (two && test(A1, B1, C1, first + 1))
            return true;
        lc = cnt > 2 ? (int32_t)~(((first + 2u) << 3) | (cnt - 3u)) : pop();
        return false;
    };

    for (;;) {
        if (code >= 0) {
            code = node(code);
        } else if (leaf_pair(code)) {
            t_out = tclose;
            return true;
        }
        if (code == DONE)
            break;
    }
    t_out = tclose;
    return hit;
}

// Closest hit over BVH + planes (hit_register_all, hit_register.c:109-123).
// code: leaf-order slot, HIT_PLANE | k, or HIT_NONE.
template <bool COUNT, int SS, bool EXT = true, bool TOP = false>
__device__ __forceinline__ bool closest_hit(const SceneDev &S, V3 o, V3 d,
                                            uint32_t *lds_code, int stride,
                                            uint32_t *ovf, uint32_t ovf_stride,
                                            float &t, uint32_t &code,
                                            float &bu, float &bv,
                                            uint32_t &nodes, uint32_t &prims,
                                            const float4 *top = nullptr,
                                            uint32_t ntop = 0)
{
    float tb = IT_MAX_F;
    uint32_t slot = HIT_NONE;
    float u = 0.0f, v = 0.0f;
    bool hit = traverse<false, COUNT, SS, EXT, TOP>(S, o, d, IT_MAX_F, lds_code,
                                                    stride, ovf, ovf_stride, tb,
                                                    slot, u, v, nodes, prims,
                                                    top, ntop);
    for (uint32_t k = 0; k < S.n_planes; k++) {
        float tp;
        if (isect_pln(o, d, S.planes[2 * k], S.planes[2 * k + 1], tp)
            && (!hit || tp < tb)) {
            hit = true;
            tb = tp;
            slot = HIT_PLANE | k;
        }
    }
    t = tb;
    code = hit ? slot : HIT_NONE;
    bu = u;
    bv = v;
    return hit;
}

// Occlusion below tlimit: BVH primitives below min(tlimit, 1000), planes
// below tlimit (check_plight_hit, src/ray_tracing/light/p_light.c:48-64).
template <bool COUNT, int SS, bool EXT = true, bool TOP = false>
__device__ __forceinline__ bool occluded(const SceneDev &S, V3 o, V3 d,
                                         float tlimit, uint32_t *lds_code,
                                         int stride, uint32_t *ovf,
                                         uint32_t ovf_stride,
                                         uint32_t &nodes, uint32_t &prims,
                                         const float4 *top = nullptr,
                                         uint32_t ntop = 0)
{
    float tl = fminf(tlimit, IT_MAX_F);
    float t;
    uint32_t slot;
    float u, v;
    if (traverse<true, COUNT, SS, EXT, TOP>(S, o, d, tl, lds_code, stride, ovf,
                                            ovf_stride, t, slot, u, v, nodes,
                                            prims, top, ntop))
        return true;
    for (uint32_t k = 0; k < S.n_planes; k++) {
        float tp;
        if (isect_pln(o, d, S.planes[2 * k], S.planes[2 * k + 1], tp)
            && tp < tlimit)
            return true;
    }
    return false;
}

}  // namespace mrt
