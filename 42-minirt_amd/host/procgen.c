/*
 * Seeded procedural scenes: the plain .rt-style scene (ground plane + three
 * spheres), displaced icospheres, a heightfield floor and an analytic HDRI
 * sky.  Every number comes from splitmix64 streams of the seed, in double,
 * so a seed gives the same scene on every host.
 */
#include "mrt_host_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.14159265358979323846

/* ---- value-noise fBm ----------------------------------------------------- */

static inline uint64_t hash4(int64_t x, int64_t y, int64_t z, uint64_t seed)
{
    uint64_t s = seed ^ ((uint64_t)x * 0x8CB92BA72F3D8DD7ull)
                 ^ ((uint64_t)y * 0xA0761D6478BD642Full)
                 ^ ((uint64_t)z * 0xE7037ED1A0B428DBull);
    return mrt_splitmix64(&s);
}

static inline double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed)
{
    return (double)(hash4(x, y, z, seed) >> 11) * (2.0 / 9007199254740992.0)
           - 1.0;
}

static inline double fade(double t)
{
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0);
}

static double value_noise(double x, double y, double z, uint64_t seed)
{
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++)
                c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty;
    double y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

/* 5-octave fBm in [-1, 1]. */
static double fbm3(double x, double y, double z, uint64_t seed)
{
    double sum = 0.0, amp = 1.0, norm = 0.0, f = 1.0;
    for (int o = 0; o < 5; o++) {
        sum += amp * value_noise(x * f, y * f, z * f, seed + (uint64_t)o);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

/* ---- materials ----------------------------------------------------------- */

static mrt_material mat_lambert(float a)
{
    mrt_material m;
    memset(&m, 0, sizeof(m));
    m.color[0] = m.color[1] = m.color[2] = a;
    m.ao = 1.0f;
    return m;
}

static mrt_material mat_metal(float a, float rough)
{
    mrt_material m = mat_lambert(a);
    m.metallic = 1.0f;
    m.roughness = rough;
    return m;
}

static mrt_material mat_glass(float ior)
{
    mrt_material m = mat_lambert(1.0f);
    m.ior = ior;
    m.transmission = 1.0f;
    return m;
}

#define MIX_GGX_LEVELS 256u
/* Palette used by the per-triangle mix: 0 lambert, 1 dielectric,
 * 2 + k GGX metal with roughness at the centre of bin k of [0.05, 0.6). */
static int add_mix_palette(mrt_host_scene *hs, uint32_t *base)
{
    mrt_material m = mat_lambert(0.8f);
    int r = mrt_hs_add_material(hs, &m, "mix_lambert");
    if (r < 0)
        return r;
    *base = (uint32_t)r;
    m = mat_glass(1.5f);
    if ((r = mrt_hs_add_material(hs, &m, "mix_glass")) < 0)
        return r;
    for (uint32_t k = 0; k < MIX_GGX_LEVELS; k++) {
        float rough = (float)(0.05 + 0.55 * ((double)k + 0.5) / MIX_GGX_LEVELS);
        m = mat_metal(0.8f, rough);
        if ((r = mrt_hs_add_material(hs, &m, NULL)) < 0)
            return r;
    }
    return MRT_OK;
}

static uint32_t draw_mix(uint64_t *st, uint32_t base)
{
    double u = mrt_u01(st);
    if (u < 0.60)
        return base;
    if (u < 0.85) {
        double r = mrt_u01(st);          /* roughness U[0.05, 0.6) */
        uint32_t k = (uint32_t)(r * MIX_GGX_LEVELS);
        if (k >= MIX_GGX_LEVELS)
            k = MIX_GGX_LEVELS - 1;
        return base + 2 + k;
    }
    return base + 1;
}

/* ---- plain scene --------------------------------------------------------- */

int mrt_gen_plain_scene(uint64_t seed, uint32_t width, uint32_t height,
                        mrt_host_scene **out)
{
    if (!out || width == 0 || height == 0)
        return MRT_E_INVALID;
    mrt_host_scene *hs = mrt_hs_new();
    if (!hs)
        return MRT_E_NOMEM;
    mrt_hs_set_resolution(hs, width, height);
    /* mat lam   204,204,204 0 0   0   0 1 0 0,0,0
     * mat metal 204,204,204 1 0.1 0   0 1 0 0,0,0
     * mat glass 255,255,255 0 0   1.5 1 1 0 0,0,0 */
    mrt_material lam = mat_lambert(204.0f / 255.0f);
    mrt_material metal = mat_metal(204.0f / 255.0f, 0.1f);
    mrt_material glass = mat_glass(1.5f);
    int ml = mrt_hs_add_material(hs, &lam, "lam");
    int mm = mrt_hs_add_material(hs, &metal, "metal");
    int mg = mrt_hs_add_material(hs, &glass, "glass");
    if (ml < 0 || mm < 0 || mg < 0) {
        mrt_hs_free(hs);
        return MRT_E_NOMEM;
    }
    float p0[3] = {0, 0, 0}, n0[3] = {0, 1, 0};
    mrt_hs_add_plane(hs, p0, n0, (uint32_t)ml);
    uint64_t st = seed;
    const float xs[3] = {-2.0f, 0.0f, 2.0f};
    const int mats[3] = {ml, mm, mg};
    for (int i = 0; i < 3; i++) {
        float jx = (float)(mrt_u01(&st) - 0.5);
        float jz = (float)(mrt_u01(&st) - 0.5);
        float c[3] = {xs[i] + jx, 1.0f, jz};
        mrt_hs_add_sphere(hs, c, 1.0f, (uint32_t)mats[i]);
    }
    mrt_light L;
    memset(&L, 0, sizeof(L));
    L.type = MRT_LIGHT_POINT;
    L.vec[0] = 0.0f;
    L.vec[1] = 5.0f;
    L.vec[2] = -2.0f;
    L.brightness = 1.0f;
    L.color[0] = L.color[1] = L.color[2] = 255.0f;
    mrt_hs_add_light(hs, &L);
    float white[3] = {255, 255, 255};
    mrt_hs_set_ambient(hs, 0.1f, white);
    mrt_camera cam;
    memset(&cam, 0, sizeof(cam));
    cam.pos[0] = 0.0f;
    cam.pos[1] = 1.5f;
    cam.pos[2] = -6.0f;
    cam.dir[2] = 1.0f;
    cam.fov_deg = 60.0f;
    cam.focus_dist = 1.0f;
    mrt_hs_set_camera(hs, &cam);
    *out = hs;
    return MRT_OK;
}

/* ---- displaced icosphere -------------------------------------------------- */

typedef struct {
    double   cx, cy, cz, radius, amp;
    uint64_t seed;
} disp_t;

static double disp_radius(const disp_t *d, const double p[3])
{
    return d->radius * (1.0 + d->amp * fbm3(p[0] * 2.0, p[1] * 2.0,
                                            p[2] * 2.0, d->seed));
}

static void normalize3(double v[3])
{
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}

static void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Position and outward surface normal of the displaced sphere along unit
 * direction p (normal from central differences of the radius field). */
static void disp_vertex(const disp_t *d, const double p[3], float pos[3],
                        float nrm[3])
{
    double r = disp_radius(d, p);
    pos[0] = (float)(d->cx + p[0] * r);
    pos[1] = (float)(d->cy + p[1] * r);
    pos[2] = (float)(d->cz + p[2] * r);
    double a[3] = {0, 1, 0};
    if (fabs(p[1]) > 0.9) {
        a[0] = 1;
        a[1] = 0;
    }
    double t1[3], t2[3];
    cross3(a, p, t1);
    normalize3(t1);
    cross3(p, t1, t2);
    const double h = 1e-4;
    double S[2][3];
    for (int k = 0; k < 2; k++) {
        const double *t = k ? t2 : t1;
        double pp[3] = {p[0] + h * t[0], p[1] + h * t[1], p[2] + h * t[2]};
        double pm[3] = {p[0] - h * t[0], p[1] - h * t[1], p[2] - h * t[2]};
        normalize3(pp);
        normalize3(pm);
        double rp = disp_radius(d, pp), rm = disp_radius(d, pm);
        for (int j = 0; j < 3; j++)
            S[k][j] = pp[j] * rp - pm[j] * rm;
    }
    double n[3];
    cross3(S[0], S[1], n);
    if (n[0] * p[0] + n[1] * p[1] + n[2] * p[2] < 0) {
        n[0] = -n[0];
        n[1] = -n[1];
        n[2] = -n[2];
    }
    normalize3(n);
    nrm[0] = (float)n[0];
    nrm[1] = (float)n[1];
    nrm[2] = (float)n[2];
}

typedef struct {
    mrt_host_scene *hs;
    const disp_t   *d;
    int             mix;
    uint32_t        mat, base;
    uint64_t        mst;
    int             err;
    int             count;
} ico_ctx;

static void emit_face(ico_ctx *c, const double a[3], const double b[3],
                      const double e[3])
{
    float v[9], n[9];
    disp_vertex(c->d, a, v, n);
    disp_vertex(c->d, b, v + 3, n + 3);
    disp_vertex(c->d, e, v + 6, n + 6);
    uint32_t m = c->mat;
    if (c->mix == 1)
        m = draw_mix(&c->mst, c->base);
    else if (c->mix == 2)
        m = (uint32_t)(mrt_u01(&c->mst) * c->mat) % c->mat;
    int r = mrt_hs_add_triangle(c->hs, v, n, m);
    if (r < 0)
        c->err = r;
    else
        c->count++;
}

static void subdivide(ico_ctx *c, const double a[3], const double b[3],
                      const double e[3], uint32_t level)
{
    if (c->err)
        return;
    if (level == 0) {
        emit_face(c, a, b, e);
        return;
    }
    double ab[3], be[3], ea[3];
    for (int k = 0; k < 3; k++) {
        ab[k] = a[k] + b[k];
        be[k] = b[k] + e[k];
        ea[k] = e[k] + a[k];
    }
    normalize3(ab);
    normalize3(be);
    normalize3(ea);
    subdivide(c, a, ab, ea, level - 1);
    subdivide(c, ab, b, be, level - 1);
    subdivide(c, ea, be, e, level - 1);
    subdivide(c, ab, be, ea, level - 1);
}

/* Point (i, j) of the freq x freq triangular grid on icosahedron face
 * (a, b, e): the direction of (freq - i - j) a + i b + e j, the three terms
 * summed in the order of the vertices' indices, so that a point on an edge
 * the face shares has the same bits in both faces (no cracks). */
static void grid_point(const double V[12][3], const int f[3], uint32_t i,
                       uint32_t j, uint32_t freq, double out[3])
{
    const double w[3] = {(double)(freq - i - j), (double)i, (double)j};
    int o[3] = {0, 1, 2};
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2 - x; y++)
            if (f[o[y]] > f[o[y + 1]]) {
                int t = o[y];
                o[y] = o[y + 1];
                o[y + 1] = t;
            }
    for (int k = 0; k < 3; k++)
        out[k] = (V[f[o[0]]][k] * w[o[0]] + V[f[o[1]]][k] * w[o[1]])
                 + V[f[o[2]]][k] * w[o[2]];
    normalize3(out);
}

/* The freq^2 triangles of face f's grid, wound as the face. */
static void geodesic_face(ico_ctx *c, const double V[12][3], const int f[3],
                          uint32_t freq)
{
    for (uint32_t j = 0; j < freq && !c->err; j++)
        for (uint32_t i = 0; i + j < freq && !c->err; i++) {
            double p00[3], p10[3], p01[3];
            grid_point(V, f, i, j, freq, p00);
            grid_point(V, f, i + 1, j, freq, p10);
            grid_point(V, f, i, j + 1, freq, p01);
            emit_face(c, p00, p10, p01);
            if (i + j + 1 < freq) {
                double p11[3];
                grid_point(V, f, i + 1, j + 1, freq, p11);
                emit_face(c, p10, p11, p01);
            }
        }
}

/* subdiv > 0: recursive 4-way splits (20 * 4^subdiv faces); freq > 0: each
 * face split into a freq x freq triangular grid (20 * freq^2 faces). */
static int icosphere(mrt_host_scene *hs, const float centre[3], float radius,
                     uint32_t subdiv, uint32_t freq, float amp, uint64_t seed,
                     int material_mix, uint32_t mat)
{
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    double V[12][3] = {{-1, t, 0}, {1, t, 0},  {-1, -t, 0}, {1, -t, 0},
                       {0, -1, t}, {0, 1, t},  {0, -1, -t}, {0, 1, -t},
                       {t, 0, -1}, {t, 0, 1},  {-t, 0, -1}, {-t, 0, 1}};
    static const int F[20][3] = {
        {0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
        {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
        {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
        {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    for (int i = 0; i < 12; i++)
        normalize3(V[i]);
    disp_t d = {centre[0], centre[1], centre[2], radius, amp, seed};
    ico_ctx c;
    memset(&c, 0, sizeof(c));
    c.hs = hs;
    c.d = &d;
    c.mix = material_mix;
    c.mat = mat;
    c.mst = seed ^ 0x5DEECE66Dull;
    if (material_mix == 1) {
        int r = add_mix_palette(hs, &c.base);
        if (r < 0)
            return r;
    } else if (material_mix == 2) {
        if (mat == 0 || mat > hs->n_mat)
            return MRT_E_INVALID;
    } else if (mat >= hs->n_mat) {
        return MRT_E_INVALID;
    }
    for (int f = 0; f < 20; f++) {
        if (freq)
            geodesic_face(&c, (const double(*)[3])V, F[f], freq);
        else
            subdivide(&c, V[F[f][0]], V[F[f][1]], V[F[f][2]], subdiv);
    }
    return c.err ? c.err : c.count;
}

int mrt_gen_icosphere(mrt_host_scene *hs, const float centre[3], float radius,
                      uint32_t subdiv, float amp, uint64_t seed,
                      int material_mix, uint32_t mat)
{
    if (!hs || !centre || !(radius > 0.0f) || subdiv > 10)
        return MRT_E_INVALID;
    return icosphere(hs, centre, radius, subdiv, 0, amp, seed, material_mix, mat);
}

int mrt_gen_icosphere_freq(mrt_host_scene *hs, const float centre[3],
                           float radius, uint32_t freq, float amp,
                           uint64_t seed, int material_mix, uint32_t mat)
{
    if (!hs || !centre || !(radius > 0.0f) || freq < 1 || freq > 1024)
        return MRT_E_INVALID;
    return icosphere(hs, centre, radius, 0, freq, amp, seed, material_mix, mat);
}

/* ---- heightfield floor ----------------------------------------------------- */

static double floor_height(double x, double z, double amp, uint64_t seed)
{
    return amp * fbm3(x * 0.8, 0.37, z * 0.8, seed);
}

/* Water: y0 plus a sum of eight seeded travelling sine ripples. */
static double water_height(double x, double z, double amp, uint64_t seed)
{
    uint64_t st = seed;
    double y = 0.0;
    for (int k = 0; k < 8; k++) {
        double ang = 2.0 * PI_D * mrt_u01(&st);
        double f = 2.0 + 10.0 * mrt_u01(&st);
        double ph = 2.0 * PI_D * mrt_u01(&st);
        y += sin(f * (x * cos(ang) + z * sin(ang)) + ph) / (1.0 + 0.3 * k);
    }
    return amp * y / 3.0;
}

typedef double (*height_fn)(double, double, double, uint64_t);

static int add_heightfield(mrt_host_scene *hs, uint32_t cells, double half,
                           double y0, double amp, uint64_t seed, height_fn fn,
                           int fixed_mat, uint32_t base);

static int add_floor(mrt_host_scene *hs, uint32_t cells, double half,
                     double amp, uint64_t seed, uint32_t base)
{
    return add_heightfield(hs, cells, half, 0.0, amp, seed, floor_height, -1,
                           base);
}

/* A cells x cells grid over [-half, half]^2 at height y0 + fn(x, z),
 * two triangles per cell with finite-difference vertex normals; material
 * fixed_mat, or drawn from the mix palette at base when fixed_mat < 0. */
static int add_heightfield(mrt_host_scene *hs, uint32_t cells, double half,
                           double y0, double amp, uint64_t seed, height_fn fn,
                           int fixed_mat, uint32_t base)
{
    uint32_t nv = cells + 1;
    float *P = malloc((size_t)nv * nv * 3 * sizeof(float));
    float *N = malloc((size_t)nv * nv * 3 * sizeof(float));
    if (!P || !N) {
        free(P);
        free(N);
        return MRT_E_NOMEM;
    }
    const double step = 2.0 * half / cells, h = 1e-4;
    for (uint32_t j = 0; j < nv; j++)
        for (uint32_t i = 0; i < nv; i++) {
            double x = -half + step * i, z = -half + step * j;
            double y = y0 + fn(x, z, amp, seed);
            double dx = (fn(x + h, z, amp, seed) - fn(x - h, z, amp, seed)) / (2 * h);
            double dz = (fn(x, z + h, amp, seed) - fn(x, z - h, amp, seed)) / (2 * h);
            double n[3] = {-dx, 1.0, -dz};
            normalize3(n);
            size_t o = ((size_t)j * nv + i) * 3;
            P[o] = (float)x;
            P[o + 1] = (float)y;
            P[o + 2] = (float)z;
            N[o] = (float)n[0];
            N[o + 1] = (float)n[1];
            N[o + 2] = (float)n[2];
        }
    uint64_t mst = seed ^ 0xF1007ull;
    int err = 0;
    for (uint32_t j = 0; j < cells && !err; j++)
        for (uint32_t i = 0; i < cells && !err; i++) {
            size_t q[4] = {(size_t)j * nv + i, (size_t)j * nv + i + 1,
                           (size_t)(j + 1) * nv + i + 1,
                           (size_t)(j + 1) * nv + i};
            const int tri[2][3] = {{0, 2, 1}, {0, 3, 2}};
            for (int t = 0; t < 2; t++) {
                float v[9], n[9];
                for (int k = 0; k < 3; k++) {
                    memcpy(v + 3 * k, P + q[tri[t][k]] * 3, 3 * sizeof(float));
                    memcpy(n + 3 * k, N + q[tri[t][k]] * 3, 3 * sizeof(float));
                }
                uint32_t m = fixed_mat >= 0 ? (uint32_t)fixed_mat
                                            : draw_mix(&mst, base);
                int r = mrt_hs_add_triangle(hs, v, n, m);
                if (r < 0) {
                    err = r;
                    break;
                }
            }
        }
    free(P);
    free(N);
    return err;
}

int mrt_gen_mesh_scene(uint64_t seed, uint32_t width, uint32_t height,
                       uint32_t subdiv, uint32_t floor_cells, float amp,
                       mrt_host_scene **out)
{
    if (!out || width == 0 || height == 0)
        return MRT_E_INVALID;
    mrt_host_scene *hs = mrt_hs_new();
    if (!hs)
        return MRT_E_NOMEM;
    mrt_hs_set_resolution(hs, width, height);
    float c[3] = {0.0f, 1.2f, 0.0f};
    int r = mrt_gen_icosphere(hs, c, 1.0f, subdiv, amp, seed, 1, 0);
    if (r < 0) {
        mrt_hs_free(hs);
        return r;
    }
    uint32_t base = 0;      /* the palette added by the icosphere */
    if (floor_cells) {
        r = add_floor(hs, floor_cells, 4.0, 0.08, seed ^ 0xABCDull, base);
        if (r < 0) {
            mrt_hs_free(hs);
            return r;
        }
    }
    mrt_light L;
    memset(&L, 0, sizeof(L));
    L.type = MRT_LIGHT_POINT;
    L.vec[1] = 5.0f;
    L.vec[2] = -2.0f;
    L.brightness = 1.0f;
    L.color[0] = L.color[1] = L.color[2] = 255.0f;
    mrt_hs_add_light(hs, &L);
    float white[3] = {255, 255, 255};
    mrt_hs_set_ambient(hs, 0.1f, white);
    mrt_camera cam;
    memset(&cam, 0, sizeof(cam));
    cam.pos[1] = 1.8f;
    cam.pos[2] = -4.0f;
    double dir[3] = {0.0, 1.1 - 1.8, 4.0};
    normalize3(dir);
    cam.dir[0] = (float)dir[0];
    cam.dir[1] = (float)dir[1];
    cam.dir[2] = (float)dir[2];
    cam.fov_deg = 60.0f;
    cam.focus_dist = 1.0f;
    mrt_hs_set_camera(hs, &cam);
    *out = hs;
    return MRT_OK;
}

/* ---- HDRI + caustics scene ---------------------------------------------- */

int mrt_gen_caustic_scene(uint64_t seed, uint32_t width, uint32_t height,
                          uint32_t subdiv, uint32_t floor_cells,
                          uint32_t water_cells, mrt_host_scene **out)
{
    if (!out)
        return MRT_E_INVALID;
    mrt_host_scene *hs = NULL;
    int r = mrt_gen_mesh_scene(seed, width, height, subdiv, floor_cells, 0.15f,
                               &hs);
    if (r < 0)
        return r;
    /* lit by the environment only */
    hs->n_light = 0;
    float white[3] = {255, 255, 255};
    mrt_hs_set_ambient(hs, 1.0f, white);
    mrt_material water = mat_glass(1.33f);
    water.color[0] = 0.92f;
    water.color[1] = 0.97f;
    water.color[2] = 1.0f;
    int wm = mrt_hs_add_material(hs, &water, "water");
    mrt_material glass = mat_glass(1.5f);
    int gm = mrt_hs_add_material(hs, &glass, "caustic_glass");
    if (wm < 0 || gm < 0) {
        mrt_hs_free(hs);
        return wm < 0 ? wm : gm;
    }
    if (water_cells) {
        r = add_heightfield(hs, water_cells, 4.0, 0.16, 0.012, seed ^ 0x77A7E5ull,
                            water_height, wm, 0);
        if (r < 0) {
            mrt_hs_free(hs);
            return r;
        }
    }
    float c[3] = {1.9f, 0.75f, -0.9f};
    r = mrt_hs_add_sphere(hs, c, 0.5f, (uint32_t)gm);
    if (r < 0) {
        mrt_hs_free(hs);
        return r;
    }
    *out = hs;
    return MRT_OK;
}

/* ---- HDRI sky ----------------------------------------------------------- */

int mrt_gen_hdri_sky(uint64_t seed, uint32_t w, uint32_t h,
                     float sun_radius_deg, float sun_radiance, float *rgb)
{
    if (!rgb || w < 2 || h < 2)
        return MRT_E_INVALID;
    uint64_t st = seed;
    double elev = (15.0 + 45.0 * mrt_u01(&st)) * PI_D / 180.0;
    double azim = 2.0 * PI_D * mrt_u01(&st);
    double sd[3] = {cos(elev) * cos(azim), sin(elev), cos(elev) * sin(azim)};
    double cos_sun = cos((double)sun_radius_deg * PI_D / 180.0);
    for (uint32_t y = 0; y < h; y++) {
        /* texel row y holds v = y / (h - 1), theta = (v - 0.5) * pi, the
         * direction convention of calc_inverse_transform_sampling_dir */
        double v = (double)y / (double)(h - 1);
        double th = (v - 0.5) * PI_D;
        for (uint32_t x = 0; x < w; x++) {
            double u = (double)x / (double)(w - 1);
            double ph = (0.5 - u) * 2.0 * PI_D;
            double d[3] = {cos(th) * cos(ph), sin(th), cos(th) * sin(ph)};
            double up = d[1];
            double r, g, b;
            if (up >= 0.0) {
                double t = pow(up, 0.4);
                r = 0.9 * (1 - t) + 0.25 * t;
                g = 0.95 * (1 - t) + 0.45 * t;
                b = 1.0 * (1 - t) + 0.9 * t;
            } else {
                double t = pow(-up, 0.5);
                r = 0.35 * (1 - t) + 0.12 * t;
                g = 0.33 * (1 - t) + 0.10 * t;
                b = 0.30 * (1 - t) + 0.08 * t;
            }
            double cs = d[0] * sd[0] + d[1] * sd[1] + d[2] * sd[2];
            if (cs >= cos_sun)
                r = g = b = sun_radiance;
            size_t o = ((size_t)y * w + x) * 3;
            rgb[o] = (float)r;
            rgb[o + 1] = (float)g;
            rgb[o + 2] = (float)b;
        }
    }
    return MRT_OK;
}
