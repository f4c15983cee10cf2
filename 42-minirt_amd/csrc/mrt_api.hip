// C ABI of libmrt_hip.so: context, scene upload, the wavefront render loop
// and read-back.  Everything here is host code around the kernels of
// mrt_kernels.hip; see include/mrt.h for the contract of each entry point.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "mrt_build.h"

#include "mrt_kernels.h"

using namespace mrt;

// Render lanes: up to this many passes of a render in flight, each on its
// own stream with its own queues.
constexpr int MAX_LANES = 4;
// Batch lanes get at least this many paths each (fewer for small passes).
constexpr uint32_t LANE_MIN_PATHS = 1u << 16;
// A pass starts once the pass before it has traced this bounce: its first,
// fullest launches then overlap the other pass's later ones, whose grids
// wait for a few long rays with most of the GPU idle (bounce 0: 1,506
// Msamples/s on the 1080p mesh workload at 256 spp in passes of 32, against
// 1,498 / 1,492 / 1,471 for bounces 1 / 2 / 3 and 1,461 for a pass split in
// two concurrent halves).
constexpr uint32_t LANE_STAGGER_BOUNCE = 0;

// Staging of mrt_trace_rays_noalloc: origins (3 floats), directions (3),
// primitive ids (1 word) and t (1 float) of up to MRT_TRACE_NOALLOC_MAX rays,
// 32 bytes a ray, 8 MiB in all.  Each GPU holds its own copy from the
// library's load until the process ends.  Host code never passes `scratch`
// itself (on the host it names a stand-in, not device memory): it uses the
// address hipGetSymbolAddress returns on each device, kept in
// g_scratch_dev[device] by mrt_init.
constexpr size_t NOALLOC_WORDS = (size_t)MRT_TRACE_NOALLOC_MAX * 8;
static __device__ __align__(256) float scratch[NOALLOC_WORDS];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_dev[MAX_DEVICES] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current).
static hipError_t scratch_on_device(int device, float **out)
{
    if (device < 0 || device >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        if (!p)
            return hipErrorInvalidSymbol;
        g_scratch_dev[device] = static_cast<float *>(p);
    }
    *out = g_scratch_dev[device];
    return hipSuccess;
}

// Host threads of a context that mrt_readback_add splits its add over (a
// 1080p frame is 6.2 M doubles): started once, so a tick starts no thread.
// run(f) calls f(0) on the caller and f(1 .. size-1) on the workers and
// returns when every call has.
class HostPool {
public:
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : th_)
            t.join();
    }
    unsigned size() const { return (unsigned)th_.size() + 1; }
    void start(unsigned n)
    {
        for (unsigned k = 1; k < n; k++)
            th_.emplace_back([this, k]() { work(k); });
    }
    void run(const std::function<void(unsigned)> &f)
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            job_ = &f;
            pending_ = (unsigned)th_.size();
            gen_++;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [this]() { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    void work(unsigned k)
    {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(unsigned)> *f;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&]() { return stop_ || gen_ != seen; });
                if (stop_)
                    return;
                seen = gen_;
                f = job_;
            }
            (*f)(k);
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0)
                done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(unsigned)> *job_ = nullptr;
    uint64_t gen_ = 0;
    unsigned pending_ = 0;
    bool stop_ = false;
};

// mrt_readback_add copies the sums in chunks and adds each as it arrives
constexpr int READBACK_CHUNKS = 8;

struct mrt_ctx {
    int          device = 0;
    float       *scratch_dev = nullptr;  // this device's `scratch`
    hipStream_t  stream = nullptr;
    uint32_t     W = 0, H = 0;
    char         err[512] = {0};
    int          num_cu = 0;
    uint32_t     trace_blocks = 0, grid_blocks = 0;
    uint32_t    *d_ovf = nullptr;    // trace-stack overflow columns, one
                                     // set per render lane
    uint32_t     ovf_blocks = 0;     // trace grid size each set covers
    uint32_t     ovf_sets = 0;
    // render lanes 1.. run on their own streams, forked from and joined to
    // the context's stream (lane 0)
    uint32_t     lanes = 2;          // MRT_OPT_LANES
    hipStream_t  lane_stream[MAX_LANES] = {};
    hipEvent_t   ev_fork = nullptr, ev_join[MAX_LANES] = {};
    hipEvent_t   ev_stag[MAX_LANES] = {}, ev_acc[MAX_LANES] = {};

    // scene
    float4      *d_nodes = nullptr, *d_prims = nullptr, *d_planes = nullptr;
    float4      *d_mats = nullptr, *d_lights = nullptr;
    float       *d_nrm = nullptr;
    bool         has_quad = false;   // the scene has cylinders / cones / ...
    float4      *d_tex = nullptr;    // textures: 3 float4 each
    uint32_t    *d_texels = nullptr;
    uint4       *d_mtex = nullptr;   // material texture slots: 3 uint4 each
    float       *d_tri_uv = nullptr;
    bool         textured = false;   // some material has a texture slot
    int          builder = MRT_BUILDER_HOST;
    bool         built_on_device = false;  // the last upload's BVH
    bool         alpha = false;      // ... a colour image with alpha < 255
    uint32_t     n_nodes = 0, n_prims = 0, n_planes = 0, n_mats = 0;
    uint32_t     bvh_depth = 0;
    std::vector<mrt_light> lights;
    std::vector<mrt_material> mats_host;
    bool         has_scene = false;
    mrt_camera   cam{};
    float        ambient_ratio = 0.0f, ambient_color[3] = {0, 0, 0};
    float        ior_all = 1.0f;
    uint32_t     flags = 0;

    // environment map
    float4      *d_env_tex = nullptr;
    float       *d_env_pdf = nullptr, *d_env_marg_inv = nullptr;
    float       *d_env_cond_inv = nullptr, *d_mis_marg = nullptr;
    float       *d_mis_cond = nullptr, *d_mis_pcell = nullptr;
    uint32_t    *d_mis_gm = nullptr, *d_mis_gc = nullptr;  // guide tables
    uint32_t     guide_m = 0, guide_c = 0;
    uint32_t     env_w = 0, env_h = 0;
    float        env_exposure = 0.0f;
    bool         has_env = false;

    // photon map
    float4      *d_ph_rec = nullptr;   // 2 per photon: position, power
    uint2       *d_range = nullptr;    // per bucket: first, one past last
    uint32_t    *d_occ = nullptr;      // per bucket: one bit, nonempty
    uint32_t     ph_n = 0, ph_table = 0, ph_k = 15;
    float        ph_cell = 0.0f, ph_radius = 0.0f, ph_norm = 1.0f;
    bool         has_photons = false;
    float        scene_c[3] = {0, 0, 0}, scene_r = 0.0f;

    // accumulation
    float       *d_accum = nullptr;
    float4      *d_tape = nullptr;    // the adjoint's tape
    size_t       tape_cap = 0;        // float4 entries
    float       *h_stage = nullptr;   // pinned staging of mrt_readback_add
    hipEvent_t   ev_rb[READBACK_CHUNKS] = {};  // its chunks' arrival
    HostPool    *pool = nullptr;      // its add's threads
    uint32_t    *d_ctr = nullptr;
    unsigned long long *d_stats = nullptr;

    // wavefront buffers
    uint32_t     cap_paths = 0, cap_shadow = 0;
    float4      *q_o[2] = {nullptr, nullptr}, *q_d[2] = {nullptr, nullptr};
    float4      *q_t[2] = {nullptr, nullptr};
    float4      *hit = nullptr, *acc = nullptr;
    float4      *s_o = nullptr, *s_d = nullptr;
    float2      *s_c = nullptr;
    uint8_t     *vis = nullptr;
    float4      *p_a = nullptr, *p_c = nullptr, *p_x = nullptr, *p_k = nullptr;
    uint4       *p_b = nullptr;

    // options
    uint64_t     seed = 0x6D696E6952542121ull;
    int          estimator = MRT_EST_REFERENCE;
    int          stats_on = 0;
    // paths in flight per wavefront pass: 128 M (64 spp at 1080p, about
    // 33 GB of queues per render lane) so that each bounce launch is long
    // against the tail of its slowest rays (1080p mesh workload at 256 spp,
    // two lanes: 32 M: 1,461, 64 M: 1,501, 128 M: 1,521, 256 M: 1,539
    // Msamples/s; earlier, at 2 spp per step: 4 M: 856, 32 M: 1,185)
    uint32_t     path_budget = 1u << 27;
    bool         trace_top = true;   // MRT_OPT_TRACE_TOP
    bool         trace_refill = true;   // MRT_OPT_TRACE_REFILL
    uint32_t     refill_blocks = 0;     // its resident grid

    // statistics
    uint64_t     paths = 0;
    double       ms[7] = {0, 0, 0, 0, 0, 0, 0};
    uint64_t     launches[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // per launch class
};

static int fail(mrt_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(ctx, call)                                                    \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return fail((ctx), MRT_E_HIP, "%s failed: %s", #call,            \
                        hipGetErrorString(e_));                              \
    } while (0)

static inline float bits_f(uint32_t u)
{
    float f;
    memcpy(&f, &u, sizeof(f));
    return f;
}

static inline uint32_t bits_u(float f)
{
    uint32_t u;
    memcpy(&u, &f, sizeof(u));
    return u;
}

template <typename T>
static void dfree(T *&p)
{
    if (p)
        (void)hipFree((void *)p);
    p = nullptr;
}

template <typename T>
static hipError_t dalloc(T *&p, size_t count)
{
    dfree(p);
    if (count == 0)
        count = 1;
    return hipMalloc((void **)&p, count * sizeof(T));
}

// Overflow columns for persistent trace grids of up to `blocks` blocks, one
// set for each of `sets` render lanes (their grids may run together).
static hipError_t ensure_ovf(mrt_ctx *c, uint32_t blocks, uint32_t sets = 1)
{
    if (blocks <= c->ovf_blocks && sets <= c->ovf_sets && c->d_ovf)
        return hipSuccess;
    blocks = std::max(blocks, c->ovf_blocks);
    sets = std::max(sets, c->ovf_sets);
    hipError_t e = hipStreamSynchronize(c->stream);
    for (int l = 1; l < MAX_LANES && e == hipSuccess; l++)
        if (c->lane_stream[l])
            e = hipStreamSynchronize(c->lane_stream[l]);
    if (e != hipSuccess)
        return e;
    c->ovf_blocks = c->ovf_sets = 0;
    e = dalloc(c->d_ovf, (size_t)sets * TRACE_OVF_BUF * blocks * TRACE_BLOCK);
    if (e == hipSuccess) {
        c->ovf_blocks = blocks;
        c->ovf_sets = sets;
    }
    return e;
}

static StackOvf stack_ovf(const mrt_ctx *c, uint32_t lane = 0)
{
    return StackOvf{c->d_ovf + (size_t)lane * TRACE_OVF_BUF * c->ovf_blocks * TRACE_BLOCK,
                    c->ovf_blocks};
}

extern "C" const char *mrt_version(void) { return "mrt_hip 0.1 gfx950"; }

extern "C" const char *mrt_last_error(const mrt_ctx *ctx)
{
    return ctx ? ctx->err : "null context";
}

extern "C" int mrt_init(int device, uint32_t width, uint32_t height,
                        mrt_ctx **out)
{
    if (!out || width == 0 || height == 0
        || (uint64_t)width * height > (1ull << 28))
        return MRT_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return MRT_E_NODEVICE;
    if (device < 0 || device >= ndev)
        return MRT_E_NODEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return MRT_E_NODEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return MRT_E_NODEVICE;
    mrt_ctx *c = new (std::nothrow) mrt_ctx();
    if (!c)
        return MRT_E_NOMEM;
    c->device = device;
    c->W = width;
    c->H = height;
    c->num_cu = prop.multiProcessorCount;
    int rc = MRT_OK;
    do {
        if (hipSetDevice(device) != hipSuccess
            || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)
                   != hipSuccess
            || hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)
                   != hipSuccess) {
            rc = MRT_E_HIP;
            break;
        }
        for (int l = 0; l < MAX_LANES && rc == MRT_OK; l++)
            if ((l && hipStreamCreateWithFlags(&c->lane_stream[l],
                                               hipStreamNonBlocking) != hipSuccess)
                || hipEventCreateWithFlags(&c->ev_join[l], hipEventDisableTiming)
                       != hipSuccess
                || hipEventCreateWithFlags(&c->ev_stag[l], hipEventDisableTiming)
                       != hipSuccess
                || hipEventCreateWithFlags(&c->ev_acc[l], hipEventDisableTiming)
                       != hipSuccess)
                rc = MRT_E_HIP;
        if (rc != MRT_OK)
            break;
        size_t npx = (size_t)width * height;
        if (dalloc(c->d_accum, npx * 3) != hipSuccess
            || dalloc(c->d_ctr, MAX_LANES * CTR_COUNT) != hipSuccess
            || dalloc(c->d_stats, STAT_COUNT) != hipSuccess) {
            rc = MRT_E_NOMEM;
            break;
        }
        if (hipMemsetAsync(c->d_accum, 0, npx * 3 * sizeof(float), c->stream)
                != hipSuccess
            || hipMemsetAsync(c->d_ctr, 0, MAX_LANES * CTR_COUNT * sizeof(uint32_t),
                              c->stream) != hipSuccess
            || hipMemsetAsync(c->d_stats, 0, STAT_COUNT * 8, c->stream)
                   != hipSuccess
            || hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = MRT_E_HIP;
            break;
        }
        if (scratch_on_device(device, &c->scratch_dev) != hipSuccess) {
            rc = MRT_E_HIP;
            break;
        }
        c->trace_blocks = (uint32_t)(c->num_cu * trace_blocks_per_cu());
        c->refill_blocks = std::min(c->trace_blocks,
                                    (uint32_t)(c->num_cu * refill_blocks_per_cu()));
        // grid of the grid-stride kernels (shade, light sums, gather): 32
        // blocks per CU, 1.6 % faster renders than 8 (no change at 128)
        c->grid_blocks = (uint32_t)c->num_cu * 32u;
        if (ensure_ovf(c, c->trace_blocks)
                != hipSuccess) {
            rc = MRT_E_NOMEM;
            break;
        }
    } while (0);
    if (rc != MRT_OK) {
        mrt_destroy(c);
        return rc;
    }
    *out = c;
    return MRT_OK;
}

static void free_wavefront(mrt_ctx *c)
{
    for (int k = 0; k < 2; k++) {
        dfree(c->q_o[k]);
        dfree(c->q_d[k]);
        dfree(c->q_t[k]);
    }
    dfree(c->hit);
    dfree(c->acc);
    dfree(c->s_o);
    dfree(c->s_d);
    dfree(c->s_c);
    dfree(c->vis);
    dfree(c->p_a);
    dfree(c->p_b);
    dfree(c->p_c);
    dfree(c->p_x);
    dfree(c->p_k);
    c->cap_paths = c->cap_shadow = 0;
}

static void free_scene(mrt_ctx *c)
{
    dfree(c->d_nodes);
    dfree(c->d_prims);
    dfree(c->d_nrm);
    dfree(c->d_planes);
    dfree(c->d_mats);
    dfree(c->d_lights);
    dfree(c->d_tex);
    dfree(c->d_texels);
    dfree(c->d_mtex);
    dfree(c->d_tri_uv);
    c->textured = c->alpha = false;
    c->has_scene = false;
}

static void free_env(mrt_ctx *c);

static void free_photons(mrt_ctx *c)
{
    dfree(c->d_ph_rec);
    dfree(c->d_range);
    dfree(c->d_occ);
    c->ph_n = 0;
    c->has_photons = false;
}

extern "C" void mrt_destroy(mrt_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    for (int l = 1; l < MAX_LANES; l++)
        if (c->lane_stream[l])
            (void)hipStreamSynchronize(c->lane_stream[l]);
    free_wavefront(c);
    free_scene(c);
    free_env(c);
    free_photons(c);
    dfree(c->d_accum);
    dfree(c->d_tape);
    if (c->h_stage)
        (void)hipHostFree(c->h_stage);
    for (hipEvent_t e : c->ev_rb)
        if (e)
            (void)hipEventDestroy(e);
    delete c->pool;
    dfree(c->d_ctr);
    dfree(c->d_stats);
    dfree(c->d_ovf);
    for (int l = 0; l < MAX_LANES; l++) {
        if (c->lane_stream[l])
            (void)hipStreamDestroy(c->lane_stream[l]);
        for (hipEvent_t e : {c->ev_join[l], c->ev_stag[l], c->ev_acc[l]})
            if (e)
                (void)hipEventDestroy(e);
    }
    if (c->ev_fork)
        (void)hipEventDestroy(c->ev_fork);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

static int upload(mrt_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (!bytes)
        return MRT_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return MRT_OK;
}

// A distance from pos that no point of the quadric exceeds: height + radius
// for cylinders and cones, height + max(a, b) sqrt(shape + (height/2c)^2)
// for hyperboloids (their radius at |z| = height / 2).
static float quadric_extent(const mrt_quadric *q)
{
    float r = q->radius;
    if (q->kind == MRT_QUADRIC_HYPERBOLOID) {
        float hc = 0.5f * q->height / q->c;
        r = fmaxf(q->a, q->b) * sqrtf(fmaxf(0.0f, q->shape + hc * hc));
    }
    return q->height + r;
}

static int validate_scene(mrt_ctx *c, const mrt_scene *s)
{
    if (s->n_materials == 0 || !s->materials)
        return fail(c, MRT_E_INVALID, "scene has no materials");
    if (s->n_lights > 64 || (s->n_lights && !s->lights))
        return fail(c, MRT_E_LIMIT, "at most 64 lights");
    for (uint32_t i = 0; i < s->n_triangles; i++)
        if (s->tri_mat[i] >= s->n_materials)
            return fail(c, MRT_E_INVALID, "triangle %u: material out of range", i);
    for (uint32_t i = 0; i < s->n_spheres; i++)
        if (s->sph_mat[i] >= s->n_materials || !(s->sph[4 * i + 3] > 0.0f))
            return fail(c, MRT_E_INVALID, "sphere %u: bad material or radius", i);
    if (s->n_quadrics && !s->quadrics)
        return fail(c, MRT_E_INVALID, "quadrics without an array");
    for (uint32_t i = 0; i < s->n_quadrics; i++) {
        const mrt_quadric &q = s->quadrics[i];
        if (q.kind > MRT_QUADRIC_HYPERBOLOID || q.mat >= s->n_materials
            || q.mat_top >= s->n_materials || q.mat_bot >= s->n_materials)
            return fail(c, MRT_E_INVALID, "quadric %u: bad kind or material", i);
    }
    if (s->n_textures && !s->textures)
        return fail(c, MRT_E_INVALID, "textures without an array");
    for (uint32_t i = 0; i < s->n_textures; i++) {
        const mrt_texture &t = s->textures[i];
        if (t.type > MRT_TEX_BLACK
            || (t.type == MRT_TEX_IMAGE
                && (t.width == 0 || t.height == 0 || !s->texels
                    || (uint64_t)t.texel0 + (uint64_t)t.width * t.height > s->n_texels)))
            return fail(c, MRT_E_INVALID, "texture %u: bad type or texels", i);
    }
    if (s->material_tex)
        for (uint32_t k = 0; k < s->n_materials; k++) {
            const uint32_t *sl = &s->material_tex[k].color;
            for (int j = 0; j < 8; j++)
                if (sl[j] != MRT_NO_TEX && sl[j] >= s->n_textures)
                    return fail(c, MRT_E_INVALID, "material %u: texture out of range", k);
        }
    if (s->n_planes && (!s->pln || !s->pln_mat))
        return fail(c, MRT_E_INVALID, "planes without arrays");
    for (uint32_t i = 0; i < s->n_planes; i++)
        if (s->pln_mat[i] >= s->n_materials)
            return fail(c, MRT_E_INVALID, "plane %u: material out of range", i);
    for (uint32_t k = 0; k < s->n_lights; k++)
        if (s->lights[k].shadow_samples > 4096)
            return fail(c, MRT_E_LIMIT, "light %u: more than 4096 shadow samples", k);
    return MRT_OK;
}

// Textures (type, w, h, texel0) (scale, c1) (c2, 0) and per material the
// slots (color, metallic, roughness, transmission) (ao, emission strength,
// emission colour, normal) (normal intensity, scale, 0, 0).  textured: a
// material has a slot; alpha: a colour slot holds an image with a texel of
// alpha < 255 (the traversal then runs the alpha test).
static void pack_textures(const mrt_scene *s, std::vector<float4> &tex,
                          std::vector<uint4> &mtex, bool &textured, bool &alpha)
{
    textured = alpha = false;
    if (!s->material_tex || !s->n_textures)
        return;
    for (uint32_t k = 0; k < s->n_materials; k++) {
        const mrt_material_tex &t = s->material_tex[k];
        const uint32_t *sl = &t.color;
        for (int j = 0; j < 8; j++)
            textured |= sl[j] != MRT_NO_TEX;
        mtex.push_back(make_uint4(t.color, t.metallic, t.roughness, t.transmission));
        mtex.push_back(make_uint4(t.ao, t.emission_strength, t.emission_color, t.normal));
        mtex.push_back(make_uint4(bits_u(t.normal_intensity), bits_u(t.scale), 0u, 0u));
        if (t.color != MRT_NO_TEX && s->textures[t.color].type == MRT_TEX_IMAGE) {
            const mrt_texture &x = s->textures[t.color];
            for (size_t i = 0; !alpha && i < (size_t)x.width * x.height; i++)
                alpha = s->texels[4 * ((size_t)x.texel0 + i) + 3] != 255;
        }
    }
    if (!textured) {
        mtex.clear();
        return;
    }
    for (uint32_t k = 0; k < s->n_textures; k++) {
        const mrt_texture &t = s->textures[k];
        tex.push_back(make_float4(bits_f(t.type), bits_f(t.width), bits_f(t.height),
                                  bits_f(t.texel0)));
        tex.push_back(make_float4(t.scale, t.c1[0], t.c1[1], t.c1[2]));
        tex.push_back(make_float4(t.c2[0], t.c2[1], t.c2[2], 0.0f));
    }
}

// Planes (point | material, normal | 0), materials and lights as the
// kernels read them; material indices are offset by mat0.
static void pack_scene(const mrt_scene *s, uint32_t mat0, std::vector<float4> &planes,
                       std::vector<float4> &mats, std::vector<float4> &lights)
{
    for (uint32_t k = 0; k < s->n_planes; k++) {
        const float *q = s->pln + 6 * (size_t)k;
        planes.push_back(make_float4(q[0], q[1], q[2], bits_f(mat0 + s->pln_mat[k])));
        planes.push_back(make_float4(q[3], q[4], q[5], 0.0f));
    }
    for (uint32_t k = 0; k < s->n_materials; k++) {
        const mrt_material &m = s->materials[k];
        mats.push_back(make_float4(m.color[0], m.color[1], m.color[2], m.metallic));
        mats.push_back(make_float4(m.roughness, m.ior, m.transmission, m.ao));
        mats.push_back(make_float4(m.emission_strength, m.emission_color[0],
                                   m.emission_color[1], m.emission_color[2]));
    }
    for (uint32_t k = 0; k < s->n_lights; k++) {
        const mrt_light &l = s->lights[k];
        lights.push_back(make_float4(l.vec[0], l.vec[1], l.vec[2], l.brightness));
        lights.push_back(make_float4(l.color[0], l.color[1], l.color[2], l.radius));
        lights.push_back(make_float4(bits_f((uint32_t)l.type),
                                     bits_f((uint32_t)l.shadow_samples), 0.0f, 0.0f));
    }
}

extern "C" int mrt_upload_scene(mrt_ctx *c, const mrt_scene *s)
{
    if (!c || !s)
        return MRT_E_INVALID;
    int vrc = validate_scene(c, s);
    if (vrc != MRT_OK)
        return vrc;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_scene(c);

    mrt_bvh bvh;
    mrt_bvh4 bvh4;
    memset(&bvh, 0, sizeof(bvh));
    memset(&bvh4, 0, sizeof(bvh4));
    DeviceBvh D;
    memset(&D, 0, sizeof(D));
    bool on_dev = false;
    int rc = MRT_OK;
    if (c->builder == MRT_BUILDER_DEVICE) {
        const char *why = "";
        rc = device_bvh_build(c->stream, s, &D, &why);
        if (rc == MRT_OK)
            on_dev = true;
        else if (rc != MRT_E_UNSUPPORTED)
            return fail(c, rc, "device BVH build failed: %s", why);
    }
    if (!on_dev) {
        rc = mrt_bvh_build(s, MRT_BVH_MAX_LEAF, &bvh);
        if (rc != MRT_OK)
            return fail(c, rc, "BVH build failed (%d)", rc);
        if (bvh.depth > MRT_BVH_MAX_DEPTH) {
            mrt_bvh_free(&bvh);
            return fail(c, MRT_E_LIMIT, "BVH depth %u exceeds %u", bvh.depth,
                        MRT_BVH_MAX_DEPTH);
        }
        rc = mrt_bvh4_build(&bvh, &bvh4);
        if (rc != MRT_OK) {
            mrt_bvh_free(&bvh);
            return fail(c, rc, "4-wide BVH build failed (%d)", rc);
        }
        if (bvh4.depth > MRT_BVH4_MAX_DEPTH) {
            mrt_bvh_free(&bvh);
            mrt_bvh4_free(&bvh4);
            return fail(c, MRT_E_LIMIT, "4-wide BVH depth %u exceeds %u", bvh4.depth,
                        MRT_BVH4_MAX_DEPTH);
        }
    } else {
        c->d_nodes = D.nodes;
        c->d_prims = D.prims;
        c->d_nrm = D.nrm;
    }
    const uint32_t n_nodes4 = on_dev ? D.n_nodes : bvh4.n_nodes;
    const uint32_t n_prims = on_dev ? D.n_prims : bvh.n_prims;
    std::vector<float4> planes, mats, lights;
    pack_scene(s, 0, planes, mats, lights);
    std::vector<float4> tex;
    std::vector<uint4> mtex;
    bool textured = false, alpha = false;
    pack_textures(s, tex, mtex, textured, alpha);
    rc = MRT_OK;
    if ((!on_dev
         && (dalloc(c->d_nodes, (size_t)bvh4.n_nodes * 8) != hipSuccess
             || dalloc(c->d_prims, (size_t)bvh.n_prims * 3) != hipSuccess
             || dalloc(c->d_nrm, (size_t)bvh.n_prims * 9) != hipSuccess))
        || dalloc(c->d_planes, planes.size()) != hipSuccess
        || dalloc(c->d_mats, mats.size()) != hipSuccess
        || dalloc(c->d_lights, lights.size()) != hipSuccess
        || (textured
            && (dalloc(c->d_tex, tex.size()) != hipSuccess
                || dalloc(c->d_texels, s->n_texels) != hipSuccess
                || dalloc(c->d_mtex, mtex.size()) != hipSuccess
                || (s->tri_uv
                    && dalloc(c->d_tri_uv, (size_t)s->n_triangles * 6) != hipSuccess))))
        rc = fail(c, MRT_E_NOMEM, "device allocation for the scene failed");
    if (rc == MRT_OK && !on_dev)
        rc = upload(c, c->d_nodes, bvh4.nodes, (size_t)bvh4.n_nodes * 128);
    if (rc == MRT_OK && !on_dev)
        rc = upload(c, c->d_prims, bvh.prims, (size_t)bvh.n_prims * 48);
    if (rc == MRT_OK && !on_dev)
        rc = upload(c, c->d_nrm, bvh.shade_nrm, (size_t)bvh.n_prims * 36);
    if (rc == MRT_OK)
        rc = upload(c, c->d_planes, planes.data(), planes.size() * 16);
    if (rc == MRT_OK)
        rc = upload(c, c->d_mats, mats.data(), mats.size() * 16);
    if (rc == MRT_OK)
        rc = upload(c, c->d_lights, lights.data(), lights.size() * 16);
    if (rc == MRT_OK && textured) {
        rc = upload(c, c->d_tex, tex.data(), tex.size() * 16);
        if (rc == MRT_OK)
            rc = upload(c, c->d_texels, s->texels, (size_t)s->n_texels * 4);
        if (rc == MRT_OK)
            rc = upload(c, c->d_mtex, mtex.data(), mtex.size() * 16);
        if (rc == MRT_OK && s->tri_uv)
            rc = upload(c, c->d_tri_uv, s->tri_uv, (size_t)s->n_triangles * 24);
    }
    if (rc == MRT_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(c, MRT_E_HIP, "scene upload failed");
    c->n_nodes = n_nodes4;
    c->n_prims = n_prims;
    c->built_on_device = on_dev;
    c->has_quad = s->n_quadrics > 0;
    c->textured = textured;
    c->alpha = alpha;
    c->bvh_depth = on_dev ? D.depth : bvh4.depth;
    mrt_bvh_free(&bvh);
    mrt_bvh4_free(&bvh4);
    if (rc != MRT_OK) {
        free_scene(c);
        return rc;
    }
    c->n_planes = s->n_planes;
    c->n_mats = s->n_materials;
    c->mats_host.assign(s->materials, s->materials + s->n_materials);
    c->lights.assign(s->lights, s->lights + s->n_lights);
    c->cam = s->camera;
    c->ambient_ratio = s->ambient_ratio;
    memcpy(c->ambient_color, s->ambient_color, sizeof(c->ambient_color));
    c->ior_all = s->ior_all > 0.0f ? s->ior_all : 1.0f;
    c->flags = s->flags;
    c->has_scene = true;
    /* bounding sphere of the BVH primitives (emission disk of environment
     * photons): box of triangle vertices and sphere extents */
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < s->n_triangles * 3; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], s->tri_pos[3 * (size_t)i + k]);
            hi[k] = fmaxf(hi[k], s->tri_pos[3 * (size_t)i + k]);
        }
    for (uint32_t i = 0; i < s->n_spheres; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], s->sph[4 * i + k] - s->sph[4 * i + 3]);
            hi[k] = fmaxf(hi[k], s->sph[4 * i + k] + s->sph[4 * i + 3]);
        }
    for (uint32_t i = 0; i < s->n_quadrics; i++) {
        const float e = quadric_extent(&s->quadrics[i]);
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], s->quadrics[i].pos[k] - e);
            hi[k] = fmaxf(hi[k], s->quadrics[i].pos[k] + e);
        }
    }
    float d2 = 0.0f;
    for (int k = 0; k < 3; k++) {
        c->scene_c[k] = 0.5f * (lo[k] + hi[k]);
        d2 = d2 + (hi[k] - lo[k]) * (hi[k] - lo[k]);
    }
    c->scene_r = 0.5f * sqrtf(d2);
    free_photons(c);
    return MRT_OK;
}

extern "C" int mrt_set_camera(mrt_ctx *c, const mrt_camera *cam)
{
    if (!c || !cam)
        return MRT_E_INVALID;
    c->cam = *cam;
    return MRT_OK;
}

static void free_env(mrt_ctx *c)
{
    dfree(c->d_env_tex);
    dfree(c->d_env_pdf);
    dfree(c->d_env_marg_inv);
    dfree(c->d_env_cond_inv);
    dfree(c->d_mis_marg);
    dfree(c->d_mis_cond);
    dfree(c->d_mis_pcell);
    dfree(c->d_mis_gm);
    dfree(c->d_mis_gc);
    c->has_env = false;
}

extern "C" int mrt_upload_hdri(mrt_ctx *c, const float *rgb, uint32_t w,
                               uint32_t h, float exposure)
{
    if (!c)
        return MRT_E_INVALID;
    if (!rgb) {                        /* NULL removes the environment */
        free_env(c);
        return MRT_OK;
    }
    if (w < 2 || h < 2 || (uint64_t)w * h > (1ull << 26))
        return fail(c, MRT_E_LIMIT, "HDRI size %ux%u out of range", w, h);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_env(c);
    // the tables are built on the device from the uploaded texels (the
    // host builders in hdri_tables.c give the same bits)
    size_t n = (size_t)w * h;
    std::vector<double> band(h - 1);
    const double pi = 3.14159265358979323846;
    for (uint32_t y = 0; y + 1 < h; y++) {
        double elo = ((double)y / (h - 1) - 0.5) * pi;
        double ehi = ((double)(y + 1) / (h - 1) - 0.5) * pi;
        band[y] = sin(ehi) - sin(elo);
    }
    const double sc = exposure == 0.0f ? 1.0 : (double)powf(2.0f, exposure);
    // guide tables: the powers of two at or above the CDFs' lengths
    uint32_t Gm = 1, Gc = 1;
    while (Gm < h - 1)
        Gm <<= 1;
    while (Gc < w - 1)
        Gc <<= 1;
    double *d_band = nullptr, *d_scratch = nullptr;
    float *d_rgb = nullptr;
    int rc = MRT_OK;
    if (dalloc(c->d_env_tex, n) != hipSuccess
        || dalloc(c->d_env_pdf, n) != hipSuccess
        || dalloc(c->d_env_marg_inv, h) != hipSuccess
        || dalloc(c->d_env_cond_inv, n) != hipSuccess
        || dalloc(c->d_mis_marg, h) != hipSuccess
        || dalloc(c->d_mis_cond, (size_t)(h - 1) * w) != hipSuccess
        || dalloc(c->d_mis_pcell, (size_t)(h - 1) * (w - 1)) != hipSuccess
        || dalloc(c->d_mis_gm, (size_t)Gm + 1) != hipSuccess
        || dalloc(c->d_mis_gc, (size_t)(h - 1) * (Gc + 1)) != hipSuccess
        || dalloc(d_band, h - 1) != hipSuccess
        || dalloc(d_scratch, 3 * (size_t)h + 2 + n) != hipSuccess
        || dalloc(d_rgb, 3 * n) != hipSuccess)
        rc = fail(c, MRT_E_NOMEM, "device allocation for the HDRI failed");
    // the texels as given (rgb), widened to float4 on the device
    if (rc == MRT_OK)
        rc = upload(c, d_rgb, rgb, n * 12);
    if (rc == MRT_OK && launch_rgb_to_float4(c->stream, d_rgb, c->d_env_tex, n)
                            != hipSuccess)
        rc = fail(c, MRT_E_HIP, "HDRI texel copy failed");
    if (rc == MRT_OK)
        rc = upload(c, d_band, band.data(), (h - 1) * sizeof(double));
    if (rc == MRT_OK) {
        hipError_t e = launch_env_tables(c->stream, c->d_env_tex, w, h, sc, d_band,
                                         d_scratch, c->d_env_pdf, c->d_env_marg_inv,
                                         c->d_env_cond_inv, c->d_mis_marg,
                                         c->d_mis_cond, c->d_mis_pcell);
        if (e == hipSuccess)
            e = launch_env_guides(c->stream, c->d_mis_marg, c->d_mis_cond, w, h,
                                  Gm, Gc, c->d_mis_gm, c->d_mis_gc);
        if (e != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(c, MRT_E_HIP, "HDRI table build failed");
    }
    dfree(d_band);
    dfree(d_scratch);
    dfree(d_rgb);
    if (rc != MRT_OK) {
        free_env(c);
        return rc;
    }
    c->guide_m = Gm;
    c->guide_c = Gc;
    c->env_w = w;
    c->env_h = h;
    c->env_exposure = exposure;
    c->has_env = true;
    return MRT_OK;
}

extern "C" int mrt_get_scene_bvh(mrt_ctx *c, uint32_t *n_nodes, uint32_t *n_prims,
                                 uint32_t *depth, int *on_device, void *nodes,
                                 void *prims, float *shade_nrm)
{
    if (!c)
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "no scene uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    if (n_nodes)
        *n_nodes = c->n_nodes;
    if (n_prims)
        *n_prims = c->n_prims;
    if (depth)
        *depth = c->bvh_depth;
    if (on_device)
        *on_device = c->built_on_device ? 1 : 0;
    if (nodes)
        HIPCHK(c, hipMemcpyAsync(nodes, c->d_nodes, (size_t)c->n_nodes * 128,
                                 hipMemcpyDeviceToHost, c->stream));
    if (prims)
        HIPCHK(c, hipMemcpyAsync(prims, c->d_prims, (size_t)c->n_prims * 48,
                                 hipMemcpyDeviceToHost, c->stream));
    if (shade_nrm)
        HIPCHK(c, hipMemcpyAsync(shade_nrm, c->d_nrm, (size_t)c->n_prims * 36,
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_get_hdri_tables(mrt_ctx *c, float *pdf_joint, float *marg_inv,
                                   float *cond_inv, float *mis_marg,
                                   float *mis_cond, float *mis_pcell)
{
    if (!c)
        return MRT_E_INVALID;
    if (!c->has_env)
        return fail(c, MRT_E_STATE, "no HDRI uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t w = c->env_w, h = c->env_h, n = w * h;
    struct {
        float       *dst;
        const float *src;
        size_t       count;
    } cp[6] = {{pdf_joint, c->d_env_pdf, n},
               {marg_inv, c->d_env_marg_inv, h},
               {cond_inv, c->d_env_cond_inv, n},
               {mis_marg, c->d_mis_marg, h},
               {mis_cond, c->d_mis_cond, (h - 1) * w},
               {mis_pcell, c->d_mis_pcell, (h - 1) * (w - 1)}};
    for (int k = 0; k < 6; k++)
        if (cp[k].dst)
            HIPCHK(c, hipMemcpyAsync(cp[k].dst, cp[k].src, cp[k].count * 4,
                                     hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

static FrameParams frame_params(const mrt_ctx *c);
static SceneDev scene_dev(const mrt_ctx *c);

static TexDev tex_dev(const mrt_ctx *c)
{
    TexDev T;
    T.tex = c->d_tex;
    T.texels = c->d_texels;
    T.mtex = c->d_mtex;
    T.tri_uv = c->d_tri_uv;
    return T;
}

static ShadeBufs scene_bufs(const mrt_ctx *c)
{
    ShadeBufs B;
    memset(&B, 0, sizeof(B));
    B.planes = c->d_planes;
    B.prims = c->d_prims;
    B.mats = c->d_mats;
    B.lights = c->d_lights;
    B.nrm = c->d_nrm;
    B.tx = tex_dev(c);
    return B;
}

extern "C" int mrt_build_photons(mrt_ctx *c, uint32_t n_per, float radius)
{
    if (!c)
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "mrt_build_photons before mrt_upload_scene");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_photons(c);
    if (n_per == 0)
        return MRT_OK;
    if (!(radius > 0.0f) || !(c->ph_norm > 0.0f))
        return fail(c, MRT_E_INVALID, "caustic radius and norm must be positive");
    PhotonEmit pe;
    memset(&pe, 0, sizeof(pe));
    for (uint32_t k = 0; k < c->lights.size(); k++)
        if (c->lights[k].type == MRT_LIGHT_POINT) {
            if (pe.n_point == 64)
                return fail(c, MRT_E_LIMIT, "caustics support at most 64 point lights");
            pe.point_idx[pe.n_point++] = k;
        }
    pe.has_env = c->has_env ? 1u : 0u;
    uint64_t total = (uint64_t)(pe.n_point + pe.has_env) * n_per;
    if (total == 0)
        return fail(c, MRT_E_INVALID, "caustics need a point light or an HDRI");
    if (total > (1ull << 31))
        return fail(c, MRT_E_LIMIT, "more than 2^31 photons");
    pe.n_per = n_per;
    pe.n_total = (uint32_t)total;
    memcpy(pe.sc, c->scene_c, sizeof(pe.sc));
    pe.sr = c->scene_r;
    const uint32_t cap = pe.n_total;
    float4 *raw_pos = nullptr, *raw_pow = nullptr;
    uint32_t *cnt = nullptr;
    unsigned long long *k_in = nullptr, *k_out = nullptr;
    uint32_t *v_in = nullptr, *v_out = nullptr;
    void *tmp = nullptr;
    int rc = MRT_OK;
    uint32_t n = 0;
    hipError_t e = hipSuccess;
    do {
        if (dalloc(raw_pos, cap) != hipSuccess || dalloc(raw_pow, cap) != hipSuccess
            || dalloc(cnt, 1) != hipSuccess) {
            rc = fail(c, MRT_E_NOMEM, "photon buffers (%u) do not fit", cap);
            break;
        }
        FrameParams fp = frame_params(c);
        if ((e = launch_fill_u32(c->stream, cnt, 0u, 1)) != hipSuccess
            || (e = launch_photon_emit(c->stream, fp, scene_dev(c), pe,
                                       scene_bufs(c), raw_pos, raw_pow, cnt,
                                       cap)) != hipSuccess
            || (e = hipMemcpyAsync(&n, cnt, 4, hipMemcpyDeviceToHost, c->stream))
                   != hipSuccess
            || (e = hipStreamSynchronize(c->stream)) != hipSuccess)
            break;
        if (n > cap)
            n = cap;
        if (n == 0)
            break;
        const uint32_t table = 2 * n;        /* cells_am = 2 photons */
        const float cell = c->ph_k ? sqrtf(radius) : radius;
        if (dalloc(k_in, n) != hipSuccess || dalloc(k_out, n) != hipSuccess
            || dalloc(v_in, n) != hipSuccess || dalloc(v_out, n) != hipSuccess
            || dalloc(c->d_ph_rec, 2 * (size_t)n) != hipSuccess
            || dalloc(c->d_range, table) != hipSuccess
            || dalloc(c->d_occ, (table + 31) / 32) != hipSuccess) {
            rc = fail(c, MRT_E_NOMEM, "photon map (%u) does not fit", n);
            break;
        }
        unsigned end_bit = 32;
        while (end_bit < 64 && (1ull << (end_bit - 32)) < table)
            end_bit++;
        size_t tmp_bytes = 0;
        if ((e = launch_photon_keys(c->stream, raw_pos, n, cell, table, k_in,
                                    v_in)) != hipSuccess
            || (e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, k_in, k_out,
                                              v_in, v_out, n, 0, end_bit,
                                              c->stream)) != hipSuccess
            || (e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1)) != hipSuccess
            || (e = rocprim::radix_sort_pairs(tmp, tmp_bytes, k_in, k_out, v_in,
                                              v_out, n, 0, end_bit, c->stream))
                   != hipSuccess
            || (e = launch_fill_u32(c->stream, (uint32_t *)c->d_range, 0u,
                                    2 * (size_t)table)) != hipSuccess
            || (e = launch_fill_u32(c->stream, c->d_occ, 0u, (table + 31) / 32))
                   != hipSuccess
            || (e = launch_photon_reorder(c->stream, raw_pos, raw_pow, k_out,
                                          v_out, n, c->d_ph_rec,
                                          c->d_range, c->d_occ)) != hipSuccess
            || (e = hipStreamSynchronize(c->stream)) != hipSuccess)
            break;
        c->ph_n = n;
        c->ph_table = table;
        c->ph_cell = cell;
        c->ph_radius = radius;
        c->has_photons = true;
    } while (0);
    (void)hipStreamSynchronize(c->stream);
    dfree(raw_pos);
    dfree(raw_pow);
    dfree(cnt);
    dfree(k_in);
    dfree(k_out);
    dfree(v_in);
    dfree(v_out);
    if (tmp)
        (void)hipFree(tmp);
    if (e != hipSuccess && rc == MRT_OK)
        rc = fail(c, MRT_E_HIP, "photon map build failed: %s", hipGetErrorString(e));
    if (rc != MRT_OK || !c->has_photons) {
        free_photons(c);
        return rc;
    }
    return MRT_OK;
}

extern "C" int mrt_get_photons(mrt_ctx *c, uint32_t max, float *pos,
                               float *power, uint32_t *ids, uint32_t *n_out)
{
    if (!c || !n_out)
        return MRT_E_INVALID;
    *n_out = c->ph_n;
    uint32_t n = c->ph_n < max ? c->ph_n : max;
    if (n == 0)
        return MRT_OK;
    std::vector<float4> rec(2 * (size_t)n);
    HIPCHK(c, hipMemcpyAsync(rec.data(), c->d_ph_rec, (size_t)n * 32,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n; i++) {
        const float4 &P = rec[2 * (size_t)i], &W = rec[2 * (size_t)i + 1];
        if (pos) {
            pos[3 * i] = P.x;
            pos[3 * i + 1] = P.y;
            pos[3 * i + 2] = P.z;
        }
        if (power) {
            power[3 * i] = W.x;
            power[3 * i + 1] = W.y;
            power[3 * i + 2] = W.z;
        }
        if (ids) {
            uint32_t g;
            memcpy(&g, &P.w, 4);
            ids[i] = g;
        }
    }
    return MRT_OK;
}

static PhotonMapDev photon_map_dev(const mrt_ctx *c)
{
    PhotonMapDev pm;
    memset(&pm, 0, sizeof(pm));
    pm.rec = c->d_ph_rec;
    pm.range = c->d_range;
    pm.occ = c->d_occ;
    pm.n = c->ph_n;
    pm.table = c->ph_table;
    pm.cell = c->ph_cell;
    pm.k = c->ph_k;
    pm.radius = c->ph_radius;
    pm.norm = c->ph_norm;
    return pm;
}

extern "C" int mrt_caustic_at(mrt_ctx *c, uint32_t n, const float *pts,
                              float *rgb)
{
    if (!c || (n && (!pts || !rgb)))
        return MRT_E_INVALID;
    if (n == 0)
        return MRT_OK;
    if (!c->has_photons) {
        memset(rgb, 0, (size_t)n * 12);
        return MRT_OK;
    }
    HIPCHK(c, hipSetDevice(c->device));
    float *d_p = nullptr, *d_o = nullptr;
    hipError_t e = hipMalloc((void **)&d_p, (size_t)n * 12);
    if (e == hipSuccess)
        e = hipMalloc((void **)&d_o, (size_t)n * 12);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_p, pts, (size_t)n * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = launch_caustic_at(c->stream, photon_map_dev(c), n, d_p, d_o);
    if (e == hipSuccess)
        e = hipMemcpyAsync(rgb, d_o, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_p);
    (void)hipFree(d_o);
    if (e != hipSuccess)
        return fail(c, MRT_E_HIP, "caustic query failed: %s", hipGetErrorString(e));
    return MRT_OK;
}

// init_viewport_values / init_viewport_vectors (src/ray_tracing/viewport.c:
// 38-88) in double, rounded to float for the kernels.
static void camera_frame(FrameParams &fp, const mrt_camera &cam, uint32_t W,
                         uint32_t H, float ambient_ratio,
                         const float ambient_color[3])
{
    double dir[3] = {cam.dir[0], cam.dir[1], cam.dir[2]};
    double l = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (int k = 0; k < 3; k++)
        dir[k] = l > 0.0 ? dir[k] / l : 0.0;
    double focus = cam.focus_dist > 0.0f ? cam.focus_dist : 1.0;
    double width = 2.0 * tan((double)cam.fov_deg * (M_PI / 180.0) / 2.0) * focus;
    double height = width / ((double)W / (double)H);
    double up[3] = {0, 1, 0};
    if (fabs(dir[1]) > 0.999) {
        up[0] = 1;
        up[1] = 0;
    }
    double nd[3] = {-dir[0], -dir[1], -dir[2]};
    double u[3] = {up[1] * nd[2] - up[2] * nd[1], up[2] * nd[0] - up[0] * nd[2],
                   up[0] * nd[1] - up[1] * nd[0]};
    l = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int k = 0; k < 3; k++)
        u[k] = l > 0.0 ? u[k] / l : 0.0;
    double cv[3] = {dir[1] * u[2] - dir[2] * u[1], dir[2] * u[0] - dir[0] * u[2],
                    dir[0] * u[1] - dir[1] * u[0]};
    double rad = focus * tan((double)cam.defocus_angle_deg * M_PI / 360.0);
    for (int k = 0; k < 3; k++) {
        double vu = u[k] * width, vv = cv[k] * height;
        double du = vu / (double)W, dv = vv / (double)H;
        double ul = ((double)cam.pos[k] + dir[k] * focus) - vu / 2 - vv / 2;
        fp.cam_o[k] = cam.pos[k];
        fp.p00[k] = (float)(ul + (du + dv) * 0.5);
        fp.du[k] = (float)du;
        fp.dv[k] = (float)dv;
        fp.disk_u[k] = (float)(u[k] * rad);
        fp.disk_v[k] = (float)(cv[k] * rad);
        fp.cam_pos[k] = cam.pos[k];
        double bg = ambient_color[k] / 255.0 * (double)ambient_ratio;
        float bgf = (float)bg;
        fp.bg_clamped[k] = bgf < 0.0f ? 0.0f : (bgf > 16.0f ? 16.0f : bgf);
    }
    fp.lens = rad > 0.0;
    fp.W = W;
    fp.H = H;
    fp.WH = W * H;
}

static FrameParams frame_params(const mrt_ctx *c)
{
    FrameParams fp;
    memset(&fp, 0, sizeof(fp));
    camera_frame(fp, c->cam, c->W, c->H, c->ambient_ratio, c->ambient_color);
    fp.ior_all = c->ior_all;
    fp.n_lights = (uint32_t)c->lights.size();
    fp.one_sided = (c->flags & MRT_SCENE_TRIANGLE_ONE_SIDED) ? 1u : 0u;
    fp.flags = c->flags;
    fp.seed = c->seed;
    fp.estimator = (uint32_t)c->estimator;
    if (c->has_env) {
        fp.has_env = 1;
        fp.env_tex = c->d_env_tex;
        fp.env_pdf = c->d_env_pdf;
        fp.env_marg_inv = c->d_env_marg_inv;
        fp.env_cond_inv = c->d_env_cond_inv;
        fp.mis_marg = c->d_mis_marg;
        fp.mis_cond = c->d_mis_cond;
        fp.mis_pcell = c->d_mis_pcell;
        fp.mis_marg_guide = c->d_mis_gm;
        fp.mis_cond_guide = c->d_mis_gc;
        fp.guide_m = c->guide_m;
        fp.guide_c = c->guide_c;
        fp.env_w = c->env_w;
        fp.env_h = c->env_h;
        float ex = c->env_exposure + c->ambient_ratio - 1.0f;
        fp.env_scale = ex == 0.0f ? 1.0f : powf(2.0f, ex);
    }
    if (c->has_photons) {
        fp.has_photons = 1;
        fp.pm = photon_map_dev(c);
    }
    fp.textured = c->textured ? 1u : 0u;
    return fp;
}

static SceneDev scene_dev(const mrt_ctx *c)
{
    SceneDev S;
    memset(&S, 0, sizeof(S));
    S.nodes = c->d_nodes;
    S.prims = c->d_prims;
    S.aux = c->d_nrm;
    S.planes = c->d_planes;
    S.n_nodes = c->n_nodes;
    S.n_planes = c->n_planes;
    S.one_sided = (c->flags & MRT_SCENE_TRIANGLE_ONE_SIDED) ? 1u : 0u;
    S.root = 0;
    S.quad = c->has_quad ? 1u : 0u;
    S.alpha = c->alpha ? 1u : 0u;
    S.tx = tex_dev(c);
    return S;
}

static uint32_t shadow_rays_per_path(const mrt_ctx *c)
{
    uint32_t n = 0;
    for (const mrt_light &l : c->lights) {
        bool soft = l.type == MRT_LIGHT_POINT && l.radius != 0.0f
                    && l.shadow_samples > 1;
        n += soft ? (uint32_t)l.shadow_samples : 1u;
    }
    if (c->has_env)
        n += 1;                        /* the environment ray */
    return n ? n : 1u;
}

static int ensure_wavefront(mrt_ctx *c, uint32_t paths, uint32_t shadow)
{
    if (paths <= c->cap_paths && shadow <= c->cap_shadow)
        return MRT_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_wavefront(c);
    bool ok = true;
    for (int k = 0; k < 2; k++)
        ok = ok && dalloc(c->q_o[k], paths) == hipSuccess
             && dalloc(c->q_d[k], paths) == hipSuccess
             && dalloc(c->q_t[k], paths) == hipSuccess;
    ok = ok && dalloc(c->hit, paths) == hipSuccess
         && dalloc(c->acc, paths) == hipSuccess
         && dalloc(c->p_a, paths) == hipSuccess
         && dalloc(c->p_b, paths) == hipSuccess
         && dalloc(c->p_c, paths) == hipSuccess
         && dalloc(c->p_x, paths) == hipSuccess
         && dalloc(c->p_k, paths) == hipSuccess
         && dalloc(c->s_o, shadow) == hipSuccess
         && dalloc(c->s_d, shadow) == hipSuccess
         && dalloc(c->s_c, shadow) == hipSuccess
         && dalloc(c->vis, shadow) == hipSuccess;
    if (!ok) {
        free_wavefront(c);
        return fail(c, MRT_E_NOMEM, "wavefront buffers (%u paths) do not fit",
                    paths);
    }
    c->cap_paths = paths;
    c->cap_shadow = shadow;
    return MRT_OK;
}

namespace {
enum { K_RAYGEN = 0, K_TRACE, K_SHADE, K_SHADOW, K_ACCUM, K_RESOLVE, K_CAUSTIC,
       K_ADVANCE };
struct EvPair {
    int         cls;
    hipEvent_t  a, b;
};
// Per-launch HIP events, recorded on the stream the launches go to (st;
// the context's stream unless a render lane sets its own).
struct Timer {
    mrt_ctx              *c;
    bool                  on;
    std::vector<EvPair>   ev;
    hipStream_t           st = nullptr;
    hipStream_t stream() const { return st ? st : c->stream; }
    void begin(int cls)
    {
        if (!on)
            return;
        EvPair p{cls, nullptr, nullptr};
        (void)hipEventCreate(&p.a);
        (void)hipEventCreate(&p.b);
        (void)hipEventRecord(p.a, stream());
        ev.push_back(p);
    }
    void end()
    {
        if (on && !ev.empty())
            (void)hipEventRecord(ev.back().b, stream());
    }
    void collect()
    {
        if (!on)
            return;
        (void)hipStreamSynchronize(c->stream);
        for (EvPair &p : ev) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                int slot = (p.cls == K_RESOLVE || p.cls == K_ADVANCE) ? K_SHADE
                           : (p.cls == K_CAUSTIC ? 6 : p.cls);
                c->ms[slot] += ms;
            }
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        ev.clear();
    }
};
}  // namespace

#define LAUNCH(ctx, tm, cls, expr)                                           \
    do {                                                                     \
        (ctx)->launches[cls]++;                                              \
        (tm).begin(cls);                                                     \
        hipError_t e_ = (expr);                                              \
        (tm).end();                                                          \
        if (e_ != hipSuccess) {                                              \
            (tm).collect();                                                  \
            return fail((ctx), MRT_E_HIP, "kernel launch failed: %s",        \
                        hipGetErrorString(e_));                              \
        }                                                                    \
    } while (0)

// One render lane: P paths of samples fp.sample0.. on stream st, using the
// queue entries from path `off` (shadow entries from off * rpp) and the
// counter block ctr.  Bounce b: its closest-hit rays and the shadow rays of
// bounce b - 1 in one launch, then the light sums of bounce b - 1, then
// the shade of bounce b; a last launch takes the shadow rays of the last
// bounce.  start() enqueues the raygen, bounce(b) bounce b's launches, so
// a caller can interleave the lanes' launches bounce by bounce.
struct LaneRun {
    mrt_ctx *c;
    hipStream_t st;
    Timer *tm;
    FrameParams fp;
    const SceneDev *S;
    bool count;
    uint32_t max_depth, P, off, rpp;
    uint32_t *ctr;
    StackOvf ovf;
    float4 *tape = nullptr;
    hipEvent_t ev_after = nullptr;   // recorded once bounce ev_bounce is traced
    uint32_t ev_bounce = 0;
    // set by start()
    size_t so = 0;
    uint32_t cap_shadow = 0, gb = 0;
    bool lit = false;
    ShadeBufs Bp;                    // the previous bounce's shade, for its light sums
    FrameParams fpp;

    int start()
    {
        so = (size_t)off * rpp;
        cap_shadow = P * rpp;
        lit = !c->lights.empty() || c->has_env || c->has_photons;
        gb = std::min((P + SHADE_BLOCK - 1) / SHADE_BLOCK, c->grid_blocks);
        memset(&Bp, 0, sizeof(Bp));
        fpp = fp;
        tm->st = st;
        LAUNCH(c, (*tm), K_RAYGEN,
               launch_raygen(st, fp, c->q_o[0] + off, c->q_d[0] + off,
                             c->q_t[0] + off, c->acc + off, P, ctr));
        tm->st = nullptr;
        return MRT_OK;
    }

    int bounce(uint32_t b)
    {
        tm->st = st;
        const int rc = bounce_launches(b);
        tm->st = nullptr;
        return rc;
    }

private:
    int bounce_launches(uint32_t b)
    {
        const bool last = b == max_depth;
        const int cur = b & 1, nxt = cur ^ 1;
        TraceWork w;
        w.q_o = c->q_o[cur] + off;
        w.q_d = c->q_d[cur] + off;
        w.q_t = c->q_t[cur] + off;
        w.hit = c->hit + off;
        w.s_o = c->s_o + so;
        w.s_d = c->s_d + so;
        w.vis = c->vis + so;
        w.cap_shadow = cap_shadow;
        w.closest = !last;
        w.shadow_ctr = (b > 0 && lit) ? ctr_shadow(b - 1) : -1;
        if (w.closest || w.shadow_ctr >= 0)
            LAUNCH(c, (*tm), K_TRACE,
                   launch_trace(st, c->trace_blocks, count, *S, ovf, w, ctr,
                                c->d_stats, c->trace_top, c->trace_refill,
                                c->refill_blocks));
        // the next pass may start (ev_after) once this bounce is traced
        if (ev_after && (b == ev_bounce || last)) {
            HIPCHK(c, hipEventRecord(ev_after, st));
            ev_after = nullptr;
        }
        if (b > 0 && lit)
            LAUNCH(c, (*tm), K_RESOLVE, launch_resolve(st, gb, fpp, Bp, ctr));
        if (last) {
            LAUNCH(c, (*tm), K_ADVANCE, launch_advance(st, ctr, c->d_stats, P, b));
            return MRT_OK;
        }
        fp.bounce = b;
        fp.last_bounce = (b + 1 == max_depth) ? 1u : 0u;
        ShadeBufs B;
        memset(&B, 0, sizeof(B));
        B.q_o = c->q_o[cur] + off;
        B.q_d = c->q_d[cur] + off;
        B.q_t = c->q_t[cur] + off;
        B.hit = c->hit + off;
        B.qn_o = c->q_o[nxt] + off;
        B.qn_d = c->q_d[nxt] + off;
        B.qn_t = c->q_t[nxt] + off;
        B.acc = c->acc + off;
        B.planes = c->d_planes;
        B.prims = c->d_prims;
        B.mats = c->d_mats;
        B.lights = c->d_lights;
        B.nrm = c->d_nrm;
        B.tx = tex_dev(c);
        B.s_o = c->s_o + so;
        B.s_d = c->s_d + so;
        B.s_c = c->s_c + so;
        B.vis = c->vis + so;
        B.p_a = c->p_a + off;
        B.p_b = c->p_b + off;
        B.p_c = c->p_c + off;
        B.p_x = c->p_x + off;
        B.p_k = c->p_k + off;
        B.cap_paths = P;
        B.cap_shadow = cap_shadow;
        B.tape = tape;
        B.tape_P = P;
        LAUNCH(c, (*tm), K_SHADE, launch_shade(st, gb, fp, B, ctr));
        if (c->has_photons)
            LAUNCH(c, (*tm), K_CAUSTIC,
                   launch_caustic(st, c->grid_blocks, fp, B, ctr, count,
                                  c->d_stats));
        LAUNCH(c, (*tm), K_ADVANCE, launch_advance(st, ctr, c->d_stats, P, b));
        Bp = B;
        fpp = fp;
        return MRT_OK;
    }
};

// A whole lane, bounce after bounce.
static int render_lane(mrt_ctx *c, hipStream_t st, Timer &tm, FrameParams fp,
                       const SceneDev &S, bool count, uint32_t max_depth,
                       uint32_t P, uint32_t off, uint32_t rpp, uint32_t *ctr,
                       const StackOvf &ovf, float4 *tape = nullptr,
                       hipEvent_t ev_after = nullptr, uint32_t ev_bounce = 0)
{
    LaneRun lr{c, st, &tm, fp, &S, count, max_depth, P, off, rpp, ctr, ovf,
               tape, ev_after, ev_bounce};
    int rc = lr.start();
    for (uint32_t b = 0; rc == MRT_OK && b <= max_depth; b++)
        rc = lr.bounce(b);
    return rc;
}

// Lanes of a render of n_samples in passes of nb samples (P64 paths, S64
// shadow-ray slots each): one per pass in flight, at most MRT_OPT_LANES,
// each with a pass's worth of queues; 0 when one pass's shadow queue
// alone is beyond 2^32 entries.
static uint32_t render_lanes(const mrt_ctx *c, uint32_t n_samples, uint32_t nb,
                             uint64_t P64, uint64_t S64)
{
    const uint32_t npass = (n_samples + nb - 1) / nb;
    uint32_t L = std::max(1u, std::min(c->lanes, npass));
    while (L > 1 && (S64 * L > 0xFFFFFFF0ull || P64 * L > 0xFFFFFFF0ull))
        L--;
    return S64 > 0xFFFFFFF0ull ? 0u : L;
}

extern "C" int mrt_reserve(mrt_ctx *c, uint32_t n_samples)
{
    if (!c)
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "mrt_reserve before mrt_upload_scene");
    if (n_samples == 0)
        return MRT_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const uint32_t WH = c->W * c->H;
    uint32_t nb = std::max(1u, c->path_budget / WH);
    nb = std::min(nb, n_samples);
    const uint64_t P64 = (uint64_t)nb * WH;
    const uint64_t S64 = P64 * shadow_rays_per_path(c);
    const uint32_t L = render_lanes(c, n_samples, nb, P64, S64);
    if (!L)
        return fail(c, MRT_E_LIMIT, "shadow queue exceeds 2^32 entries");
    int rc = ensure_wavefront(c, (uint32_t)(P64 * L), (uint32_t)(S64 * L));
    if (rc != MRT_OK)
        return rc;
    if (ensure_ovf(c, c->trace_blocks, L) != hipSuccess)
        return fail(c, MRT_E_NOMEM, "trace stack buffer does not fit");
    return MRT_OK;
}

extern "C" int mrt_render(mrt_ctx *c, uint32_t first_sample, uint32_t n_samples,
                          uint32_t max_depth)
{
    if (!c)
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "mrt_render before mrt_upload_scene");
    if (max_depth > 4095)
        return fail(c, MRT_E_LIMIT, "max_depth above 4095");
    if (n_samples == 0 || max_depth == 0)
        return MRT_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const uint32_t WH = c->W * c->H;
    uint32_t nb = c->path_budget / WH;
    if (nb < 1)
        nb = 1;
    if (nb > n_samples)
        nb = n_samples;
    const uint64_t P64 = (uint64_t)nb * WH;
    const uint32_t rpp = shadow_rays_per_path(c);
    const uint64_t S64 = P64 * rpp;
    // Lanes: the render's passes go round-robin to up to MRT_OPT_LANES
    // streams, each with a pass's worth of queues, counters and stack
    // columns.  A pass starts once the one before it has traced bounce
    // LANE_STAGGER_BOUNCE, so the launches of one fill the GPU while the
    // other's late bounces wait for their longest rays.  Each pass adds its
    // sums after the pass before it (ev_acc), so the image is the same sum
    // in the same order whatever the lanes.
    const uint32_t L = render_lanes(c, n_samples, nb, P64, S64);
    if (!L)
        return fail(c, MRT_E_LIMIT, "shadow queue exceeds 2^32 entries");
    // A render of one pass (a drop-in tick of 1 spp) splits that pass into
    // contiguous parts traced at once on the lanes: each bounce launch is
    // short, and the parts' launches overlap each other's slowest rays.
    const uint32_t Ls = n_samples <= nb
        ? std::max(1u, std::min(c->lanes, n_samples * WH / LANE_MIN_PATHS)) : 1u;
    int rc = ensure_wavefront(c, (uint32_t)(P64 * L), (uint32_t)(S64 * L));
    if (rc != MRT_OK)
        return rc;
    if (ensure_ovf(c, c->trace_blocks, std::max(L, Ls)) != hipSuccess)
        return fail(c, MRT_E_NOMEM, "trace stack buffer does not fit");
    FrameParams fp = frame_params(c);
    SceneDev S = scene_dev(c);
    const bool count = (c->stats_on & MRT_STATS_COUNT) != 0;
    const bool timing = (c->stats_on & MRT_STATS_TIME) != 0;
    Timer tm{c, timing, {}};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing) {
        (void)hipEventCreate(&t0);
        (void)hipEventCreate(&t1);
        (void)hipEventRecord(t0, c->stream);
    }
    if (L > 1 || Ls > 1) {
        // every lane after what the context's stream holds (uploads, clears)
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        for (uint32_t l = 1; l < std::max(L, Ls); l++)
            HIPCHK(c, hipStreamWaitEvent(c->lane_stream[l], c->ev_fork, 0));
    }
    if (Ls > 1) {
        // part l: slots [o0, o0 + Pl) of the pass, numbered as in the whole
        // pass (fp.slot0), in the pass's queue entries from o0, with lane
        // l's counters and stack columns; the pass's sums after every part
        // The parts' launches are enqueued bounce by bounce, part after
        // part: a tick starts with the GPU idle, and a part enqueued whole
        // would hold the next part back by its ~35 launches' host time.
        const uint32_t P = n_samples * WH;
        FrameParams fl = fp;
        fl.sample0 = first_sample;
        LaneRun parts[MAX_LANES];
        for (uint32_t p = 0; p < Ls; p++) {
            const uint32_t o0 = (uint32_t)((uint64_t)P * p / Ls);
            const uint32_t Pl = (uint32_t)((uint64_t)P * (p + 1) / Ls) - o0;
            fl.slot0 = o0;
            parts[p] = LaneRun{c, p ? c->lane_stream[p] : c->stream, &tm, fl, &S,
                               count, max_depth, Pl, o0, rpp,
                               c->d_ctr + (size_t)p * CTR_COUNT, stack_ovf(c, p)};
            rc = parts[p].start();
            if (rc != MRT_OK)
                return rc;
        }
        for (uint32_t b = 0; b <= max_depth; b++)
            for (uint32_t p = 0; p < Ls; p++) {
                rc = parts[p].bounce(b);
                if (rc != MRT_OK)
                    return rc;
            }
        for (uint32_t p = 1; p < Ls; p++) {
            HIPCHK(c, hipEventRecord(c->ev_join[p], parts[p].st));
            HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join[p], 0));
        }
        LAUNCH(c, tm, K_ACCUM,
               launch_accumulate(c->stream, c->acc, c->d_accum, WH, n_samples,
                                 c->d_stats));
    }
    uint32_t l = 0;
    for (uint32_t s = 0, k = 0; Ls == 1 && s < n_samples; s += nb, k++) {
        const uint32_t n = (n_samples - s) < nb ? (n_samples - s) : nb;
        const uint32_t P = n * WH;
        const uint32_t prev = l;
        l = k % L;
        hipStream_t st = l ? c->lane_stream[l] : c->stream;
        FrameParams fl = fp;
        fl.sample0 = first_sample + s;
        fl.slot0 = 0;
        if (k && L > 1)
            HIPCHK(c, hipStreamWaitEvent(st, c->ev_stag[prev], 0));
        const uint32_t off = (uint32_t)(P64 * l);
        rc = render_lane(c, st, tm, fl, S, count, max_depth, P, off, rpp,
                         c->d_ctr + (size_t)l * CTR_COUNT, stack_ovf(c, l),
                         nullptr, L > 1 ? c->ev_stag[l] : nullptr,
                         LANE_STAGGER_BOUNCE);
        if (rc != MRT_OK)
            return rc;
        if (k && L > 1)
            HIPCHK(c, hipStreamWaitEvent(st, c->ev_acc[prev], 0));
        tm.st = st;
        LAUNCH(c, tm, K_ACCUM,
               launch_accumulate(st, c->acc + off, c->d_accum, WH, n,
                                 c->d_stats));
        tm.st = nullptr;
        if (L > 1)
            HIPCHK(c, hipEventRecord(c->ev_acc[l], st));
    }
    // the context's stream after the last pass's sums (which follow every
    // pass before it)
    if (l)
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_acc[l], 0));
    c->paths += (uint64_t)n_samples * WH;
    if (timing) {
        (void)hipEventRecord(t1, c->stream);
        tm.collect();
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t0, t1) == hipSuccess)
            c->ms[5] += ms;
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    }
    return MRT_OK;
}

extern "C" int mrt_synchronize(mrt_ctx *c)
{
    if (!c)
        return MRT_E_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_readback_f32(mrt_ctx *c, float *out)
{
    if (!c || !out)
        return MRT_E_INVALID;
    size_t n = (size_t)c->W * c->H * 3;
    HIPCHK(c, hipMemcpyAsync(out, c->d_accum, n * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_readback_add(mrt_ctx *c, double *float_render)
{
    if (!c || !float_render)
        return MRT_E_INVALID;
    const size_t n = (size_t)c->W * c->H * 3;
    // set-up at the first call; each part is made once, so a call after a
    // failed set-up picks up where it stopped
    if (!c->h_stage && hipHostMalloc((void **)&c->h_stage, n * sizeof(float),
                                     hipHostMallocDefault) != hipSuccess) {
        c->h_stage = nullptr;
        return fail(c, MRT_E_NOMEM, "pinned read-back buffer (%zu floats)", n);
    }
    for (hipEvent_t &e : c->ev_rb)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            e = nullptr;
            return fail(c, MRT_E_HIP, "read-back events");
        }
    if (!c->pool) {
        // a 1080p frame is 6.2 M doubles: the add runs on up to 8 threads
        unsigned nt = std::thread::hardware_concurrency();
        nt = std::max(1u, std::min(nt, 8u));
        if (n < (1u << 20))
            nt = 1;
        HostPool *pool = new (std::nothrow) HostPool();
        if (!pool)
            return fail(c, MRT_E_NOMEM, "read-back threads");
        try {
            pool->start(nt);
        } catch (...) {
            // the threads that did start are joined by the destructor
            delete pool;
            return fail(c, MRT_E_NOMEM, "read-back threads");
        }
        c->pool = pool;
    }
    // get_result adds one dispatch's samples per call: the device sums are
    // handed over and zeroed in stream order, so the next render starts
    // from zero and every sample reaches float_render exactly once.  The
    // copy goes in chunks, and each chunk is added while the next ones
    // cross PCIe.
    const int K = n >= (1u << 20) ? READBACK_CHUNKS : 1;
    for (int k = 0; k < K; k++) {
        const size_t a = n * k / K, b = n * (k + 1) / K;
        HIPCHK(c, hipMemcpyAsync(c->h_stage + a, c->d_accum + a,
                                 (b - a) * sizeof(float), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipEventRecord(c->ev_rb[k], c->stream));
    }
    HIPCHK(c, hipMemsetAsync(c->d_accum, 0, n * sizeof(float), c->stream));
    const float *src = c->h_stage;
    const unsigned T = c->pool->size();
    std::atomic<int> bad{0};
    c->pool->run([&](unsigned t) {
        for (int k = 0; k < K; k++) {
            // polled, not a blocking wait: a tick's chunks arrive within a
            // millisecond, and a blocked thread's wake-up costs a good part
            // of that
            hipError_t e;
            while ((e = hipEventQuery(c->ev_rb[k])) == hipErrorNotReady)
                std::this_thread::yield();
            if (e != hipSuccess) {
                bad = 1;
                return;
            }
            const size_t a = n * k / K, len = n * (k + 1) / K - a;
            const size_t lo = a + len * t / T, hi = a + len * (t + 1) / T;
            for (size_t i = lo; i < hi; i++)
                float_render[i] += (double)src[i];
        }
    });
    if (bad)
        return fail(c, MRT_E_HIP, "read-back copy failed");
    return MRT_OK;
}

extern "C" int mrt_clear(mrt_ctx *c)
{
    if (!c)
        return MRT_E_INVALID;
    size_t n = (size_t)c->W * c->H * 3;
    HIPCHK(c, hipMemsetAsync(c->d_accum, 0, n * sizeof(float), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_render_aov(mrt_ctx *c, int aov, uint32_t sample, void *out)
{
    if (!c || !out || aov < MRT_AOV_PRIM_ID || aov > MRT_AOV_PRIMS)
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "mrt_render_aov before mrt_upload_scene");
    HIPCHK(c, hipSetDevice(c->device));
    size_t WH = (size_t)c->W * c->H;
    size_t bytes = WH * (aov == MRT_AOV_NORMAL ? 12 : 4);
    void *d = nullptr;
    HIPCHK(c, hipMalloc(&d, bytes));
    FrameParams fp = frame_params(c);
    SceneDev S = scene_dev(c);
    hipError_t e = launch_aov(c->stream, fp, S, c->d_prims, c->d_nrm, sample,
                              aov, d);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return fail(c, MRT_E_HIP, "AOV render failed: %s", hipGetErrorString(e));
    return MRT_OK;
}

extern "C" int mrt_trace_rays(mrt_ctx *c, uint32_t n, const float *orig,
                              const float *dir, uint32_t *prim_id, float *t)
{
    if (!c || (n && (!orig || !dir || !prim_id || !t)))
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "mrt_trace_rays before mrt_upload_scene");
    if (n == 0)
        return MRT_OK;
    HIPCHK(c, hipSetDevice(c->device));
    float *d_o = nullptr, *d_d = nullptr, *d_t = nullptr;
    uint32_t *d_p = nullptr;
    size_t v = (size_t)n * 3 * sizeof(float);
    hipError_t e = hipMalloc((void **)&d_o, v);
    if (e == hipSuccess)
        e = hipMalloc((void **)&d_d, v);
    if (e == hipSuccess)
        e = hipMalloc((void **)&d_t, (size_t)n * 4);
    if (e == hipSuccess)
        e = hipMalloc((void **)&d_p, (size_t)n * 4);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_o, orig, v, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_d, dir, v, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = launch_trace_rays(c->stream, scene_dev(c), c->d_prims, n, d_o, d_d,
                              d_p, d_t);
    if (e == hipSuccess)
        e = hipMemcpyAsync(prim_id, d_p, (size_t)n * 4, hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_o);
    (void)hipFree(d_d);
    (void)hipFree(d_t);
    (void)hipFree(d_p);
    if (e != hipSuccess)
        return fail(c, MRT_E_HIP, "trace_rays failed: %s", hipGetErrorString(e));
    return MRT_OK;
}

extern "C" int mrt_trace_rays_noalloc(mrt_ctx *c, uint32_t n,
                                      const float *orig, const float *dir,
                                      uint32_t *prim_id, float *t)
{
    if (!c || (n && (!orig || !dir || !prim_id || !t)))
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE,
                    "mrt_trace_rays_noalloc before mrt_upload_scene");
    if (n > MRT_TRACE_NOALLOC_MAX)
        return fail(c, MRT_E_LIMIT,
                    "mrt_trace_rays_noalloc takes at most %u rays (got %u)",
                    MRT_TRACE_NOALLOC_MAX, n);
    if (n == 0)
        return MRT_OK;
    if (!c->scratch_dev)
        return fail(c, MRT_E_STATE, "no staging address for this device");
    HIPCHK(c, hipSetDevice(c->device));
    // fixed partition of the staging array: every region is written (host
    // copies, then the kernel for each of the n rays) before it is read
    const size_t M = MRT_TRACE_NOALLOC_MAX;
    float *d_o = c->scratch_dev;
    float *d_d = c->scratch_dev + 3 * M;
    uint32_t *d_p = reinterpret_cast<uint32_t *>(c->scratch_dev + 6 * M);
    float *d_t = c->scratch_dev + 7 * M;
    const size_t v = (size_t)n * 3 * sizeof(float);
    HIPCHK(c, hipMemcpyAsync(d_o, orig, v, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_d, dir, v, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_trace_rays(c->stream, scene_dev(c), c->d_prims, n, d_o,
                                d_d, d_p, d_t));
    HIPCHK(c, hipMemcpyAsync(prim_id, d_p, (size_t)n * 4,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_set_seed(mrt_ctx *c, uint64_t seed)
{
    if (!c)
        return MRT_E_INVALID;
    c->seed = seed;
    return MRT_OK;
}

extern "C" int mrt_set_option(mrt_ctx *c, int option, double value)
{
    if (!c)
        return MRT_E_INVALID;
    switch (option) {
    case MRT_OPT_SEED:
        // a double holds integers exactly only up to 2^53: larger keys go
        // through mrt_set_seed, so that no key is silently rounded
        if (!(value >= 0.0) || value > 9007199254740992.0
            || value != floor(value))
            return fail(c, MRT_E_INVALID,
                        "seed must be an integer in [0, 2^53] here; use "
                        "mrt_set_seed for a full 64-bit key");
        c->seed = (uint64_t)value;
        return MRT_OK;
    case MRT_OPT_ESTIMATOR:
        if (value != MRT_EST_REFERENCE && value != MRT_EST_MIS)
            return fail(c, MRT_E_INVALID, "unknown estimator");
        c->estimator = (int)value;
        return MRT_OK;
    case MRT_OPT_STATS:
        if (!(value >= 0.0) || value > 3.0)
            return fail(c, MRT_E_INVALID, "stats flags out of range");
        c->stats_on = (int)value;
        return MRT_OK;
    case MRT_OPT_PATH_BUDGET:
        if (!(value >= 1.0) || value > 1.0e9)
            return fail(c, MRT_E_INVALID, "path budget out of range");
        c->path_budget = (uint32_t)value;
        return MRT_OK;
    case MRT_OPT_TRACE_REFILL:
        if (value != 0.0 && value != 1.0)
            return fail(c, MRT_E_INVALID, "trace refill is 0 or 1");
        c->trace_refill = value != 0.0;
        return MRT_OK;
    case MRT_OPT_TRACE_TOP:
        if (value != 0.0 && value != 1.0)
            return fail(c, MRT_E_INVALID, "trace top cache is 0 or 1");
        c->trace_top = value != 0.0;
        return MRT_OK;
    case MRT_OPT_BLOCK:
        return fail(c, MRT_E_UNSUPPORTED, "the trace block size is fixed when the library is compiled");
    case MRT_OPT_PHOTON_K:
        if (!(value >= 0.0) || value > 32.0)
            return fail(c, MRT_E_LIMIT, "photon k must be in 0..32");
        c->ph_k = (uint32_t)value;
        return MRT_OK;
    case MRT_OPT_PHOTON_NORM:
        if (!(value > 0.0))
            return fail(c, MRT_E_INVALID, "photon norm must be positive");
        c->ph_norm = (float)value;
        return MRT_OK;
    case MRT_OPT_BUILDER:
        if (value != MRT_BUILDER_HOST && value != MRT_BUILDER_DEVICE)
            return fail(c, MRT_E_INVALID, "unknown BVH builder");
        c->builder = (int)value;
        return MRT_OK;
    case MRT_OPT_LANES:
        if (!(value >= 1.0) || value > (double)MAX_LANES || value != floor(value))
            return fail(c, MRT_E_INVALID, "render lanes must be 1..%d", MAX_LANES);
        c->lanes = (uint32_t)value;
        return MRT_OK;
    default:
        return fail(c, MRT_E_INVALID, "unknown option %d", option);
    }
}

extern "C" int mrt_get_stats(mrt_ctx *c, mrt_stats *out)
{
    if (!c || !out)
        return MRT_E_INVALID;
    unsigned long long st[STAT_COUNT];
    HIPCHK(c, hipMemcpyAsync(st, c->d_stats, sizeof(st), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof(*out));
    out->paths = c->paths;
    out->rays_closest = st[STAT_RAYS_CLOSEST];
    out->rays_shadow = st[STAT_RAYS_SHADOW];
    out->nodes_closest = st[STAT_NODES_CLOSEST];
    out->prims_closest = st[STAT_PRIMS_CLOSEST];
    out->nodes_shadow = st[STAT_NODES_SHADOW];
    out->prims_shadow = st[STAT_PRIMS_SHADOW];
    out->nonfinite = st[STAT_NONFINITE];
    out->launches_trace = c->launches[K_TRACE];
    out->launches_shadow = c->launches[K_SHADOW];
    out->launches_shade = c->launches[K_SHADE];
    out->ms_raygen = c->ms[K_RAYGEN];
    out->ms_trace = c->ms[K_TRACE];
    out->ms_shade = c->ms[K_SHADE];
    out->ms_shadow = c->ms[K_SHADOW];
    out->ms_accum = c->ms[K_ACCUM];
    out->ms_total = c->ms[5];
    out->ms_caustic = c->ms[6];
    out->gather_queries = st[STAT_GATHER_QUERIES];
    out->gather_photons = st[STAT_GATHER_PHOTONS];
    out->gather_cells = st[STAT_GATHER_CELLS];
    out->gather_lookups = st[STAT_GATHER_LOOKUPS];
    return MRT_OK;
}

extern "C" int mrt_reset_stats(mrt_ctx *c)
{
    if (!c)
        return MRT_E_INVALID;
    HIPCHK(c, hipMemsetAsync(c->d_stats, 0, STAT_COUNT * 8, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->paths = 0;
    for (double &m : c->ms)
        m = 0.0;
    for (uint64_t &l : c->launches)
        l = 0;
    return MRT_OK;
}

// ---- differentiable rendering -------------------------------------------------
extern "C" int mrt_diff_set_materials(mrt_ctx *c, uint32_t n, const float *albedo,
                                      const float *roughness, const float *emission)
{
    if (!c)
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "no scene uploaded");
    if (n != c->n_mats)
        return fail(c, MRT_E_INVALID, "%u materials given, the scene has %u", n,
                    c->n_mats);
    for (uint32_t k = 0; k < n; k++) {
        mrt_material &m = c->mats_host[k];
        if (albedo)
            for (int j = 0; j < 3; j++)
                m.color[j] = albedo[3 * k + j];
        if (roughness)
            m.roughness = roughness[k];
        if (emission) {
            m.emission_strength = 1.0f;
            for (int j = 0; j < 3; j++)
                m.emission_color[j] = emission[3 * k + j];
        }
    }
    std::vector<float4> mats;
    for (const mrt_material &m : c->mats_host) {
        mats.push_back(make_float4(m.color[0], m.color[1], m.color[2], m.metallic));
        mats.push_back(make_float4(m.roughness, m.ior, m.transmission, m.ao));
        mats.push_back(make_float4(m.emission_strength, m.emission_color[0],
                                   m.emission_color[1], m.emission_color[2]));
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_mats, mats.data(), mats.size() * 16,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_diff_set_hdri(mrt_ctx *c, const float *rgb)
{
    if (!c || !rgb)
        return MRT_E_INVALID;
    if (!c->has_env)
        return fail(c, MRT_E_STATE, "no HDRI uploaded");
    const size_t n = (size_t)c->env_w * c->env_h;
    std::vector<float4> tex(n);
    for (size_t i = 0; i < n; i++)
        tex[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.0f);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_env_tex, tex.data(), n * 16,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_adjoint(mrt_ctx *c, uint32_t first_sample, uint32_t n_samples,
                           uint32_t max_depth, const float *d_image, float *g_mat,
                           float *g_hdri)
{
    if (!c || !d_image || !g_mat)
        return MRT_E_INVALID;
    if (!c->has_scene)
        return fail(c, MRT_E_STATE, "mrt_adjoint before mrt_upload_scene");
    if (max_depth == 0 || max_depth > (uint32_t)ADJ_MAXD)
        return fail(c, MRT_E_LIMIT, "adjoint depth must be in 1..%d", ADJ_MAXD);
    if (c->n_mats > (uint32_t)ADJ_MAXMAT)
        return fail(c, MRT_E_LIMIT, "adjoint supports at most %d materials",
                    ADJ_MAXMAT);
    if (c->has_photons)
        return fail(c, MRT_E_UNSUPPORTED, "no adjoint through the photon map");
    if (c->textured)
        return fail(c, MRT_E_UNSUPPORTED, "no adjoint through textured materials");
    if (!c->lights.empty() && !(c->flags & MRT_SCENE_LIGHT_EXACT))
        return fail(c, MRT_E_UNSUPPORTED,
                    "lights need MRT_SCENE_LIGHT_EXACT (truncated sums have no "
                    "derivative)");
    if (c->has_env && c->estimator != MRT_EST_MIS)
        return fail(c, MRT_E_UNSUPPORTED, "the adjoint follows the MIS estimator");
    if (g_hdri && !c->has_env)
        return fail(c, MRT_E_STATE, "HDRI gradients without an HDRI");
    const size_t ng = (size_t)c->n_mats * ADJ_NG;
    std::vector<float> zero_mat(ng, 0.0f);
    memcpy(g_mat, zero_mat.data(), ng * 4);
    const size_t ntex = (size_t)c->env_w * c->env_h * 3;
    if (g_hdri)
        memset(g_hdri, 0, ntex * 4);
    if (n_samples == 0)
        return MRT_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const uint32_t WH = c->W * c->H;
    float *d_img = nullptr, *d_gm = nullptr, *d_gt = nullptr;
    hipError_t e = hipSuccess;
    int rc = MRT_OK;
    do {
        if (dalloc(d_img, (size_t)WH * 3) != hipSuccess
            || dalloc(d_gm, ng) != hipSuccess
            || (g_hdri && dalloc(d_gt, ntex) != hipSuccess)) {
            rc = fail(c, MRT_E_NOMEM, "adjoint buffers do not fit");
            break;
        }
        if ((e = hipMemcpyAsync(d_img, d_image, (size_t)WH * 12,
                                hipMemcpyHostToDevice, c->stream)) != hipSuccess
            || (e = launch_fill_u32(c->stream, (uint32_t *)d_gm, 0u, ng)) != hipSuccess
            || (g_hdri && (e = launch_fill_u32(c->stream, (uint32_t *)d_gt, 0u, ntex))
                              != hipSuccess))
            break;
        // Record: the forward wavefront (the same kernels, queues and RNG
        // keys as mrt_render, one lane) replays the paths and writes each
        // vertex's record to the tape; backward: one thread per path walks
        // its records in reverse.  At most 16 M paths per pass (a tape of
        // (2 + 4 depth) x 16 B per path: 9 GB at depth 8).
        FrameParams fp = frame_params(c);
        fp.slot0 = 0;
        SceneDev S = scene_dev(c);
        AdjBufs A;
        A.dimg = d_img;
        A.g_mat = d_gm;
        A.g_tex = d_gt;
        A.n_mat = c->n_mats;
        A.inv_spp = 1.0f / (float)n_samples;
        uint32_t nb = std::min(c->path_budget, 1u << 24) / WH;
        nb = nb < 1 ? 1 : (nb > n_samples ? n_samples : nb);
        const uint32_t P = nb * WH;
        const uint32_t rpp = shadow_rays_per_path(c);
        const uint64_t S64 = (uint64_t)P * rpp;
        if (S64 > 0xFFFFFFF0ull) {
            rc = fail(c, MRT_E_LIMIT, "shadow queue exceeds 2^32 entries");
            break;
        }
        if ((rc = ensure_wavefront(c, P, (uint32_t)S64)) != MRT_OK)
            break;
        if (ensure_ovf(c, c->trace_blocks, 1) != hipSuccess) {
            rc = fail(c, MRT_E_NOMEM, "trace stack buffer does not fit");
            break;
        }
        const size_t tape_n = tape_row(max_depth, 0) * (size_t)P;
        if (tape_n > c->tape_cap) {
            dfree(c->d_tape);
            c->tape_cap = 0;
            if (dalloc(c->d_tape, tape_n) != hipSuccess) {
                rc = fail(c, MRT_E_NOMEM, "adjoint tape (%zu bytes) does not fit",
                          tape_n * 16);
                break;
            }
            c->tape_cap = tape_n;
        }
        ShadeBufs Bb;
        memset(&Bb, 0, sizeof(Bb));
        Bb.mats = c->d_mats;
        Bb.tape = c->d_tape;
        Timer tm{c, false, {}};
        for (uint32_t s = 0; s < n_samples && e == hipSuccess && rc == MRT_OK;
             s += nb) {
            const uint32_t n = (n_samples - s) < nb ? (n_samples - s) : nb;
            const uint32_t Pn = n * WH;
            fp.sample0 = first_sample + s;
            Bb.tape_P = Pn;
            if ((e = launch_tape_init(c->stream, c->d_tape, Pn, max_depth))
                != hipSuccess)
                break;
            rc = render_lane(c, c->stream, tm, fp, S, false, max_depth, Pn, 0,
                             rpp, c->d_ctr, stack_ovf(c, 0), c->d_tape);
            if (rc != MRT_OK)
                break;
            // one resident wave of blocks (4 per CU with the texel hash's
            // LDS), so each block's table gathers many paths before its
            // flush
            const uint32_t blocks = std::min((Pn + 255) / 256,
                                             (uint32_t)c->num_cu * 4u);
            e = launch_adjoint_back(c->stream, blocks, fp, Bb, A, Pn);
        }
        if (rc != MRT_OK)
            break;
        if (e != hipSuccess)
            break;
        if ((e = hipMemcpyAsync(g_mat, d_gm, ng * 4, hipMemcpyDeviceToHost,
                                c->stream)) != hipSuccess
            || (g_hdri && (e = hipMemcpyAsync(g_hdri, d_gt, ntex * 4,
                                              hipMemcpyDeviceToHost, c->stream))
                              != hipSuccess)
            || (e = hipStreamSynchronize(c->stream)) != hipSuccess)
            break;
    } while (0);
    (void)hipStreamSynchronize(c->stream);
    dfree(d_img);
    dfree(d_gm);
    dfree(d_gt);
    if (e != hipSuccess && rc == MRT_OK)
        rc = fail(c, MRT_E_HIP, "adjoint failed: %s", hipGetErrorString(e));
    return rc;
}

// ---- batches ------------------------------------------------------------------
namespace {
// A scene of a batch, copied from the caller's arrays.
struct BatchSrc {
    std::vector<float>        tri_pos, tri_nrm, sph, pln;
    std::vector<uint32_t>     tri_mat, sph_mat, pln_mat;
    std::vector<mrt_material> mats;
    std::vector<mrt_light>    lights;
    std::vector<mrt_quadric>  quad;
    mrt_scene                 view;
    uint32_t                  W = 0, H = 0, spp = 0;
    uint64_t                  seed = 0;
};
}  // namespace

struct mrt_batch {
    mrt_ctx                *c = nullptr;   // device, stream, queues, stats
    std::vector<BatchSrc *> src;
    bool                    committed = false;
    float4                 *d_nodes = nullptr, *d_prims = nullptr;
    float4                 *d_planes = nullptr, *d_mats = nullptr;
    float4                 *d_lights = nullptr;
    float                  *d_nrm = nullptr, *d_accum = nullptr;
    BatchScene             *d_bs = nullptr;
    BatchCtl               *d_ctl = nullptr;
    uint2                  *d_slot = nullptr;
    uint32_t                cap_slot = 0;
    std::vector<BatchScene> hbs;
    uint64_t                total = 0;
    size_t                  accum_n = 0;
    uint32_t                max_rays = 1;
    bool                    has_lights = false;
    bool                    has_quad = false;
    // the batched refill trace kernel applies (triangle / sphere scenes of
    // one sidedness, at most BATCH_LDS_SCENES of them) on its resident grid
    bool                    refill_ok = false;
    bool                    one_sided = false;
    uint32_t                refill_blocks = 0;
    uint32_t                n_planes_all = 0;
    mrt_batch_tuning        tune{};
    hipGraph_t              graph = nullptr;
    hipGraphExec_t          gexec = nullptr;
    uint32_t                g_key[7] = {0, 0, 0, 0, 0, 0, 0};
};

template <typename T>
static void vcopy(std::vector<T> &v, const T *p, size_t n)
{
    v.assign(p ? p : nullptr, p ? p + n : nullptr);
}

static void batch_free_device(mrt_batch *b)
{
    if (b->gexec)
        (void)hipGraphExecDestroy(b->gexec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    dfree(b->d_nodes);
    dfree(b->d_prims);
    dfree(b->d_planes);
    dfree(b->d_mats);
    dfree(b->d_lights);
    dfree(b->d_nrm);
    dfree(b->d_accum);
    dfree(b->d_bs);
    dfree(b->d_ctl);
    dfree(b->d_slot);
    b->cap_slot = 0;
}

extern "C" int mrt_batch_create(int device, mrt_batch **out)
{
    if (!out)
        return MRT_E_INVALID;
    *out = nullptr;
    mrt_ctx *c = nullptr;
    int rc = mrt_init(device, 1, 1, &c);
    if (rc != MRT_OK)
        return rc;
    mrt_batch *b = new (std::nothrow) mrt_batch();
    if (!b) {
        mrt_destroy(c);
        return MRT_E_NOMEM;
    }
    b->c = c;
    // the autotuner's choice on the 64-scene batched workload
    b->tune.path_budget = 1u << 27;
    b->tune.trace_blocks_per_cu = c->trace_blocks / (uint32_t)c->num_cu;
    b->tune.shade_blocks_per_cu = 32;
    b->tune.use_graph = 0;
    b->tune.trace_refill = 1;
    b->refill_blocks = std::min(c->trace_blocks,
                                (uint32_t)(c->num_cu * refill_blocks_per_cu(true)));
    *out = b;
    return MRT_OK;
}

extern "C" void mrt_batch_destroy(mrt_batch *b)
{
    if (!b)
        return;
    if (b->c) {
        (void)hipSetDevice(b->c->device);
        (void)hipStreamSynchronize(b->c->stream);
    }
    batch_free_device(b);
    for (BatchSrc *p : b->src)
        delete p;
    mrt_destroy(b->c);
    delete b;
}

extern "C" const char *mrt_batch_last_error(const mrt_batch *b)
{
    return b ? b->c->err : "null batch";
}

extern "C" int mrt_batch_add_scene(mrt_batch *b, const mrt_scene *s,
                                   uint32_t width, uint32_t height, uint32_t spp,
                                   uint64_t seed, uint32_t *id)
{
    if (!b || !s)
        return MRT_E_INVALID;
    mrt_ctx *c = b->c;
    if (b->committed)
        return fail(c, MRT_E_STATE, "batch already committed");
    if (width == 0 || height == 0 || spp == 0)
        return fail(c, MRT_E_INVALID, "empty resolution or sample count");
    if ((uint64_t)width * height * spp > 0xFFFFFFFFull)
        return fail(c, MRT_E_LIMIT, "more than 2^32 paths in one scene");
    if (b->src.size() >= 4096)
        return fail(c, MRT_E_LIMIT, "at most 4096 scenes per batch");
    int rc = validate_scene(c, s);
    if (rc != MRT_OK)
        return rc;
    if (s->material_tex && s->n_textures)
        for (uint32_t k = 0; k < s->n_materials; k++) {
            const uint32_t *sl = &s->material_tex[k].color;
            for (int j = 0; j < 8; j++)
                if (sl[j] != MRT_NO_TEX)
                    return fail(c, MRT_E_UNSUPPORTED,
                                "textured materials are not batched");
        }
    BatchSrc *p = new (std::nothrow) BatchSrc();
    if (!p)
        return fail(c, MRT_E_NOMEM, "out of host memory");
    vcopy(p->tri_pos, s->tri_pos, (size_t)s->n_triangles * 9);
    if (s->tri_nrm)
        vcopy(p->tri_nrm, s->tri_nrm, (size_t)s->n_triangles * 9);
    vcopy(p->tri_mat, s->tri_mat, s->n_triangles);
    vcopy(p->sph, s->sph, (size_t)s->n_spheres * 4);
    vcopy(p->sph_mat, s->sph_mat, s->n_spheres);
    vcopy(p->pln, s->pln, (size_t)s->n_planes * 6);
    vcopy(p->pln_mat, s->pln_mat, s->n_planes);
    vcopy(p->mats, s->materials, s->n_materials);
    vcopy(p->lights, s->lights, s->n_lights);
    vcopy(p->quad, s->quadrics, s->n_quadrics);
    p->view = *s;
    p->view.quadrics = p->quad.data();
    p->view.tri_pos = p->tri_pos.data();
    p->view.tri_nrm = s->tri_nrm ? p->tri_nrm.data() : nullptr;
    p->view.tri_mat = p->tri_mat.data();
    p->view.sph = p->sph.data();
    p->view.sph_mat = p->sph_mat.data();
    p->view.pln = p->pln.data();
    p->view.pln_mat = p->pln_mat.data();
    p->view.materials = p->mats.data();
    p->view.lights = p->lights.data();
    p->W = width;
    p->H = height;
    p->spp = spp;
    p->seed = seed;
    if (id)
        *id = (uint32_t)b->src.size();
    b->src.push_back(p);
    return MRT_OK;
}

extern "C" int mrt_batch_commit(mrt_batch *b)
{
    if (!b)
        return MRT_E_INVALID;
    mrt_ctx *c = b->c;
    if (b->committed)
        return MRT_OK;
    const size_t n = b->src.size();
    if (n == 0)
        return fail(c, MRT_E_STATE, "batch has no scenes");
    // BVHs, built in parallel
    std::vector<mrt_bvh> bv(n);
    std::vector<mrt_bvh4> bv4(n);
    std::vector<int> brc(n, MRT_OK);
    std::atomic<size_t> next{0};
    unsigned nth = std::thread::hardware_concurrency();
    nth = std::max(1u, std::min({nth, 16u, (unsigned)n}));
    for (size_t i = 0; i < n; i++) {
        memset(&bv[i], 0, sizeof(bv[i]));
        memset(&bv4[i], 0, sizeof(bv4[i]));
    }
    // with the device builder, the scenes it takes are built one after
    // the other on the stream and stay on the device; the host threads
    // build the rest
    std::vector<DeviceBvh> dv(n);
    std::vector<char> on_dev(n, 0);
    for (size_t i = 0; i < n; i++)
        memset(&dv[i], 0, sizeof(dv[i]));
    auto free_built = [&]() {
        for (size_t i = 0; i < n; i++) {
            mrt_bvh_free(&bv[i]);
            mrt_bvh4_free(&bv4[i]);
            if (on_dev[i]) {
                (void)hipFree(dv[i].nodes);
                (void)hipFree(dv[i].prims);
                (void)hipFree(dv[i].nrm);
                on_dev[i] = 0;
            }
        }
    };
    if (c->builder == MRT_BUILDER_DEVICE) {
        HIPCHK(c, hipSetDevice(c->device));
        for (size_t i = 0; i < n; i++) {
            const char *why = "";
            int r = device_bvh_build(c->stream, &b->src[i]->view, &dv[i], &why);
            if (r == MRT_OK)
                on_dev[i] = 1;
            else if (r != MRT_E_UNSUPPORTED)
                brc[i] = r;
        }
    }
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < n;) {
            if (on_dev[i] || brc[i] != MRT_OK)
                continue;
            brc[i] = mrt_bvh_build(&b->src[i]->view, MRT_BVH_MAX_LEAF, &bv[i]);
            if (brc[i] == MRT_OK)
                brc[i] = mrt_bvh4_build(&bv[i], &bv4[i]);
        }
    };
    // threads that cannot be started leave their share to the others (the
    // scenes are claimed from one counter), at worst to this thread
    std::vector<std::thread> pool;
    try {
        pool.reserve(nth);
        for (unsigned t = 1; t < nth; t++)
            pool.emplace_back(work);
    } catch (...) {
    }
    work();
    for (std::thread &t : pool)
        t.join();
    int rc = MRT_OK;
    for (size_t i = 0; i < n && rc == MRT_OK; i++) {
        if (brc[i] != MRT_OK)
            rc = fail(c, brc[i], "scene %zu: BVH build failed", i);
        else if (!on_dev[i]
                 && (bv[i].depth > MRT_BVH_MAX_DEPTH || bv4[i].depth > MRT_BVH4_MAX_DEPTH))
            rc = fail(c, MRT_E_LIMIT, "scene %zu: BVH depth %u / %u exceeds %u / %u",
                      i, bv[i].depth, bv4[i].depth, MRT_BVH_MAX_DEPTH,
                      MRT_BVH4_MAX_DEPTH);
    }
    // offsets of every scene in the shared arrays; planes, materials and
    // lights packed on the host (small)
    std::vector<uint32_t> node0s(n), prim0s(n), mat0s(n);
    std::vector<float4> planes, mats, lights;
    b->hbs.assign(n, BatchScene{});
    uint64_t path0 = 0, n_nodes_all = 0, n_prims_all = 0;
    size_t accum0 = 0;
    b->max_rays = 1;
    b->has_lights = false;
    b->has_quad = false;
    b->n_planes_all = 0;
    bool sides_agree = true;
    for (size_t i = 0; i < n && rc == MRT_OK; i++) {
        const BatchSrc &S = *b->src[i];
        const uint32_t nn4 = on_dev[i] ? dv[i].n_nodes : bv4[i].n_nodes;
        const uint32_t np = on_dev[i] ? dv[i].n_prims : bv[i].n_prims;
        b->has_quad = b->has_quad || S.view.n_quadrics > 0;
        if (n_prims_all + np >= (1ull << 28) || n_nodes_all + nn4 >= (1ull << 31)) {
            rc = fail(c, MRT_E_LIMIT, "batch exceeds 2^28 primitives");
            break;
        }
        node0s[i] = (uint32_t)n_nodes_all;
        prim0s[i] = (uint32_t)n_prims_all;
        mat0s[i] = (uint32_t)(mats.size() / 3);
        const uint32_t plane0 = (uint32_t)(planes.size() / 2);
        const uint32_t light0 = (uint32_t)(lights.size() / 3);
        n_nodes_all += nn4;
        n_prims_all += np;
        pack_scene(&S.view, mat0s[i], planes, mats, lights);
        BatchScene &T = b->hbs[i];
        memset(&T, 0, sizeof(T));
        camera_frame(T.fp, S.view.camera, S.W, S.H, S.view.ambient_ratio,
                     S.view.ambient_color);
        T.fp.ior_all = S.view.ior_all > 0.0f ? S.view.ior_all : 1.0f;
        T.fp.n_lights = S.view.n_lights;
        T.fp.one_sided = (S.view.flags & MRT_SCENE_TRIANGLE_ONE_SIDED) ? 1u : 0u;
        T.fp.flags = S.view.flags;
        T.fp.seed = S.seed;
        T.fp.estimator = MRT_EST_REFERENCE;
        T.fp.light0 = light0;
        T.fp.plane0 = plane0;
        T.light0 = light0;
        T.flags = T.fp.flags;
        T.root = nn4 ? node0s[i] : 0u;
        T.n_nodes = nn4;
        T.n_planes = S.view.n_planes;
        T.accum0 = (uint32_t)accum0;
        T.path0 = path0;
        T.paths = (uint64_t)S.W * S.H * S.spp;
        path0 += T.paths;
        accum0 += (size_t)S.W * S.H * 3;
        uint32_t rays = 0;
        for (const mrt_light &l : S.lights) {
            bool soft = l.type == MRT_LIGHT_POINT && l.radius != 0.0f
                        && l.shadow_samples > 1;
            rays += soft ? (uint32_t)l.shadow_samples : 1u;
        }
        b->max_rays = std::max(b->max_rays, rays);
        b->has_lights = b->has_lights || !S.lights.empty();
        b->n_planes_all += S.view.n_planes;
        if (i == 0)
            b->one_sided = T.fp.one_sided != 0;
        sides_agree = sides_agree && (T.fp.one_sided != 0) == b->one_sided;
    }
    if (rc == MRT_OK && accum0 >= (1ull << 32))
        rc = fail(c, MRT_E_LIMIT, "batch images exceed 2^32 floats");
    if (rc != MRT_OK) {
        free_built();
        return rc;
    }
    b->refill_ok = !b->has_quad && sides_agree && n <= BATCH_LDS_SCENES
                   && b->refill_blocks > 0;
    HIPCHK(c, hipSetDevice(c->device));
    batch_free_device(b);
    BatchCtl ctl{0, path0, (uint32_t)n, 0};
    if (dalloc(b->d_nodes, n_nodes_all * 8) != hipSuccess
        || dalloc(b->d_prims, n_prims_all * 3) != hipSuccess
        || dalloc(b->d_nrm, n_prims_all * 9) != hipSuccess
        || dalloc(b->d_planes, planes.size()) != hipSuccess
        || dalloc(b->d_mats, mats.size()) != hipSuccess
        || dalloc(b->d_lights, lights.size()) != hipSuccess
        || dalloc(b->d_accum, accum0) != hipSuccess
        || dalloc(b->d_bs, n) != hipSuccess || dalloc(b->d_ctl, 1) != hipSuccess) {
        batch_free_device(b);
        free_built();
        return fail(c, MRT_E_NOMEM, "device allocation for the batch failed");
    }
    // every scene's arrays into place, then moved by its offsets
    for (size_t i = 0; i < n && rc == MRT_OK; i++) {
        const uint32_t nn4 = on_dev[i] ? dv[i].n_nodes : bv4[i].n_nodes;
        const uint32_t np = on_dev[i] ? dv[i].n_prims : bv[i].n_prims;
        float4 *dn = b->d_nodes + (size_t)node0s[i] * 8;
        float4 *dp = b->d_prims + (size_t)prim0s[i] * 3;
        float *dr = b->d_nrm + (size_t)prim0s[i] * 9;
        hipMemcpyKind k = on_dev[i] ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const void *sn = on_dev[i] ? (const void *)dv[i].nodes : (const void *)bv4[i].nodes;
        const void *sp = on_dev[i] ? (const void *)dv[i].prims : (const void *)bv[i].prims;
        const void *sr = on_dev[i] ? (const void *)dv[i].nrm : (const void *)bv[i].shade_nrm;
        if ((nn4 && hipMemcpyAsync(dn, sn, (size_t)nn4 * 128, k, c->stream) != hipSuccess)
            || (np && hipMemcpyAsync(dp, sp, (size_t)np * 48, k, c->stream) != hipSuccess)
            || (np && hipMemcpyAsync(dr, sr, (size_t)np * 36, k, c->stream) != hipSuccess)
            || launch_rebase(c->stream, dn, nn4, dp, dr, np, node0s[i], prim0s[i],
                             mat0s[i]) != hipSuccess)
            rc = fail(c, MRT_E_HIP, "scene %zu: copy into the batch failed", i);
    }
    if (rc == MRT_OK)
        rc = upload(c, b->d_planes, planes.data(), planes.size() * 16);
    if (rc == MRT_OK)
        rc = upload(c, b->d_mats, mats.data(), mats.size() * 16);
    if (rc == MRT_OK)
        rc = upload(c, b->d_lights, lights.data(), lights.size() * 16);
    if (rc == MRT_OK)
        rc = upload(c, b->d_bs, b->hbs.data(), n * sizeof(BatchScene));
    if (rc == MRT_OK)
        rc = upload(c, b->d_ctl, &ctl, sizeof(ctl));
    if (rc == MRT_OK
        && (hipMemsetAsync(b->d_accum, 0, accum0 * 4, c->stream) != hipSuccess
            || hipStreamSynchronize(c->stream) != hipSuccess))
        rc = fail(c, MRT_E_HIP, "batch upload failed");
    (void)hipStreamSynchronize(c->stream);
    free_built();
    if (rc != MRT_OK) {
        batch_free_device(b);
        return rc;
    }
    b->total = path0;
    b->accum_n = accum0;
    b->committed = true;
    for (BatchSrc *p : b->src)
        delete p;
    b->src.clear();
    return MRT_OK;
}

// Queues for P paths per pass: one pass split over L lanes (a captured
// pass), or L passes in flight on L lanes (pipelined).
static int batch_buffers(mrt_batch *b, uint32_t P, uint32_t L, bool pipelined)
{
    mrt_ctx *c = b->c;
    const uint64_t Q64 = pipelined ? (uint64_t)P * L : P;
    const uint64_t S64 = Q64 * b->max_rays;
    if (S64 > 0xFFFFFFF0ull || Q64 > 0xFFFFFFF0ull)
        return fail(c, MRT_E_LIMIT, "shadow queue exceeds 2^32 entries");
    int rc = ensure_wavefront(c, (uint32_t)Q64, (uint32_t)S64);
    if (rc != MRT_OK)
        return rc;
    const uint32_t blocks = std::max((uint32_t)c->num_cu * b->tune.trace_blocks_per_cu,
                                     b->refill_blocks);
    if (ensure_ovf(c, blocks, L) != hipSuccess)
        return fail(c, MRT_E_NOMEM, "trace stack buffer does not fit");
    if (Q64 > b->cap_slot) {
        if (dalloc(b->d_slot, Q64) != hipSuccess) {
            b->cap_slot = 0;
            return fail(c, MRT_E_NOMEM, "slot table (%llu) does not fit",
                        (unsigned long long)Q64);
        }
        b->cap_slot = (uint32_t)Q64;
    }
    return MRT_OK;
}

// The launch shapes and shared buffers of a batch's passes.
struct BatchShape {
    SceneDev  S;
    ShadeBufs B;
    uint32_t  tb, gb, rb;    // trace grid, shade grid, refill grid (0: k_trace)
};

static BatchShape batch_shape(const mrt_batch *b, uint32_t P)
{
    const mrt_ctx *c = b->c;
    BatchShape sh;
    memset(&sh, 0, sizeof(sh));
    SceneDev &S = sh.S;
    S.nodes = b->d_nodes;
    S.prims = b->d_prims;
    S.aux = b->d_nrm;
    S.planes = b->d_planes;
    S.quad = b->has_quad ? 1u : 0u;
    // for the batched refill kernel: the sidedness every scene shares, and
    // whether any scene has planes (their total)
    S.one_sided = b->one_sided ? 1u : 0u;
    S.n_planes = b->n_planes_all;
    ShadeBufs &B = sh.B;
    B.acc = c->acc;
    B.planes = b->d_planes;
    B.prims = b->d_prims;
    B.mats = b->d_mats;
    B.lights = b->d_lights;
    B.nrm = b->d_nrm;
    B.s_o = c->s_o;
    B.s_d = c->s_d;
    B.s_c = c->s_c;
    B.vis = c->vis;
    B.p_a = c->p_a;
    B.p_b = c->p_b;
    B.p_c = c->p_c;
    B.p_x = c->p_x;
    B.p_k = c->p_k;
    B.cap_paths = c->cap_paths;
    B.cap_shadow = c->cap_shadow;
    B.bs = b->d_bs;
    B.slot_info = b->d_slot;
    sh.tb = (uint32_t)c->num_cu * b->tune.trace_blocks_per_cu;
    sh.gb = std::min((P + SHADE_BLOCK - 1) / SHADE_BLOCK,
                     (uint32_t)c->num_cu * b->tune.shade_blocks_per_cu);
    sh.rb = (b->tune.trace_refill && b->refill_ok) ? b->refill_blocks : 0u;
    return sh;
}

#define BSTEP(cls, expr)                                                     \
    do {                                                                     \
        c->launches[cls]++;                                                  \
        if (tm)                                                              \
            tm->begin(cls);                                                  \
        e = (expr);                                                          \
        if (tm)                                                              \
            tm->end();                                                       \
        if (e != hipSuccess)                                                 \
            return e;                                                        \
    } while (0)

// One render lane's part of a pass: paths [o0, o0 + Pl) of the pass of P
// paths starting at global path `base` (BASE_FROM_CTL: the device
// cursor), on stream st with lane `lane`'s counters and stack columns and
// the queue entries from qoff (shadow entries from qoff * max_rays).  When
// ev_after is given it is recorded once bounce LANE_STAGGER_BOUNCE is traced.
static hipError_t batch_lane(mrt_batch *b, const BatchShape &sh, uint32_t depth,
                             hipStream_t st, uint32_t lane,
                             unsigned long long base, uint32_t P, uint32_t o0,
                             uint32_t Pl, uint32_t qoff, bool count, Timer *tm,
                             hipEvent_t ev_after, uint32_t top_base = 0,
                             uint32_t ntop = 0, uint32_t s_lo = 0,
                             uint32_t s_hi = ~0u)
{
    mrt_ctx *c = b->c;
    hipError_t e;
    const uint32_t R = b->max_rays;
    const size_t so = (size_t)qoff * R;
    uint32_t *ctr = c->d_ctr + (size_t)lane * CTR_COUNT;
    const StackOvf ovf = stack_ovf(c, lane);
    if (tm)
        tm->st = st;
    ShadeBufs Bl = sh.B;
    Bl.acc = sh.B.acc + qoff;
    Bl.s_o = sh.B.s_o + so;
    Bl.s_d = sh.B.s_d + so;
    Bl.s_c = sh.B.s_c + so;
    Bl.vis = sh.B.vis + so;
    Bl.p_a = sh.B.p_a + qoff;
    Bl.p_b = sh.B.p_b + qoff;
    Bl.p_c = sh.B.p_c + qoff;
    Bl.p_x = sh.B.p_x + qoff;
    Bl.p_k = sh.B.p_k + qoff;
    Bl.slot_info = sh.B.slot_info + qoff;
    Bl.cap_paths = Pl;
    Bl.cap_shadow = Pl * R;
    const uint32_t gbl = std::min((Pl + SHADE_BLOCK - 1) / SHADE_BLOCK, sh.gb);
    BSTEP(K_RAYGEN, launch_braygen(st, Bl, b->d_ctl, base, s_lo, s_hi,
                                   c->q_o[0] + qoff, c->q_d[0] + qoff,
                                   c->q_t[0] + qoff, o0, Pl, P, ctr));
    // the order of the plain render: the closest-hit rays of bounce k with
    // the shadow rays of bounce k - 1, the light sums of k - 1, the shade of k
    FrameParams fl;
    memset(&fl, 0, sizeof(fl));
    ShadeBufs Bp = Bl;
    FrameParams fpp = fl;
    const uint32_t n_scenes = (uint32_t)b->hbs.size();
    for (uint32_t k = 0; k <= depth; k++) {
        const bool last = k == depth;
        const int cur = k & 1, nxt = cur ^ 1;
        TraceWork w;
        w.q_o = c->q_o[cur] + qoff;
        w.q_d = c->q_d[cur] + qoff;
        w.q_t = c->q_t[cur] + qoff;
        w.hit = c->hit + qoff;
        w.s_o = Bl.s_o;
        w.s_d = Bl.s_d;
        w.vis = Bl.vis;
        w.cap_shadow = Bl.cap_shadow;
        w.closest = !last;
        w.shadow_ctr = (k > 0 && b->has_lights) ? ctr_shadow(k - 1) : -1;
        if (w.closest || w.shadow_ctr >= 0)
            BSTEP(K_TRACE, launch_btrace(st, sh.tb, count, sh.S, ovf, Bl, w, ctr,
                                         c->d_stats, sh.rb, n_scenes, top_base,
                                         ntop));
        if (ev_after && (k == LANE_STAGGER_BOUNCE || last)) {
            if ((e = hipEventRecord(ev_after, st)) != hipSuccess)
                return e;
            ev_after = nullptr;
        }
        if (k > 0 && b->has_lights)
            BSTEP(K_RESOLVE, launch_bresolve(st, gbl, fpp, Bp, ctr));
        if (last) {
            BSTEP(K_ADVANCE, launch_advance(st, ctr, c->d_stats, Pl, k));
            break;
        }
        fl.bounce = k;
        fl.last_bounce = (k + 1 == depth) ? 1u : 0u;
        ShadeBufs Bk = Bl;
        Bk.q_o = c->q_o[cur] + qoff;
        Bk.q_d = c->q_d[cur] + qoff;
        Bk.q_t = c->q_t[cur] + qoff;
        Bk.hit = c->hit + qoff;
        Bk.qn_o = c->q_o[nxt] + qoff;
        Bk.qn_d = c->q_d[nxt] + qoff;
        Bk.qn_t = c->q_t[nxt] + qoff;
        BSTEP(K_SHADE, launch_bshade(st, gbl, fl, Bk, ctr));
        BSTEP(K_ADVANCE, launch_advance(st, ctr, c->d_stats, Pl, k));
        Bp = Bk;
        fpp = fl;
    }
    if (tm)
        tm->st = nullptr;
    return hipSuccess;
}

// One pass of P paths at the device cursor, split into contiguous parts on
// the render lanes (joined before the accumulation; inside a captured graph
// the fork and join events carry the dependencies), then the cursor moves
// on.  This is the pass a graph captures and replays.
static hipError_t batch_pass(mrt_batch *b, uint32_t depth, uint32_t P,
                             bool count, Timer *tm)
{
    mrt_ctx *c = b->c;
    const BatchShape sh = batch_shape(b, P);
    hipError_t e;
    const uint32_t L = std::max(1u, std::min(c->lanes, P / LANE_MIN_PATHS));
    if (L > 1) {
        if ((e = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess)
            return e;
        for (uint32_t l = 1; l < L; l++)
            if ((e = hipStreamWaitEvent(c->lane_stream[l], c->ev_fork, 0))
                != hipSuccess)
                return e;
    }
    for (uint32_t l = 0; l < L; l++) {
        const uint32_t o0 = (uint32_t)((uint64_t)P * l / L);
        const uint32_t Pl = (uint32_t)((uint64_t)P * (l + 1) / L) - o0;
        hipStream_t st = l ? c->lane_stream[l] : c->stream;
        if ((e = batch_lane(b, sh, depth, st, l, BASE_FROM_CTL, P, o0, Pl, o0,
                            count, tm, nullptr)) != hipSuccess)
            return e;
        if (l) {
            if ((e = hipEventRecord(c->ev_join[l], st)) != hipSuccess)
                return e;
            if ((e = hipStreamWaitEvent(c->stream, c->ev_join[l], 0)) != hipSuccess)
                return e;
        }
    }
    BSTEP(K_ACCUM, launch_baccumulate(c->stream, sh.B, b->d_ctl, BASE_FROM_CTL,
                                      b->d_accum, P, c->d_ctr, c->d_stats));
    BSTEP(K_ACCUM, launch_bnext(c->stream, b->d_ctl, P));
    return hipSuccess;
}

// The scenes [s_lo, s_hi] that global paths [base, base + P) fall in, and
// the one holding most of them: the batched refill kernel keeps the top of
// its tree in LDS for that pass (every scene's tree starts at its root with
// its top levels first).
static void pass_top(const mrt_batch *b, uint64_t base, uint32_t P,
                     uint32_t &top_base, uint32_t &ntop, uint32_t &s_lo,
                     uint32_t &s_hi)
{
    const std::vector<BatchScene> &h = b->hbs;
    size_t lo = 0, hi = h.size();
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (h[mid].path0 + h[mid].paths <= base)
            lo = mid + 1;
        else
            hi = mid;
    }
    const uint64_t end = base + P;
    uint64_t best = 0;
    top_base = 0;
    ntop = 0;
    s_lo = (uint32_t)std::min(lo, h.size() - 1);
    s_hi = s_lo;
    for (size_t i = lo; i < h.size() && h[i].path0 < end; i++) {
        s_hi = (uint32_t)i;
        const uint64_t a = std::max<uint64_t>(base, h[i].path0);
        const uint64_t z = std::min<uint64_t>(end, h[i].path0 + h[i].paths);
        if (z > a && z - a > best && h[i].n_nodes) {
            best = z - a;
            top_base = h[i].root;
            ntop = std::min<uint32_t>((uint32_t)TOP_NODES, h[i].n_nodes);
        }
    }
}

// Passes of P paths from global path 0 to `total`, launched directly and
// pipelined over the render lanes as mrt_render's passes are: pass k runs
// whole on lane k % L with that lane's queues, starts once pass k - 1 has
// traced bounce LANE_STAGGER_BOUNCE, and adds its sums after pass k - 1's,
// so the sums are those of the captured pass replayed pass after pass.
static hipError_t batch_pipelined(mrt_batch *b, uint32_t depth, uint32_t P,
                                  uint32_t L, uint64_t total, bool count, Timer *tm)
{
    mrt_ctx *c = b->c;
    const BatchShape sh = batch_shape(b, P);
    hipError_t e;
    if (L > 1) {
        if ((e = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess)
            return e;
        for (uint32_t l = 1; l < L; l++)
            if ((e = hipStreamWaitEvent(c->lane_stream[l], c->ev_fork, 0))
                != hipSuccess)
                return e;
    }
    uint32_t l = 0;
    uint64_t k = 0;
    for (uint64_t base = 0; base < total; base += P, k++) {
        const uint32_t prev = l;
        l = (uint32_t)(k % L);
        hipStream_t st = l ? c->lane_stream[l] : c->stream;
        if (k && L > 1 && (e = hipStreamWaitEvent(st, c->ev_stag[prev], 0)) != hipSuccess)
            return e;
        const uint32_t qoff = P * l;
        uint32_t top_base = 0, ntop = 0, s_lo = 0, s_hi = 0;
        pass_top(b, base, P, top_base, ntop, s_lo, s_hi);
        if ((e = batch_lane(b, sh, depth, st, l, base, P, 0, P, qoff, count, tm,
                            L > 1 ? c->ev_stag[l] : nullptr, top_base, ntop,
                            s_lo, s_hi))
            != hipSuccess)
            return e;
        if (k && L > 1 && (e = hipStreamWaitEvent(st, c->ev_acc[prev], 0)) != hipSuccess)
            return e;
        ShadeBufs Bl = sh.B;
        Bl.acc = sh.B.acc + qoff;
        Bl.slot_info = sh.B.slot_info + qoff;
        if (tm)
            tm->st = st;
        BSTEP(K_ACCUM, launch_baccumulate(st, Bl, b->d_ctl, base, b->d_accum, P,
                                          c->d_ctr + (size_t)l * CTR_COUNT,
                                          c->d_stats));
        if (tm)
            tm->st = nullptr;
        if (L > 1 && (e = hipEventRecord(c->ev_acc[l], st)) != hipSuccess)
            return e;
    }
    // the context's stream after the last pass's sums
    if (l && (e = hipStreamWaitEvent(c->stream, c->ev_acc[l], 0)) != hipSuccess)
        return e;
    return hipSuccess;
}
#undef BSTEP

// Renders global paths [0, total) of the batch in passes of P.
static int batch_run(mrt_batch *b, uint32_t depth, uint64_t total)
{
    mrt_ctx *c = b->c;
    uint32_t P = (uint32_t)std::min<uint64_t>(b->tune.path_budget, total);
    if (P == 0)
        return MRT_OK;
    const uint64_t passes = (total + P - 1) / P;
    const bool count = (c->stats_on & MRT_STATS_COUNT) != 0;
    const bool timing = (c->stats_on & MRT_STATS_TIME) != 0;
    const bool graph = b->tune.use_graph && !count && !timing;
    // lanes: a captured pass splits into up to MRT_OPT_LANES parts of at
    // least LANE_MIN_PATHS; direct launches pipeline up to that many passes
    const uint32_t L = graph
        ? std::max(1u, std::min(c->lanes, P / LANE_MIN_PATHS))
        : (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(c->lanes, passes));
    int rc = batch_buffers(b, P, L, !graph);
    if (rc != MRT_OK)
        return rc;
    BatchCtl ctl{0, total, (uint32_t)b->hbs.size(), 0};
    HIPCHK(c, hipMemcpyAsync(b->d_ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice,
                             c->stream));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing) {
        (void)hipEventCreate(&t0);
        (void)hipEventCreate(&t1);
        (void)hipEventRecord(t0, c->stream);
    }
    if (graph) {
        const uint32_t key[7] = {depth, P, b->tune.trace_blocks_per_cu,
                                 b->tune.shade_blocks_per_cu, c->cap_shadow,
                                 c->lanes, b->tune.trace_refill};
        if (!b->gexec || memcmp(key, b->g_key, sizeof(key)) != 0) {
            if (b->gexec)
                (void)hipGraphExecDestroy(b->gexec);
            if (b->graph)
                (void)hipGraphDestroy(b->graph);
            b->gexec = nullptr;
            b->graph = nullptr;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = batch_pass(b, depth, P, false, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g)
                    (void)hipGraphDestroy(g);
                return fail(c, MRT_E_HIP, "pass capture failed: %s",
                            hipGetErrorString(e != hipSuccess ? e : e2));
            }
            b->graph = g;
            HIPCHK(c, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
            memcpy(b->g_key, key, sizeof(key));
        }
        for (uint64_t k = 0; k < passes; k++)
            HIPCHK(c, hipGraphLaunch(b->gexec, c->stream));
    } else {
        Timer tm{c, timing, {}};
        hipError_t e = batch_pipelined(b, depth, P, L, total, count, &tm);
        if (e != hipSuccess) {
            tm.collect();
            return fail(c, MRT_E_HIP, "kernel launch failed: %s",
                        hipGetErrorString(e));
        }
        tm.collect();
    }
    c->paths += total;
    if (timing) {
        (void)hipEventRecord(t1, c->stream);
        (void)hipStreamSynchronize(c->stream);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t0, t1) == hipSuccess)
            c->ms[5] += ms;
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    }
    return MRT_OK;
}

extern "C" int mrt_batch_render(mrt_batch *b, uint32_t max_depth)
{
    if (!b)
        return MRT_E_INVALID;
    mrt_ctx *c = b->c;
    if (!b->committed) {
        int rc = mrt_batch_commit(b);
        if (rc != MRT_OK)
            return rc;
    }
    if (max_depth > 4095)
        return fail(c, MRT_E_LIMIT, "max_depth above 4095");
    if (max_depth == 0)
        return MRT_OK;
    HIPCHK(c, hipSetDevice(c->device));
    return batch_run(b, max_depth, b->total);
}

extern "C" int mrt_batch_readback(mrt_batch *b, uint32_t id, float *out)
{
    if (!b || !out)
        return MRT_E_INVALID;
    mrt_ctx *c = b->c;
    if (!b->committed || id >= b->hbs.size())
        return fail(c, MRT_E_INVALID, "no committed scene %u", id);
    const BatchScene &T = b->hbs[id];
    HIPCHK(c, hipMemcpyAsync(out, b->d_accum + T.accum0,
                             (size_t)T.fp.WH * 3 * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

extern "C" int mrt_batch_clear(mrt_batch *b)
{
    if (!b)
        return MRT_E_INVALID;
    mrt_ctx *c = b->c;
    if (b->committed)
        HIPCHK(c, hipMemsetAsync(b->d_accum, 0, b->accum_n * 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRT_OK;
}

static int check_tuning(mrt_batch *b, const mrt_batch_tuning *t)
{
    if (t->path_budget < 1024 || t->path_budget > (1u << 27))
        return fail(b->c, MRT_E_LIMIT, "path budget must be in 2^10..2^27");
    if (t->trace_blocks_per_cu < 1 || t->trace_blocks_per_cu > 32
        || t->shade_blocks_per_cu < 1 || t->shade_blocks_per_cu > 32)
        return fail(b->c, MRT_E_LIMIT, "blocks per CU must be in 1..32");
    if (t->use_graph > 1 || t->trace_refill > 1)
        return fail(b->c, MRT_E_INVALID, "use_graph and trace_refill are 0 or 1");
    return MRT_OK;
}

extern "C" int mrt_batch_set_tuning(mrt_batch *b, const mrt_batch_tuning *t)
{
    if (!b || !t)
        return MRT_E_INVALID;
    int rc = check_tuning(b, t);
    if (rc != MRT_OK)
        return rc;
    b->tune = *t;
    return MRT_OK;
}

extern "C" int mrt_batch_get_tuning(const mrt_batch *b, mrt_batch_tuning *t)
{
    if (!b || !t)
        return MRT_E_INVALID;
    *t = b->tune;
    return MRT_OK;
}

// Times the first `trial` paths of the batch under each candidate shape and
// keeps the fastest; the image sums are cleared afterwards.
extern "C" int mrt_batch_autotune(mrt_batch *b, uint32_t max_depth,
                                  mrt_batch_tuning *chosen)
{
    if (!b)
        return MRT_E_INVALID;
    mrt_ctx *c = b->c;
    if (!b->committed) {
        int rc = mrt_batch_commit(b);
        if (rc != MRT_OK)
            return rc;
    }
    if (max_depth == 0 || max_depth > 4095)
        return fail(c, MRT_E_LIMIT, "max_depth must be in 1..4095");
    HIPCHK(c, hipSetDevice(c->device));
    const int saved_stats = c->stats_on;
    c->stats_on = 0;
    const uint32_t occ = std::max(1u, c->trace_blocks / (uint32_t)c->num_cu);
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    mrt_batch_tuning best = b->tune;
    double best_ms = 1e300;
    int rc = MRT_OK;
    // times the first `trial` paths under t (after a warm-up render that
    // also captures the graph) and keeps the fastest
    auto try_one = [&](mrt_batch_tuning t, uint64_t trial) {
        b->tune = t;
        rc = batch_run(b, max_depth, trial);
        if (rc != MRT_OK)
            return;
        (void)hipEventRecord(e0, c->stream);
        rc = batch_run(b, max_depth, trial);
        (void)hipEventRecord(e1, c->stream);
        if (rc != MRT_OK)
            return;
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        t.ms_per_mpath = ms / ((double)trial / 1e6);
        if (t.ms_per_mpath < best_ms) {
            best_ms = t.ms_per_mpath;
            best = t;
        }
    };
    // stage 1, the launch shape at 16 M-path passes over the first 64 M
    // paths: trace kernel (the refill kernel on its resident grid, or
    // k_trace on one of two grids), shade grid, and a captured pass
    // replayed or passes pipelined on the lanes
    const uint64_t trial1 = std::min<uint64_t>(b->total, 1ull << 26);
    const uint32_t pb1 = (uint32_t)std::min<uint64_t>(1u << 24, trial1);
    for (uint32_t tk = 0; tk < 3 && rc == MRT_OK; tk++)
        for (uint32_t sbc : {8u, 32u})
            for (uint32_t graph : {1u, 0u}) {
                const uint32_t refill = tk == 0 ? 1u : 0u;
                if ((refill && !b->refill_ok) || rc != MRT_OK)
                    continue;
                const uint32_t tbc = tk == 2 ? std::max(1u, occ / 2) : occ;
                try_one(mrt_batch_tuning{std::max(pb1, 1024u), tbc, sbc, graph,
                                         refill, 0.0},
                        trial1);
            }
    // stage 2, the path budget of that shape over the first 256 M paths
    // (longer passes: each bounce launch is long against its slowest rays)
    const uint64_t trial2 = std::min<uint64_t>(b->total, 1ull << 28);
    const mrt_batch_tuning shape = best;
    best_ms = 1e300;
    for (uint32_t pb : {1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26, 1u << 27}) {
        if (rc != MRT_OK || (pb > trial2 && pb != (1u << 22)))
            continue;
        mrt_batch_tuning t = shape;
        t.path_budget = pb;
        try_one(t, trial2);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->stats_on = saved_stats;
    b->tune = best;
    c->paths = 0;
    if (rc != MRT_OK)
        return rc;
    rc = mrt_batch_clear(b);
    if (rc == MRT_OK)
        rc = mrt_reset_stats(c);
    if (chosen)
        *chosen = best;
    return rc;
}

extern "C" int mrt_batch_set_option(mrt_batch *b, int option, double value)
{
    if (!b)
        return MRT_E_INVALID;
    switch (option) {
    case MRT_OPT_STATS:
    case MRT_OPT_LANES:
    case MRT_OPT_BUILDER:
        if (option == MRT_OPT_BUILDER && b->committed)
            return fail(b->c, MRT_E_STATE, "batch already committed");
        return mrt_set_option(b->c, option, value);
    case MRT_OPT_PATH_BUDGET: {
        mrt_batch_tuning t = b->tune;
        if (!(value >= 1024.0) || value > (double)(1u << 27))
            return fail(b->c, MRT_E_LIMIT, "path budget must be in 2^10..2^27");
        t.path_budget = (uint32_t)value;
        return mrt_batch_set_tuning(b, &t);
    }
    default:
        return fail(b->c, MRT_E_UNSUPPORTED, "option %d does not apply to batches",
                    option);
    }
}

extern "C" int mrt_batch_get_stats(mrt_batch *b, mrt_stats *out)
{
    if (!b)
        return MRT_E_INVALID;
    return mrt_get_stats(b->c, out);
}
